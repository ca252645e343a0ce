/*
 * engine.cpp -- host side scene manager: rt_SCENE walker and update stage.
 *
 * The arithmetic here is float, evaluated in source order (the library is
 * built with -ffp-contract=off), because every value produced is consumed
 * bit-for-bit by the tracer: a different rounding in a clip value or an
 * inverse matrix moves hit points.  QuadRay's references for each step are
 * cited inline; lists are built with complete membership (QuadRay with its
 * CPU list-pruning options off), see RT_OPTS_PRUNING in qr_format.h.
 */
#include "engine.h"
#include "image.h"

#include <algorithm>
#include <functional>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace qr {

const rt_mat4 iden4 =
{
    {1.0f, 0.0f, 0.0f, 0.0f},
    {0.0f, 1.0f, 0.0f, 0.0f},
    {0.0f, 0.0f, 1.0f, 0.0f},
    {0.0f, 0.0f, 0.0f, 1.0f},
};

static const float QR_INF = FLT_MAX;

static inline float fmaxq(float a, float b) { return a > b ? a : b; }
static inline float fminq(float a, float b) { return a < b ? a : b; }
static inline int   isign(float a) { return a < 0.0f ? -1 : a > 0.0f ? +1 : 0; }
/* square root that is 0 for non-positive input (QuadRay RT_SQRT32) */
static inline float sqrtq(float a) { return a <= 0.0f ? 0.0f : sqrtf(a); }

/* sin/cos of an angle in degrees, exact at multiples of 90
 * (QuadRay RT_SINA32/RT_COSA32, core/config/rtbase.h:858-875) */
static float sina(float a)
{
    if (a == -270.0f) return +1.0f;
    if (a == -180.0f) return  0.0f;
    if (a ==  -90.0f) return -1.0f;
    if (a ==    0.0f) return  0.0f;
    if (a ==  +90.0f) return +1.0f;
    if (a == +180.0f) return  0.0f;
    if (a == +270.0f) return -1.0f;
    return sinf((float)(a * 3.14159265358979323846 / 180.0));
}

static float cosa(float a)
{
    if (a == -270.0f) return  0.0f;
    if (a == -180.0f) return -1.0f;
    if (a ==  -90.0f) return  0.0f;
    if (a ==    0.0f) return +1.0f;
    if (a ==  +90.0f) return  0.0f;
    if (a == +180.0f) return -1.0f;
    if (a == +270.0f) return  0.0f;
    return cosf((float)(a * 3.14159265358979323846 / 180.0));
}

/* ------------------------------------------------------------------------ */
/* matrices (QuadRay core/engine/rtgeom.cpp:60-200)                          */
/* ------------------------------------------------------------------------ */

static void mat_mul_vec(float *vp, const float (*m1)[4], const float *v1)
{
    for (int i = 0; i < 4; i++)
    {
        vp[i] = m1[0][i] * v1[0] + m1[1][i] * v1[1] +
                m1[2][i] * v1[2] + m1[3][i] * v1[3];
    }
}

void mat_mul_mat(rt_mat4 mp, const float (*m1)[4], const float (*m2)[4])
{
    for (int i = 0; i < 4; i++)
    {
        mat_mul_vec(mp[i], m1, m2[i]);
    }
}

void mat_from_transform(rt_mat4 mp, const rt_TRANSFORM3D *t, bool with_scale)
{
    rt_mat4 mt;
    float sx = with_scale ? t->scl[RT_X] : 1.0f;
    float sy = with_scale ? t->scl[RT_Y] : 1.0f;
    float sz = with_scale ? t->scl[RT_Z] : 1.0f;
    rt_mat4 sc = {{sx, 0, 0, 0}, {0, sy, 0, 0}, {0, 0, sz, 0}, {0, 0, 0, 1}};

    float s = sina(t->rot[RT_X]), c = cosa(t->rot[RT_X]);
    rt_mat4 rx = {{1, 0, 0, 0}, {0, +c, +s, 0}, {0, -s, +c, 0}, {0, 0, 0, 1}};
    s = sina(t->rot[RT_Y]); c = cosa(t->rot[RT_Y]);
    rt_mat4 ry = {{+c, 0, -s, 0}, {0, 1, 0, 0}, {+s, 0, +c, 0}, {0, 0, 0, 1}};
    s = sina(t->rot[RT_Z]); c = cosa(t->rot[RT_Z]);
    rt_mat4 rz = {{+c, +s, 0, 0}, {-s, +c, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    rt_mat4 ps = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0},
                  {t->pos[RT_X], t->pos[RT_Y], t->pos[RT_Z], 1}};

    mat_mul_mat(mt, rx, sc);
    mat_mul_mat(mp, ry, mt);
    mat_mul_mat(mt, rz, mp);
    mat_mul_mat(mp, ps, mt);
}

void mat_inverse(rt_mat4 mp, const float (*m1)[4])
{
    memset(mp, 0, sizeof(rt_mat4));

    float a = m1[1][1] * m1[2][2] - m1[2][1] * m1[1][2];
    float b = m1[2][1] * m1[0][2] - m1[0][1] * m1[2][2];
    float c = m1[0][1] * m1[1][2] - m1[1][1] * m1[0][2];
    float d = m1[2][0] * m1[1][2] - m1[1][0] * m1[2][2];
    float e = m1[0][0] * m1[2][2] - m1[2][0] * m1[0][2];
    float f = m1[0][2] * m1[1][0] - m1[0][0] * m1[1][2];
    float g = m1[1][0] * m1[2][1] - m1[2][0] * m1[1][1];
    float h = m1[2][0] * m1[0][1] - m1[0][0] * m1[2][1];
    float l = m1[0][0] * m1[1][1] - m1[1][0] * m1[0][1];
    float q = 1.0f / (m1[0][0] * a + m1[1][0] * b + m1[2][0] * c);

    mp[0][0] = a * q; mp[0][1] = b * q; mp[0][2] = c * q;
    mp[1][0] = d * q; mp[1][1] = e * q; mp[1][2] = f * q;
    mp[2][0] = g * q; mp[2][1] = h * q; mp[2][2] = l * q;
}

/* ------------------------------------------------------------------------ */
/* construction                                                             */
/* ------------------------------------------------------------------------ */

Engine::Engine(rt_SCENE *scn, int x_res, int y_res, int x_row)
    : scn_(scn)
{
    if (scn == nullptr)
    {
        throw Error("null-pointer in scene");
    }
    if (scn->lock != RT_NULL)
    {
        throw Error("scene data is locked by another instance");
    }
    if (x_res <= 0 || y_res <= 0 || std::abs(x_row) < x_res)
    {
        throw Error("frambuffer's dimensions are not valid");
    }

    x_res_ = x_res;
    y_res_ = y_res;
    x_row_ = x_row;

    /* QuadRay rt_Scene::rt_Scene, core/engine/engine.cpp:2900-2904 */
    factor_ = 1.0f / (float)x_res;
    aspect_ = (float)y_res * factor_;
    opts_ = (int)(RT_OPTS_FULL & ~scn->opts);

    memset(&rootobj_, 0, sizeof(rootobj_));
    rootobj_.trm.scl[RT_I] = 1.0f;
    rootobj_.trm.scl[RT_J] = 1.0f;
    rootobj_.trm.scl[RT_K] = 1.0f;
    rootobj_.obj = scn->root;

    if (!RT_IS_ARRAY(&scn->root))
    {
        throw Error("scene's root is not an array");
    }

    try
    {
        root_ = make_array(nullptr, &rootobj_);

        if (cameras_.empty())
        {
            throw Error("scene doesn't contain camera");
        }
    }
    catch (...)
    {
        for (Object *o : objects_)
        {
            if (o->obj != &rootobj_) o->obj->trm = o->otm;
            delete o;
        }
        objects_.clear();
        throw;
    }

    /* the camera list head is the camera constructed last */
    cam_ = cameras_.back();
    cam_idx_ = 0;

    scn->lock = this;
}

Engine::~Engine()
{
    for (Object *o : objects_)
    {
        if (o->obj != &rootobj_)
        {
            o->obj->trm = o->otm;
        }
        delete o;
    }
    if (scn_ != nullptr && scn_->lock == this)
    {
        scn_->lock = RT_NULL;
    }
}

static Object *new_object(std::vector<Object *> &own, Object::Kind kind,
                          Object *parent, rt_OBJECT *obj)
{
    if (obj == nullptr)
    {
        throw Error("null-pointer in object");
    }
    Object *o = new Object();
    own.push_back(o);
    o->kind = kind;
    o->tag = obj->obj.tag;
    o->parent = parent;
    o->obj = obj;
    o->otm = obj->trm;
    o->trm = &obj->trm;
    memcpy(o->mtx, iden4, sizeof(rt_mat4));
    memcpy(o->inv, iden4, sizeof(rt_mat4));
    obj->time = -1;
    return o;
}

/* QuadRay rt_Array::rt_Array, core/engine/object.cpp:1183-1647 */
Object *Engine::make_array(Object *parent, rt_OBJECT *obj)
{
    Object *a = new_object(objects_, Object::ARRAY, parent, obj);
    arrays_.push_back(a);

    rt_OBJECT *arr = (rt_OBJECT *)obj->obj.pobj;
    for (int i = 0; i < obj->obj.obj_num; i++)
    {
        Object *c = make_object(a, &arr[i]);
        if (c != nullptr)
        {
            a->children.push_back(c);
        }
    }

    /* emission of LIGHT materials: the light accumulated in this array
     * (object.cpp:1347-1388), used by the path tracer */
    for (Object *c : a->children)
    {
        if (c->kind != Object::SURFACE) continue;
        Material *ms[2] = {&c->outer, &c->inner};
        for (Material *m : ms)
        {
            if (m->props & QR_PROP_LIGHT)
            {
                m->rec.col[0] = a->col_hdr[RT_R] * 100.0f;
                m->rec.col[1] = a->col_hdr[RT_G] * 100.0f;
                m->rec.col[2] = a->col_hdr[RT_B] * 100.0f;
                m->rec.e_src  = a->col_hdr[RT_A] * 100.0f;
            }
            else
            {
                m->rec.col[0] = m->rec.col[1] = m->rec.col[2] = 0.0f;
                m->rec.e_src = 0.0f;
            }
        }
    }

    build_relations(a, obj);
    return a;
}

Object *Engine::make_object(Object *parent, rt_OBJECT *obj)
{
    switch (obj->obj.tag)
    {
        case RT_TAG_ARRAY:
        return make_array(parent, obj);

        case RT_TAG_CAMERA:
        {
            /* QuadRay rt_Camera::rt_Camera, object.cpp:426-452 */
            Object *o = new_object(objects_, Object::CAMERA, parent, obj);
            o->cam = (rt_CAMERA *)obj->obj.pobj;
            rt_CAMERA *cam = o->cam;
            if (cam->col.val != 0x0)
            {
                cam->col.hdr[RT_R] = ((cam->col.val >> 0x10) & 0xFF) / 255.0f;
                cam->col.hdr[RT_G] = ((cam->col.val >> 0x08) & 0xFF) / 255.0f;
                cam->col.hdr[RT_B] = ((cam->col.val >> 0x00) & 0xFF) / 255.0f;
            }
            o->pov = cam->vpt[0] <= 0.0f ? 1.0f :
                     cam->vpt[0] <= 2 * 0.01f ? 2 * 0.01f : cam->vpt[0];
            cameras_.push_back(o);
            return o;
        }

        case RT_TAG_LIGHT:
        {
            /* QuadRay rt_Light::rt_Light, object.cpp:590-631 */
            Object *o = new_object(objects_, Object::LIGHT, parent, obj);
            o->lgt = (rt_LIGHT *)obj->obj.pobj;
            rt_LIGHT *lgt = o->lgt;
            if (lgt->col.val != 0x0)
            {
                lgt->col.hdr[RT_R] = ((lgt->col.val >> 0x10) & 0xFF) / 255.0f;
                lgt->col.hdr[RT_G] = ((lgt->col.val >> 0x08) & 0xFF) / 255.0f;
                lgt->col.hdr[RT_B] = ((lgt->col.val >> 0x00) & 0xFF) / 255.0f;
            }
            qr_light &r = o->lrec;
            memset(&r, 0, sizeof(r));
            r.t_max = 1.0f;
            r.col[0] = lgt->col.hdr[RT_R] * lgt->lum[1];
            r.col[1] = lgt->col.hdr[RT_G] * lgt->lum[1];
            r.col[2] = lgt->col.hdr[RT_B] * lgt->lum[1];
            r.l_src = lgt->lum[1];
            r.a_qdr = lgt->atn[3];
            r.a_lnr = lgt->atn[2];
            r.a_cnt = lgt->atn[1] + 1.0f;
            r.a_rng = lgt->atn[0];

            parent->col_hdr[RT_R] += r.col[0];
            parent->col_hdr[RT_G] += r.col[1];
            parent->col_hdr[RT_B] += r.col[2];
            parent->col_hdr[RT_R] += lgt->col.hdr[RT_R] * lgt->lum[0];
            parent->col_hdr[RT_G] += lgt->col.hdr[RT_G] * lgt->lum[0];
            parent->col_hdr[RT_B] += lgt->col.hdr[RT_B] * lgt->lum[0];
            parent->col_hdr[RT_A] += lgt->lum[0] + lgt->lum[1];
            lights_.push_back(o);
            return o;
        }

        case RT_TAG_PLANE:
        case RT_TAG_CYLINDER:
        case RT_TAG_SPHERE:
        case RT_TAG_CONE:
        case RT_TAG_PARABOLOID:
        case RT_TAG_HYPERBOLOID:
        case RT_TAG_PARACYLINDER:
        case RT_TAG_HYPERCYLINDER:
        case RT_TAG_HYPERPARABOLOID:
        {
            /* QuadRay rt_Surface::rt_Surface, object.cpp:2366-2402 */
            Object *o = new_object(objects_, Object::SURFACE, parent, obj);
            o->srf = (rt_SURFACE *)obj->obj.pobj;
            init_material(o->outer, &o->srf->side_outer,
                          obj->obj.pmat_outer ? obj->obj.pmat_outer :
                                                o->srf->side_outer.pmat);
            init_material(o->inner, &o->srf->side_inner,
                          obj->obj.pmat_inner ? obj->obj.pmat_inner :
                                                o->srf->side_inner.pmat);
            o->id = (int)surfaces_.size();
            surfaces_.push_back(o);
            return o;
        }

        default:
        return nullptr;
    }
}

/* QuadRay rt_Material::rt_Material, object.cpp:4038-4180 */
void Engine::init_material(Material &m, const rt_SIDE *sd, const rt_MATERIAL *mat)
{
    if (mat == nullptr)
    {
        throw Error("null-pointer in material");
    }
    m.mat = mat;
    m.sd = sd;
    memset(&m.rec, 0, sizeof(m.rec));

    const rt_TEX *tx = &mat->tex;
    if (tx->x_dim == 0 && tx->y_dim == 0 && tx->ptex == nullptr)
    {
        m.x_dim = m.y_dim = 1;
        m.texels.assign(1, tx->col.val);
    }
    else if (tx->x_dim == 0 && tx->y_dim == 0)
    {
        /* a texture file, loaded once per name (rt_Material::resolve_texture,
         * object.cpp:4201-4225; load_image, rtimag.cpp:47-186) */
        const std::string name = (const char *)tx->ptex;
        auto it = textures_.find(name);
        if (it == textures_.end())
        {
            Texture t;
            if (!bmp_load(data_root() + "data/textures/" + name, t.texels, t.x_dim, t.y_dim))
            {
                throw Error("failed to load image");
            }
            it = textures_.emplace(name, std::move(t)).first;
        }
        m.x_dim = it->second.x_dim;
        m.y_dim = it->second.y_dim;
        m.texels = it->second.texels;
        m.tex_src = &it->second;
    }
    else
    {
        m.x_dim = tx->x_dim;
        m.y_dim = tx->y_dim;
        const uint32_t *p = (const uint32_t *)tx->ptex;
        m.texels.assign(p, p + (size_t)m.x_dim * m.y_dim);
        m.tex_src = p;
    }

    uint32_t props = QR_PROP_NORMAL;
    props |= (opts_ & RT_OPTS_GAMMA) == 0 ? QR_PROP_GAMMA : 0;
    props |= (opts_ & RT_OPTS_FRESNEL) == 0 ? QR_PROP_FRESNEL : 0;
    props |= mat->tag == RT_MAT_LIGHT ? QR_PROP_LIGHT : 0;
    props |= mat->tag == RT_MAT_METAL ? QR_PROP_METAL : 0;
    props |= mat->prp[1] == 0.0f ? QR_PROP_OPAQUE : 0;
    props |= mat->prp[1] == 1.0f ? QR_PROP_TRANSP : 0;
    props |= m.x_dim == 1 && m.y_dim == 1 ? 0 : QR_PROP_TEXTURE;
    props |= mat->prp[0] == 0.0f ? 0 : QR_PROP_REFLECT;
    props |= mat->prp[2] == 1.0f ? 0 : QR_PROP_REFRACT;
    props |= mat->lgt[0] == 0.0f ? 0 : QR_PROP_DIFFUSE;
    props |= mat->lgt[1] == 0.0f ? 0 : QR_PROP_SPECULAR;
    if (mat->prp[0] + mat->prp[1] >= 1.0f)
    {
        props &= ~(QR_PROP_DIFFUSE | QR_PROP_SPECULAR);
    }
    m.props = props;

    /* texture side transform: axis map of a multiple-of-90 rotation */
    float t[2][2];
    t[0][0] = +cosa(sd->rot);
    t[0][1] = +sina(sd->rot);
    t[1][0] = -sina(sd->rot);
    t[1][1] = +cosa(sd->rot);

    int sgn[2] = {1, 1}, match = 0;
    for (int i = 0; i < 2; i++)
    {
        for (int j = 0; j < 2; j++)
        {
            if (fabsf(t[i][0]) == iden4[j][0] && fabsf(t[i][1]) == iden4[j][1])
            {
                m.map[i] = j;
                sgn[i] = isign(t[i][j]);
                match++;
            }
        }
    }
    if (match < 2)
    {
        m.map[0] = 0; sgn[0] = 1;
        m.map[1] = 1; sgn[1] = 1;
    }

    qr_material &r = m.rec;
    r.tmap[0] = m.map[0];
    r.tmap[1] = m.map[1];
    m.scl[0] = m.x_dim / (sd->scl[RT_X] * sgn[0]);
    m.scl[1] = m.y_dim / (sd->scl[RT_Y] * sgn[1]);
    r.xscal = m.scl[0];
    r.yscal = m.scl[1];
    r.xoffs = sd->pos[m.map[0]];
    r.yoffs = sd->pos[m.map[1]];
    r.xmask = (uint32_t)(m.x_dim - 1);
    r.ymask = (uint32_t)(m.y_dim - 1);
    int xd = m.x_dim, lg2 = 0;
    while (xd >>= 1) lg2++;
    r.yshft = (uint32_t)lg2;

    float f = 1.0f - (mat->prp[0] + mat->prp[1]);
    f = fmaxq(f, 0.0f);
    r.l_dff = mat->lgt[0] * f;
    r.l_spc = mat->lgt[1] * f;
    r.l_pow = (uint32_t)(mat->lgt[2] * 16.0f);
    r.c_rfl = mat->prp[0];
    r.c_trn = mat->prp[1];
    r.c_rfr = mat->prp[2];
    r.rfr_2 = mat->prp[2] * mat->prp[2];
    r.c_rcp = 1.0f / mat->prp[2];
    r.ext_2 = mat->prp[3] * mat->prp[3];
    if (!(mat->prp[1] == 0.0f || mat->prp[1] == 1.0f || mat->prp[2] != 1.0f))
    {
        r.c_rfr = mat->prp[3];
        r.rfr_2 = mat->prp[3] * mat->prp[3];
    }
}

/* relations of an array's scene data into per-surface relation templates
 * (QuadRay object.cpp:1390-1601) */
void Engine::build_relations(Object *a, rt_OBJECT *obj)
{
    const rt_RELATION *rel = obj->obj.prel;
    int acc = 0;
    std::vector<RelElem> seg;

    const std::vector<Object *> *sel_l = &a->children;
    const std::vector<Object *> *sel_r = &a->children;

    for (int i = 0; i < obj->obj.rel_num; i++)
    {
        const rt_RELATION &r = rel[i];
        if (r.obj1 >= (int)sel_l->size() || r.obj2 >= (int)sel_r->size())
        {
            continue;
        }

        std::vector<RelElem> elm;
        bool bound = false;

        switch (r.rel)
        {
            case RT_REL_INDEX_ARRAY:
            if (r.obj1 >= 0 && r.obj2 >= -1 &&
                (*sel_l)[r.obj1]->kind == Object::ARRAY)
            {
                sel_l = &(*sel_l)[r.obj1]->children;
            }
            if (r.obj1 >= -1 && r.obj2 >= 0 &&
                (*sel_r)[r.obj2]->kind == Object::ARRAY)
            {
                sel_r = &(*sel_r)[r.obj2]->children;
            }
            break;

            case RT_REL_MINUS_INNER:
            case RT_REL_MINUS_OUTER:
            {
                if (r.obj1 == -1 && r.obj2 >= 0 && acc == 0)
                {
                    acc = 1;
                    seg.clear();
                    seg.push_back({-1, nullptr}); /* accumulator enter */
                }
                RelElem e = {0, nullptr};
                if (r.obj1 >= -1 && r.obj2 >= 0)
                {
                    e = {r.rel, (*sel_r)[r.obj2]};
                    sel_r = &a->children;
                    elm.push_back(e);
                }
                if (r.obj1 == -1 && r.obj2 >= 0)
                {
                    seg.push_back(e);
                }
            }
            break;

            case RT_REL_MINUS_ACCUM:
            if (r.obj1 >= 0 && r.obj2 == -1 && acc == 1)
            {
                acc = 0;
                seg.push_back({+1, nullptr}); /* accumulator leave */
                elm = seg;
                seg.clear();
            }
            break;

            case RT_REL_BOUND_ARRAY:
            case RT_REL_UNTIE_ARRAY:
            if (r.obj1 == -1 && r.obj2 == -1)
            {
                bound = true;
            }
            if (r.obj1 == -1 && r.obj2 >= 0 &&
                (*sel_r)[r.obj2]->kind == Object::ARRAY)
            {
                bound = true;
            }
            break;

            case RT_REL_BOUND_INDEX:
            case RT_REL_UNTIE_INDEX:
            if (r.obj1 == -1 && r.obj2 >= 0)
            {
                bound = true;
            }
            if (r.obj1 >= 0 && r.obj2 >= 0 &&
                (*sel_l)[r.obj1]->kind == Object::ARRAY)
            {
                bound = true;
            }
            break;

            default:
            break;
        }

        if (r.obj1 >= 0 && !elm.empty())
        {
            add_relation((*sel_l)[r.obj1], elm);
            sel_l = &a->children;
        }
        if (bound)
        {
            /* bounding volumes only steer CPU list pruning; the selection
             * resets are kept so later relations index the same objects */
            if (r.obj1 >= 0) sel_l = &a->children;
            if (r.obj2 >= 0) sel_r = &a->children;
        }
    }
}

/* QuadRay rt_Surface::add_relation / rt_Array::add_relation,
 * object.cpp:1652-1662, 2407-2479: elements are inserted at the head */
void Engine::add_relation(Object *o, const std::vector<RelElem> &lst)
{
    if (o->kind == Object::ARRAY)
    {
        for (Object *c : o->children)
        {
            add_relation(c, lst);
        }
        return;
    }
    if (o->kind != Object::SURFACE)
    {
        return;
    }
    for (const RelElem &e : lst)
    {
        if (e.target == nullptr)
        {
            o->rel.insert(o->rel.begin(), e);
        }
        else if (e.target->kind == Object::ARRAY)
        {
            for (Object *c : e.target->children)
            {
                std::vector<RelElem> one(1, RelElem{e.data, c});
                add_relation(o, one);
            }
        }
        else if (e.target->kind == Object::SURFACE)
        {
            o->rel.insert(o->rel.begin(), e);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* update stage                                                             */
/* ------------------------------------------------------------------------ */

int Engine::set_opts(int opts)
{
    opts_ = (int)(opts & ~scn_->opts);
    return opts;
}

/* QuadRay rt_Camera::update_action, object.cpp:495-573: the camera's
 * transform in the caller's scene data moves by its per-50-ms speeds (dps:
 * right, forward, up; drt: turn, tilt) times the time since the camera's
 * last update, along the heading of that update; the tilt stays within
 * [-180, 0] degrees and the turn within (-180, 180) */
void Engine::update_action(rt_time time, int action)
{
    if (cam_ == nullptr) return;
    Object *c = cam_;
    rt_TRANSFORM3D *trm = c->trm;
    const rt_CAMERA *cam = c->cam;
    const float t = (time - c->obj->time) / 50.0f;
    switch (action)
    {
        case 3:     /* RT_CAMERA_MOVE_UP */
        trm->pos[RT_Z] += cam->dps[RT_K] * t;
        break;
        case 2:     /* RT_CAMERA_MOVE_DOWN */
        trm->pos[RT_Z] -= cam->dps[RT_K] * t;
        break;
        case 4:     /* RT_CAMERA_MOVE_LEFT */
        trm->pos[RT_X] -= cam->dps[RT_I] * t * c->hor_cos;
        trm->pos[RT_Y] -= cam->dps[RT_I] * t * c->hor_sin;
        break;
        case 5:     /* RT_CAMERA_MOVE_RIGHT */
        trm->pos[RT_X] += cam->dps[RT_I] * t * c->hor_cos;
        trm->pos[RT_Y] += cam->dps[RT_I] * t * c->hor_sin;
        break;
        case 6:     /* RT_CAMERA_MOVE_BACK */
        trm->pos[RT_X] += cam->dps[RT_J] * t * c->hor_sin;
        trm->pos[RT_Y] -= cam->dps[RT_J] * t * c->hor_cos;
        break;
        case 7:     /* RT_CAMERA_MOVE_FORWARD */
        trm->pos[RT_X] -= cam->dps[RT_J] * t * c->hor_sin;
        trm->pos[RT_Y] += cam->dps[RT_J] * t * c->hor_cos;
        break;
        case 12:    /* RT_CAMERA_ROTATE_LEFT */
        trm->rot[RT_Z] += cam->drt[RT_I] * t;
        if (trm->rot[RT_Z] >= +180.0f) trm->rot[RT_Z] -= +360.0f;
        break;
        case 13:    /* RT_CAMERA_ROTATE_RIGHT */
        trm->rot[RT_Z] -= cam->drt[RT_I] * t;
        if (trm->rot[RT_Z] <= -180.0f) trm->rot[RT_Z] += +360.0f;
        break;
        case 11:    /* RT_CAMERA_ROTATE_UP */
        if (trm->rot[RT_X] < 0.0f)
        {
            trm->rot[RT_X] += cam->drt[RT_J] * t;
            if (trm->rot[RT_X] > 0.0f) trm->rot[RT_X] = 0.0f;
        }
        break;
        case 10:    /* RT_CAMERA_ROTATE_DOWN */
        if (trm->rot[RT_X] > -180.0f)
        {
            trm->rot[RT_X] -= cam->drt[RT_J] * t;
            if (trm->rot[RT_X] < -180.0f) trm->rot[RT_X] = -180.0f;
        }
        break;
        default:    /* crouch, jump, lean: no movement in QuadRay either */
        break;
    }
}

int Engine::next_cam()
{
    /* camera list runs from the last constructed camera backwards */
    int n = (int)cameras_.size();
    if (cam_idx_ + 1 < n)
    {
        cam_idx_++;
    }
    else
    {
        cam_idx_ = 0;
    }
    cam_ = cameras_[n - 1 - cam_idx_];
    return cam_idx_;
}

/* QuadRay rt_Object::update_status, object.cpp:175-216 */
void Engine::update_status(Object *o, rt_time time, int flags, Object *trnode)
{
    rt_OBJECT *obj = o->obj;
    if (obj->f_anim != nullptr && obj->time != time)
    {
        obj->f_anim(time, obj->time < 0 ? 0 : obj->time, o->trm, RT_NULL);
    }
    obj->time = time;

    o->obj_changed = flags & FLAG_OBJ;
    if (obj->f_anim != nullptr)
    {
        o->obj_changed |= FLAG_OBJ;
    }
    if (o->obj_changed == 0)
    {
        return;
    }
    o->obj_has_trm = (flags & FLAG_SCL) | (flags & FLAG_ROT);
    o->trnode = trnode;
}

/* QuadRay rt_Object::update_matrix, object.cpp:221-389 */
void Engine::update_matrix(Object *o, const float (*mtx)[4])
{
    if (o->obj_changed == 0)
    {
        return;
    }
    const rt_TRANSFORM3D *trm = o->trm;
    int i, c;

    o->mtx_has_trm = 0;

    const float fsc[] = {-1.0f, +1.0f};
    for (i = 0, c = 0; i < 2; i++)
    {
        if (trm->scl[RT_X] == fsc[i]) c++;
        if (trm->scl[RT_Y] == fsc[i]) c++;
        if (trm->scl[RT_Z] == fsc[i]) c++;
    }
    o->mtx_has_trm |= (c == 3) ? 0 : FLAG_SCL;

    const float frt[] = {-270.0f, -180.0f, -90.0f, 0.0f, +90.0f, +180.0f, +270.0f};
    for (i = 0, c = 0; i < 7; i++)
    {
        if (trm->rot[RT_X] == frt[i]) c++;
        if (trm->rot[RT_Y] == frt[i]) c++;
        if (trm->rot[RT_Z] == frt[i]) c++;
    }
    o->mtx_has_trm |= (c == 3) ? 0 : FLAG_ROT;

    if ((o->mtx_has_trm & FLAG_ROT) == 0)
    {
        rt_mat4 tm;
        mat_from_transform(tm, trm, true);
        mat_mul_mat(o->mtx, mtx, tm);

        if (o->obj_has_trm == FLAG_SCL)
        {
            o->mtx_has_trm = o->obj_has_trm;
            o->obj_has_trm = 0;
        }

        /* axis mapping of a trivial transform */
        for (i = 0; i < 3; i++)
        {
            for (int j = 0; j < 3; j++)
            {
                if ((o->mtx[i][0] != 0.0f) == (iden4[j][0] != 0.0f) &&
                    (o->mtx[i][1] != 0.0f) == (iden4[j][1] != 0.0f) &&
                    (o->mtx[i][2] != 0.0f) == (iden4[j][2] != 0.0f))
                {
                    o->map[i] = j;
                    o->sgn[i] = isign(o->mtx[i][j]);
                    o->scl[j] = fabsf(o->mtx[i][j]);
                }
            }
        }
        o->map[RT_L] = RT_W;
        o->sgn[RT_L] = 1;
        o->scl[RT_W] = 1.0f;
    }

    if ((o->mtx_has_trm & FLAG_ROT) != 0 && o->trnode == nullptr)
    {
        rt_mat4 tm;
        mat_from_transform(tm, trm, false);
        mat_mul_mat(o->mtx, mtx, tm);
    }
    if ((o->mtx_has_trm & FLAG_ROT) != 0 && o->trnode != nullptr)
    {
        rt_mat4 tm, tmp;
        mat_from_transform(tm, trm, false);
        mat_mul_mat(tmp, o->trnode->mtx, mtx);
        mat_mul_mat(o->mtx, tmp, tm);
    }
    if ((o->mtx_has_trm & FLAG_ROT) != 0)
    {
        o->trnode = o;
        o->obj_has_trm |= FLAG_ROT;
        for (i = 0; i < 4; i++)
        {
            o->map[i] = i;
            o->sgn[i] = 1;
        }
        o->scl[RT_X] = trm->scl[RT_X];
        o->scl[RT_Y] = trm->scl[RT_Y];
        o->scl[RT_Z] = trm->scl[RT_Z];
        o->scl[RT_W] = 1.0f;
    }

    if ((o->obj_has_trm & FLAG_ROT) != 0 && (opts_ & RT_OPTS_FSCALE) == 0)
    {
        o->obj_has_trm |= FLAG_SCL;
    }

    bool tarray = (opts_ & RT_OPTS_TARRAY) != 0;
    if (o->trnode != nullptr && o->trnode != o &&
        (!tarray || o->tag > RT_TAG_SURFACE_MAX))
    {
        rt_mat4 tmp;
        mat_mul_mat(tmp, o->trnode->mtx, o->mtx);
        memcpy(o->mtx, tmp, sizeof(rt_mat4));

        o->trnode = o;
        o->obj_has_trm |= o->mtx_has_trm;
        for (i = 0; i < 4; i++)
        {
            o->map[i] = i;
            o->sgn[i] = 1;
            o->scl[i] = 1.0f;
        }
    }
}

/* QuadRay update_object of arrays (object.cpp:1683-1781), surfaces, cameras
 * and lights (matrix deferred to update_fields) */
void Engine::update_object(Object *o, rt_time time, int flags, Object *trnode,
                           const float (*mtx)[4])
{
    if (o->kind != Object::ARRAY)
    {
        update_status(o, time, flags, trnode);
        o->pmtx = mtx;
        return;
    }

    /* arrays always refresh the whole subtree here: the tables are rebuilt
     * from scratch on every update */
    update_status(o, time, flags | FLAG_OBJ, trnode);
    update_matrix(o, mtx);

    o->cmtx = o->mtx;
    if (o->trnode == o)
    {
        memcpy(o->scm, iden4, sizeof(rt_mat4));
        o->scm[0][0] = o->scl[0];
        o->scm[1][1] = o->scl[1];
        o->scm[2][2] = o->scl[2];
        o->cmtx = o->scm;
    }

    for (Object *c : o->children)
    {
        update_object(c, time, flags | o->mtx_has_trm | o->obj_changed,
                      o->trnode, o->cmtx);
    }
}

void Engine::update_fields(Object *o)
{
    if (o->obj_changed == 0)
    {
        return;
    }
    switch (o->kind)
    {
        case Object::ARRAY:
        if (o->trnode == o)
        {
            mat_inverse(o->inv, o->mtx);
        }
        break;

        case Object::CAMERA:
        update_matrix(o, o->pmtx);
        /* the heading camera actions move along (object.cpp:485-486) */
        o->hor_sin = sina(o->trm->rot[RT_Z]);
        o->hor_cos = cosa(o->trm->rot[RT_Z]);
        break;

        case Object::LIGHT:
        update_matrix(o, o->pmtx);
        o->lrec.pos[0] = o->pos()[RT_X];
        o->lrec.pos[1] = o->pos()[RT_Y];
        o->lrec.pos[2] = o->pos()[RT_Z];
        break;

        case Object::SURFACE:
        /* the record, coefficients and clip box follow from the matrix
         * (fields.h, at emit_tables or on the device) */
        update_matrix(o, o->pmtx);
        break;
    }
}

/* ------------------------------------------------------------------------ */
/* custom clip lists (QuadRay rt_SceneThread::sclip, engine.cpp:1850-1979)  */
/* ------------------------------------------------------------------------ */

void Engine::build_clip_list(Object *s, std::vector<ClipElem> &out)
{
    /* serial numbers identify elements while positions shift on insertion */
    struct E { ClipElem c; int serial; int last_serial; };
    std::vector<E> lst;
    int serial = 0;

    for (const RelElem &r : s->rel)
    {
        if (r.target == nullptr)
        {
            E e = {{r.data < 0 ? QR_CLIP_ENTER : QR_CLIP_LEAVE, nullptr, 0, -1},
                   serial++, -1};
            lst.insert(lst.begin(), e);
            continue;
        }
        if (r.target->kind != Object::SURFACE)
        {
            continue;
        }
        Object *c = r.target;
        E e = {{QR_CLIP_SURFACE, c, r.data, -1}, serial++, -1};

        if (c->trnode != nullptr && c->trnode != c)
        {
            int acc = 0, found = -1;
            for (int k = 0; k < (int)lst.size(); k++)
            {
                const ClipElem &n = lst[k].c;
                if (acc == 0 && n.kind == QR_CLIP_TRNODE && n.obj == c->trnode)
                {
                    found = k;
                    break;
                }
                if (n.kind == QR_CLIP_SURFACE || n.kind == QR_CLIP_TRNODE)
                {
                    continue;
                }
                if (acc == 0 && n.kind == QR_CLIP_LEAVE)
                {
                    break;
                }
                if (acc == 0 && n.kind == QR_CLIP_ENTER)
                {
                    acc = 1;
                }
                if (acc == 1 && n.kind == QR_CLIP_LEAVE)
                {
                    acc = 0;
                }
            }
            if (found < 0)
            {
                lst.insert(lst.begin(), e);
                E t = {{QR_CLIP_TRNODE, c->trnode, 0, -1}, serial++, e.serial};
                lst.insert(lst.begin(), t);
            }
            else
            {
                lst.insert(lst.begin() + found + 1, e);
            }
        }
        else
        {
            lst.insert(lst.begin(), e);
        }
    }

    out.clear();
    for (const E &e : lst)
    {
        out.push_back(e.c);
    }
    for (size_t k = 0; k < lst.size(); k++)
    {
        if (lst[k].c.kind != QR_CLIP_TRNODE) continue;
        for (size_t m = 0; m < lst.size(); m++)
        {
            if (lst[m].serial == lst[k].last_serial)
            {
                out[k].last = (int)m;
            }
        }
    }
}

/* surface list: surfaces grouped under their transform node, QuadRay
 * rt_SceneThread::insert/filter (engine.cpp:1145-1242, 1683-1764) with the
 * surface registry traversed from its head (last constructed first) */
void Engine::build_surface_list()
{
    /* A lone surface is a node of its own; the grouped surfaces of a
     * transform node form one node, placed where the group's last surface
     * (in construction order) stands and holding its surfaces in that
     * order.  (This is the order QuadRay's list gets by walking the
     * surfaces backwards and inserting each node at the front.) */
    const bool tarray = (opts_ & RT_OPTS_TARRAY) != 0;
    auto grouped = [&](const Object *s)
    {
        return tarray && s->trnode != nullptr && s->trnode != s;
    };
    std::map<const Object *, int> last_of;      /* group -> last surface */
    std::map<const Object *, std::vector<Object *>> members;
    for (int k = 0; k < (int)surfaces_.size(); k++)
    {
        Object *s = surfaces_[k];
        if (!grouped(s)) continue;
        last_of[s->trnode] = k;
        members[s->trnode].push_back(s);
    }

    slist_.clear();
    for (int k = 0; k < (int)surfaces_.size(); k++)
    {
        Object *s = surfaces_[k];
        if (!grouped(s))
        {
            slist_.push_back({s, -1});
            continue;
        }
        if (last_of[s->trnode] != k) continue;
        Object *t = s->trnode;
        const std::vector<Object *> &sub = members[t];
        if (t->kind == Object::ARRAY)
        {
            const int at = (int)slist_.size();
            slist_.push_back({t, at + (int)sub.size()});
            for (Object *m : sub) slist_.push_back({m, -1});
        }
        else
        {
            slist_.push_back({t, -1});
        }
    }
}

void Engine::update(rt_time time)
{
    /* phase 0.5: hierarchical transforms */
    update_object(root_, time, 0, nullptr, iden4);

    /* phase 1: fields of arrays, cameras, lights, surfaces */
    for (Object *a : arrays_) update_fields(a);
    for (Object *c : cameras_) update_fields(c);
    for (Object *l : lights_) update_fields(l);
    for (Object *s : surfaces_) update_fields(s);

    finish_update();
}

/* phase 2 on the host: clip lists, the surface list, the records and the
 * culling programs, and the structure they were built with */
void Engine::finish_update()
{
    std::vector<std::vector<ClipElem>> clips(surfaces_.size());
    for (int k = (int)surfaces_.size() - 1; k >= 0; k--)
    {
        build_clip_list(surfaces_[k], clips[k]);
    }
    build_surface_list();
    emit_tables(clips);

    sig_.clear();
    for (Object *a : arrays_) sig_.push_back({a, a->trnode, a->obj_has_trm});
    for (Object *o : surfaces_) sig_.push_back({o, o->trnode, o->obj_has_trm});
    sig_opts_ = opts_;
    stale_ = false;
}

/* the part of a record's inputs that an update changes */
void Engine::geom_transform(Object *o, qr_geom &g) const
{
    memcpy(g.mtx, o->mtx, sizeof(g.mtx));
    const Object *t = o->trnode;
    memcpy(g.tmtx, t != nullptr ? t->mtx : iden4, sizeof(g.tmtx));
    for (int a = 0; a < 4; a++)
    {
        g.scl[a] = o->scl[a];
        g.map[a] = o->map[a];
        g.sgn[a] = o->sgn[a];
    }
}

bool Engine::update_transforms(rt_time time)
{
    update_object(root_, time, 0, nullptr, iden4);
    for (Object *a : arrays_) update_fields(a);
    for (Object *c : cameras_) update_fields(c);
    for (Object *l : lights_) update_fields(l);
    for (Object *s : surfaces_) update_fields(s);

    bool same = sig_opts_ == opts_ && !sig_.empty() && t_geom_.size() == t_surf_.size();
    for (size_t k = 0; same && k < sig_.size(); k++)
    {
        const Sig &g = sig_[k];
        same = g.trnode == g.o->trnode && g.trm == g.o->obj_has_trm;
    }
    if (!same)
    {
        finish_update();
        return false;
    }
    for (Object *a : arrays_)
    {
        if (a->rec >= 0) geom_transform(a, t_geom_[(size_t)a->rec]);
    }
    for (Object *o : surfaces_) geom_transform(o, t_geom_[(size_t)o->rec]);
    t_lgt_.clear();
    for (Object *l : lights_) t_lgt_.push_back(l->lrec);
    tab_.lights = t_lgt_.data();
    stale_ = true;
    return true;
}

void Engine::materialize()
{
    if (stale_) finish_update();
}

/* the inputs of one record (fields.h): a surface, or an array that is its
 * own transform node */
void Engine::make_geom(Object *o, qr_geom &g) const
{
    memset(&g, 0, sizeof(g));
    geom_transform(o, g);
    const Object *t = o->trnode;
    g.tn_rec = t != nullptr ? t->rec : -1;
    g.self_tn = t == o ? 1 : 0;
    g.obj_has_trm = o->obj_has_trm;
    g.elem = -1;
    g.mat[0] = g.mat[1] = -1;
    if (o->kind == Object::ARRAY)
    {
        g.tag = RT_TAG_ARRAY;
        g.id = -1;
        return;
    }
    /* the kernels keep one local hit per root: a surface is traced in a
     * transform node's frame exactly when it has a transform (a rotation
     * always makes a node, and everything under a node inherits the
     * rotation's flag; a scale alone folds into the axis map) */
    if ((g.obj_has_trm != 0) != (g.tn_rec >= 0))
    {
        throw Error("surface transform state not supported");
    }
    g.tag = o->tag;
    g.id = o->id;
    for (int a = 0; a < 3; a++)
    {
        g.smin[a] = o->srf->min[a];
        g.smax[a] = o->srf->max[a];
    }
    const void *p = o->obj->obj.pobj;
    switch (o->tag)
    {
        case RT_TAG_CYLINDER: g.prm[0] = ((const rt_CYLINDER *)p)->rad; break;
        case RT_TAG_SPHERE: g.prm[0] = ((const rt_SPHERE *)p)->rad; break;
        case RT_TAG_CONE: g.prm[0] = ((const rt_CONE *)p)->rat; break;
        case RT_TAG_PARABOLOID: g.prm[0] = ((const rt_PARABOLOID *)p)->par; break;
        case RT_TAG_PARACYLINDER: g.prm[0] = ((const rt_PARACYLINDER *)p)->par; break;
        case RT_TAG_HYPERBOLOID:
        g.prm[0] = ((const rt_HYPERBOLOID *)p)->rat;
        g.prm[1] = ((const rt_HYPERBOLOID *)p)->hyp;
        break;
        case RT_TAG_HYPERCYLINDER:
        g.prm[0] = ((const rt_HYPERCYLINDER *)p)->rat;
        g.prm[1] = ((const rt_HYPERCYLINDER *)p)->hyp;
        break;
        case RT_TAG_HYPERPARABOLOID:
        g.prm[0] = ((const rt_HYPERPARABOLOID *)p)->pr1;
        g.prm[1] = ((const rt_HYPERPARABOLOID *)p)->pr2;
        break;
        default: break;
    }
    const Material *ms[2] = {&o->outer, &o->inner};
    for (int side = 0; side < 2; side++)
    {
        const Material *m = ms[side];
        g.tex[side][0] = m->scl[0];
        g.tex[side][1] = m->scl[1];
        g.tex[side][2] = m->sd->pos[m->map[0]];
        g.tex[side][3] = m->sd->pos[m->map[1]];
        g.tex_map[side][0] = m->map[0];
        g.tex_map[side][1] = m->map[1];
        g.props[side] = m->props;
    }
}

void Engine::emit_tables(const std::vector<std::vector<ClipElem>> &clips)
{
    t_surf_.clear();
    t_mat_.clear();
    t_lgt_.clear();
    t_clip_.clear();
    t_elem_.clear();
    t_texel_.clear();
    std::map<std::pair<const void *, size_t>, int32_t> tex_at;

    for (Object *a : arrays_) a->rec = -1;

    /* records: surfaces in construction order, then trnode arrays */
    for (Object *s : surfaces_) s->rec = (int)t_surf_.size(), t_surf_.emplace_back();
    for (Object *a : arrays_)
    {
        if (a->trnode == a)
        {
            a->rec = (int)t_surf_.size();
            t_surf_.emplace_back();
        }
    }

    /* every record's inputs */
    t_geom_.assign(t_surf_.size(), qr_geom{});
    for (Object *a : arrays_)
    {
        if (a->rec >= 0) make_geom(a, t_geom_[a->rec]);
    }
    for (Object *s : surfaces_) make_geom(s, t_geom_[s->rec]);

    for (size_t k = 0; k < surfaces_.size(); k++)
    {
        Object *s = surfaces_[k];
        qr_geom &g = t_geom_[s->rec];

        /* materials and texels: a texture shared by several materials is
         * stored once */
        Material *ms[2] = {&s->outer, &s->inner};
        for (int side = 0; side < 2; side++)
        {
            Material *m = ms[side];
            auto key = std::make_pair(m->tex_src, m->texels.size());
            auto at = m->tex_src != nullptr ? tex_at.find(key) : tex_at.end();
            if (at != tex_at.end())
            {
                m->rec.tex_off = at->second;
            }
            else
            {
                m->rec.tex_off = (int32_t)t_texel_.size();
                t_texel_.insert(t_texel_.end(), m->texels.begin(), m->texels.end());
                if (m->tex_src != nullptr) tex_at.emplace(key, m->rec.tex_off);
            }
            g.mat[side] = (int32_t)t_mat_.size();
            t_mat_.push_back(m->rec);
        }

        /* custom clip list */
        g.clip_beg = (int32_t)t_clip_.size();
        g.clip_cnt = (int32_t)clips[k].size();
        g.adjust = (opts_ & RT_OPTS_ADJUST) != 0 && !clips[k].empty() && s->trnode != s;
        for (const ClipElem &e : clips[k])
        {
            qr_clip c;
            c.kind = e.kind;
            c.surf = e.obj != nullptr ? e.obj->rec : -1;
            c.rel = e.kind == QR_CLIP_SURFACE ? e.rel :
                    e.kind == QR_CLIP_TRNODE ? e.last : 0;
            c.pad = 0;
            t_clip_.push_back(c);
        }
    }

    /* the records themselves (the device update kernel runs the same) */
    t_box_.assign(4 * t_surf_.size(), {0.0f, 0.0f, 0.0f, 0.0f});
    for (size_t r = 0; r < t_surf_.size(); r++)
    {
        const qr_geom &g = t_geom_[r];
        if (g.tag == RT_TAG_ARRAY)
        {
            qf_array_record(g, t_surf_[r]);
            continue;
        }
        float sci[4], scj[4];
        qf_shape(g, sci, scj);
        qf_bounds(t_geom_.data(), t_clip_.data(), (int)r, t_box_[4 * r].v, t_box_[4 * r + 1].v,
                  t_box_[4 * r + 2].v, t_box_[4 * r + 3].v);
        qf_surface_record(g, sci, scj, t_box_[4 * r].v, t_box_[4 * r + 1].v,
                          t_box_[4 * r + 2].v, t_box_[4 * r + 3].v, t_surf_[r]);
        if (g.tag == RT_TAG_PLANE)
        {
            for (int side = 0; side < 2; side++)
            {
                float tx[4];
                qf_plane_tex(g, side, tx);
                qr_material &m = t_mat_[(size_t)g.mat[side]];
                m.xscal = tx[0];
                m.yscal = tx[1];
                m.xoffs = tx[2];
                m.yoffs = tx[3];
            }
        }
    }

    for (Object *l : lights_)
    {
        t_lgt_.push_back(l->lrec);
    }
    for (const ListElem &e : slist_)
    {
        t_geom_[(size_t)e.obj->rec].elem = (int32_t)t_elem_.size();
        t_elem_.push_back({e.obj->rec, e.last});
    }

    tab_.n_surf = (int32_t)t_surf_.size();
    tab_.n_mat = (int32_t)t_mat_.size();
    tab_.n_lgt = (int32_t)t_lgt_.size();
    tab_.n_clip = (int32_t)t_clip_.size();
    tab_.n_elem = (int32_t)t_elem_.size();
    tab_.n_texel = (int32_t)t_texel_.size();
    tab_.n_user = (int32_t)surfaces_.size();
    tab_.pad = 0;
    tab_.surfaces = t_surf_.data();
    tab_.materials = t_mat_.data();
    tab_.lights = t_lgt_.data();
    tab_.clips = t_clip_.data();
    tab_.elems = t_elem_.data();
    tab_.texels = t_texel_.data();

    /* the culling programs: refitted when the elements and their boxed
     * steps are those the current programs were built for, else built */
    std::vector<int32_t> key;
    elem_inputs(key);
    if (!t_step_.empty() && key == topo_key_)
    {
        refit_cull_program();
        refit_ = true;
    }
    else
    {
        build_cull_program();
        topo_key_.swap(key);
        topo_id_++;
        refit_ = false;
        /* reference point of the world boxes' allowance: the middle of their
         * centres' span (kept while the programs are refitted: any point
         * gives a box at least as large as the allowance requires) */
        float lo[3] = {QR_INF, QR_INF, QR_INF}, hi[3] = {-QR_INF, -QR_INF, -QR_INF};
        for (const qr_cull_step &st : t_step_)
        {
            if (st.elem >= 0) continue;
            for (int a = 0; a < 3; a++)
            {
                if (!std::isfinite(st.c[a])) continue;
                lo[a] = std::min(lo[a], st.c[a]);
                hi[a] = std::max(hi[a], st.c[a]);
            }
        }
        for (int a = 0; a < 3; a++) cull_ref_[a] = lo[a] <= hi[a] ? 0.5f * (lo[a] + hi[a]) : 0.0f;
        cull_ref_[3] = 0.0f;
    }
    tab_.n_step = (int32_t)t_step_.size();
    tab_.pad2 = 0;
    tab_.steps = t_step_.data();
    tab_.n_step_any = (int32_t)t_step_any_.size();
    tab_.pad3 = 0;
    tab_.steps_any = t_step_any_.data();
}

/* ------------------------------------------------------------------------ */
/* culling program (no counterpart in QuadRay's list walk)                  */
/* ------------------------------------------------------------------------ */

/* Per-element inputs of the programs' boxes (qr_refit_elem) and the key
 * the programs are built for:
 *  - world box: every hit the tracer can accept on a surface lies in the
 *    shape's extent within its own clip box (the bound QuadRay derives for
 *    its tiling, before clipper tightening), mapped through the transform
 *    node and grown for fp32 rounding of computed hit points
 *    (refit_world_box); none if the extent is unbounded;
 *  - local box: for a transformed surface, the same extent in its own
 *    tracing frame (the frame of the walk's DFF_I / RAY_I registers: the
 *    transform node's axes, origin at the surface), grown by 2e-3 of its
 *    magnitude, as centre and half extents.
 * The key holds, per element, its record, group end and which boxes it
 * has: the programs' structure depends on nothing else. */
void Engine::elem_inputs(std::vector<int32_t> &key)
{
    const int n = (int)t_elem_.size();
    t_refit_.assign((size_t)n, qr_refit_elem{});
    key.assign(4 * (size_t)n + 1, 0);
    key[4 * (size_t)n] = (int32_t)t_surf_.size();
    for (int e = 0; e < n; e++)
    {
        qr_refit_elem &r = t_refit_[(size_t)e];
        const size_t rec = (size_t)t_elem_[e].surf;
        const qr_geom &g = t_geom_[rec];
        if (g.tag != RT_TAG_ARRAY)
        {
            qf_refit_elem(g, t_box_[4 * rec].v, t_box_[4 * rec + 1].v, t_box_[4 * rec + 2].v,
                          t_box_[4 * rec + 3].v, r);
        }
        key[4 * (size_t)e + 0] = t_elem_[e].surf;
        key[4 * (size_t)e + 1] = t_elem_[e].last;
        key[4 * (size_t)e + 2] = r.world;
        key[4 * (size_t)e + 3] = r.local;
    }
}

/* The programs' boxes again from the current element inputs, their steps
 * unchanged: a box step bounds the world boxes of the elements it ranges
 * over (a group's transform node has none), an element step with a local
 * box takes its element's.  A box is the union the build would have formed
 * for the same ranges. */
void Engine::refit_cull_program()
{
    const int n = (int)t_refit_.size();
    std::vector<float> wb(6 * (size_t)n);
    std::vector<char> ok((size_t)n);
    for (int e = 0; e < n; e++)
    {
        ok[(size_t)e] = (char)refit_world_box(&t_refit_[(size_t)e], &wb[6 * (size_t)e],
                                              &wb[6 * (size_t)e + 3]);
    }
    for (std::vector<qr_cull_step> *prog : {&t_step_, &t_step_any_})
    {
        std::vector<qr_cull_step> &ps = *prog;
        for (size_t p = 0; p < ps.size(); p++)
        {
            qr_cull_step &st = ps[p];
            if (st.elem >= 0)
            {
                if (st.skip == -2)
                {
                    const qr_refit_elem &r = t_refit_[(size_t)st.elem];
                    for (int a = 0; a < 3; a++)
                    {
                        st.c[a] = r.lc[a];
                        st.h[a] = r.lh[a];
                    }
                }
                continue;
            }
            float lo[3] = {QR_INF, QR_INF, QR_INF}, hi[3] = {-QR_INF, -QR_INF, -QR_INF};
            for (int q = (int)p + 1; q < st.skip; q++)
            {
                const int e = ps[(size_t)q].elem;
                if (e < 0 || !ok[(size_t)e]) continue;
                for (int a = 0; a < 3; a++)
                {
                    lo[a] = std::min(lo[a], wb[6 * (size_t)e + a]);
                    hi[a] = std::max(hi[a], wb[6 * (size_t)e + 3 + a]);
                }
            }
            refit_centre(lo, hi, st.c, st.h);
        }
    }
}

/* Two culling programs over the top-level units of the list (a transform-
 * node group, kept whole and in list order, or a lone surface):
 *
 *   t_step_      nearest-hit walks: units in list order, under a tree of
 *                contiguous ranges split where the summed surface area of
 *                the two halves' boxes, weighted by their element counts,
 *                is smallest (so a large surface such as a ground plane gets
 *                a range of its own instead of widening every box above it);
 *   t_step_any_  any-hit (shadow) walks, whose result does not depend on the
 *                order of units: a spatial tree, units ordered by the same
 *                cost over their box centres along the best axis.
 *
 * A range box is emitted only where it is clearly smaller than the box above
 * it.  Every unit has its world box; a transformed surface's element step
 * carries its box in its own frame (skip -2), tested after the element's
 * transform, and an untransformed grouped surface gets a world box step. */
void Engine::build_cull_program()
{
    struct Box { float lo[3], hi[3]; bool ok; };
    struct Unit { int beg, end; Box box; };   /* elements [beg, end] */

    const int n = (int)t_elem_.size();
    std::vector<Box> ebox(n);
    for (int e = 0; e < n; e++)
    {
        ebox[e].ok = refit_world_box(&t_refit_[(size_t)e], ebox[e].lo, ebox[e].hi) != 0;
    }
    auto join = [](Box &d, const Box &b)
    {
        if (!d.ok || !b.ok) { d.ok = false; return; }
        for (int a = 0; a < 3; a++)
        {
            d.lo[a] = std::min(d.lo[a], b.lo[a]);
            d.hi[a] = std::max(d.hi[a], b.hi[a]);
        }
    };
    /* surface area (the chance a random ray meets the box); unbounded: huge */
    auto area = [](const Box &b) -> double
    {
        if (!b.ok) return 1e30;
        const double x = (double)b.hi[0] - b.lo[0], y = (double)b.hi[1] - b.lo[1],
                     z = (double)b.hi[2] - b.lo[2];
        return 2.0 * (x * y + y * z + z * x);
    };
    constexpr double SHRINK = 0.5;     /* emit a range box below this area ratio */
    constexpr double SHRINK_ANY = 0.5; /* the same for the any-hit program
                                          (0.25-0.9 measured within 1%) */

    std::vector<Unit> units;
    for (int e = 0; e < n; )
    {
        Unit u;
        u.beg = e;
        u.end = t_elem_[e].last >= e ? t_elem_[e].last : e;
        if (u.end > u.beg)
        {
            u.box.ok = true;
            bool first = true;
            for (int k = u.beg + 1; k <= u.end; k++)
            {
                if (first) { u.box = ebox[k]; first = false; }
                else join(u.box, ebox[k]);
            }
        }
        else
        {
            u.box = ebox[e];
        }
        units.push_back(u);
        e = u.end + 1;
    }

    std::vector<qr_cull_step> *out = nullptr;
    auto box_step = [&](const Box &b) -> size_t
    {
        qr_cull_step st;
        refit_centre(b.lo, b.hi, st.c, st.h);
        st.elem = -1;
        st.skip = -1;
        out->push_back(st);
        return out->size() - 1;
    };
    auto elem_step = [&](int e)
    {
        qr_cull_step st;
        memset(&st, 0, sizeof(st));
        st.elem = e;
        st.skip = -1;
        const qr_refit_elem &r = t_refit_[(size_t)e];
        if (r.local)
        {
            for (int a = 0; a < 3; a++)
            {
                st.c[a] = r.lc[a];
                st.h[a] = r.lh[a];
            }
            st.skip = -2;
        }
        out->push_back(st);
    };
    auto emit_unit = [&](const Unit &u)
    {
        size_t b = u.box.ok ? box_step(u.box) : (size_t)-1;
        elem_step(u.beg);
        for (int k = u.beg + 1; k <= u.end; k++)
        {
            const bool local = t_refit_[(size_t)k].local != 0;
            size_t kb = ebox[k].ok && !local ? box_step(ebox[k]) : (size_t)-1;
            elem_step(k);
            if (kb != (size_t)-1) (*out)[kb].skip = (int32_t)out->size();
        }
        if (b != (size_t)-1) (*out)[b].skip = (int32_t)out->size();
    };
    auto weight = [&](const Unit &u) { return (double)(u.end - u.beg + 1); };

    /* a tree over units ids[lo, hi): best split point by area times weight;
     * "spatial" first orders the range along the best axis */
    std::vector<int> ids(units.size());
    std::function<void(int, int, double, bool)> emit =
        [&](int lo, int hi, double above, bool spatial)
    {
        if (hi - lo == 1)
        {
            emit_unit(units[ids[lo]]);
            return;
        }
        Box b = units[ids[lo]].box;
        for (int k = lo + 1; k < hi; k++) join(b, units[ids[k]].box);
        const double ab = area(b);
        const double shrink = spatial ? SHRINK_ANY : SHRINK;
        size_t st = b.ok && ab < shrink * above ? box_step(b) : (size_t)-1;
        const double here = st != (size_t)-1 ? ab : above;

        const int m = hi - lo;
        std::vector<double> pre(m + 1), suf(m + 1);
        double best = 1e300;
        int best_k = lo + m / 2, best_axis = -1;
        for (int axis = spatial ? 0 : -1; axis < (spatial ? 3 : 0); axis++)
        {
            if (axis >= 0)
            {
                auto key = [&](int id)
                {
                    const Box &x = units[id].box;
                    return x.ok ? 0.5 * ((double)x.lo[axis] + x.hi[axis]) : -1e30;
                };
                std::stable_sort(ids.begin() + lo, ids.begin() + hi,
                                 [&](int p, int q) { return key(p) < key(q); });
            }
            Box acc;
            double w = 0.0;
            pre[0] = 0.0;
            for (int k = 0; k < m; k++)
            {
                const Unit &u = units[ids[lo + k]];
                if (k == 0) acc = u.box; else join(acc, u.box);
                w += weight(u);
                pre[k + 1] = area(acc) * w;
            }
            w = 0.0;
            suf[m] = 0.0;
            for (int k = m - 1; k >= 0; k--)
            {
                const Unit &u = units[ids[lo + k]];
                if (k == m - 1) acc = u.box; else join(acc, u.box);
                w += weight(u);
                suf[k] = area(acc) * w;
            }
            for (int k = 1; k < m; k++)
            {
                const double c = pre[k] + suf[k];
                if (c < best)
                {
                    best = c;
                    best_k = lo + k;
                    best_axis = axis;
                }
            }
        }
        if (spatial && best_axis != 2)
        {
            /* restore the chosen axis' order (the last sort was along z) */
            auto key = [&](int id)
            {
                const Box &x = units[id].box;
                return x.ok ? 0.5 * ((double)x.lo[best_axis] + x.hi[best_axis]) : -1e30;
            };
            std::stable_sort(ids.begin() + lo, ids.begin() + hi,
                             [&](int p, int q) { return key(p) < key(q); });
        }
        emit(lo, best_k, here, spatial);
        emit(best_k, hi, here, spatial);
        if (st != (size_t)-1) (*out)[st].skip = (int32_t)out->size();
    };

    t_step_.clear();
    t_step_any_.clear();
    if (units.empty()) return;
    for (int pass = 0; pass < 2; pass++)
    {
        out = pass == 0 ? &t_step_ : &t_step_any_;
        for (size_t k = 0; k < units.size(); k++) ids[k] = (int)k;
        emit(0, (int)units.size(), 1e300, pass == 1);
    }
}

void Engine::camera_pose(float *pose) const
{
    for (int r = 0; r < 4; r++)
    {
        for (int a = 0; a < 3; a++) pose[3 * r + a] = cam_->mtx[r][a];
    }
}

/* camera basis and frame stepping (QuadRay engine.cpp:3078-3098,
 * 3301-3318, 3514-3697) */
void Engine::frame_desc(qr_frame_desc *fd, int fsaa, int depth, int x_res,
                        int y_res, const float *pose) const
{
    memset(fd, 0, sizeof(*fd));
    /* another resolution than the scene's: its camera steps as rt_Scene's
     * constructor derives them for that size (engine.cpp:2900-2904) */
    float factor = factor_, aspect = aspect_;
    if (x_res <= 0 || y_res <= 0)
    {
        x_res = x_res_;
        y_res = y_res_;
    }
    else
    {
        factor = 1.0f / (float)x_res;
        aspect = (float)y_res * factor;
    }
    const Object *c = cam_;
    const float *pos = c->mtx[3];
    const float *hor0 = c->mtx[0];
    const float *ver0 = c->mtx[1];
    const float *nrm = c->mtx[2];
    if (pose != nullptr)
    {
        hor0 = pose;
        ver0 = pose + 3;
        nrm = pose + 6;
        pos = pose + 9;
    }

    float h = -0.5f * 1.0f;
    float v = -0.5f * aspect;
    float dir[3], hor[3], ver[3];
    for (int a = 0; a < 3; a++) dir[a] = nrm[a] * c->pov;
    for (int a = 0; a < 3; a++) dir[a] += hor0[a] * h;
    for (int a = 0; a < 3; a++) dir[a] += ver0[a] * v;
    for (int a = 0; a < 3; a++) hor[a] = hor0[a] * factor;
    for (int a = 0; a < 3; a++) ver[a] = ver0[a] * factor;
    for (int a = 0; a < 3; a++) dir[a] += hor[a] * 0.5f;
    for (int a = 0; a < 3; a++) dir[a] += ver[a] * 0.5f;

    for (int a = 0; a < 3; a++)
    {
        fd->org[a] = pos[a];
        fd->dir[a] = dir[a];
        fd->hor[a] = hor[a];
        fd->ver[a] = ver[a];
    }

    /* ambient: camera, then lights from the registry head (last first) */
    const rt_CAMERA *cam = c->cam;
    for (int a = 0; a < 3; a++) fd->amb[a] = cam->col.hdr[a] * cam->lum[0];
    fd->amb[3] = cam->lum[0];
    for (int k = (int)lights_.size() - 1; k >= 0; k--)
    {
        const rt_LIGHT *l = lights_[k]->lgt;
        for (int a = 0; a < 3; a++) fd->amb[a] += l->col.hdr[a] * l->lum[0];
        fd->amb[3] += l->lum[0];
    }

    fd->t_min = c->pov;
    fd->x_res = x_res;
    fd->y_res = y_res;
    fd->x_row = x_row_;
    fd->fsaa = fsaa;
    fd->depth = depth;
    fd->gamma = (opts_ & RT_OPTS_GAMMA) == 0 ? QR_PROP_GAMMA : 0;

    /* sub-sample grid: rotated for 4X, alternating for 2X */
    const float as = 0.25f, ar = 0.08f;
    if (fsaa == 2)
    {
        fd->aa_h[0] = (-ar - as); fd->aa_v[0] = (+ar - as);
        fd->aa_h[1] = (-ar + as); fd->aa_v[1] = (-ar - as);
        fd->aa_h[2] = (+ar - as); fd->aa_v[2] = (+ar + as);
        fd->aa_h[3] = (+ar + as); fd->aa_v[3] = (-ar + as);
    }
    else if (fsaa == 1)
    {
        fd->aa_h[0] = (-ar + as); fd->aa_v[0] = (+ar + as);
        fd->aa_h[1] = (+ar - as); fd->aa_v[1] = (-ar - as);
        fd->aa_h[2] = (+ar - as); fd->aa_v[2] = (+ar + as);
        fd->aa_h[3] = (-ar + as); fd->aa_v[3] = (-ar - as);
    }
}

} /* namespace qr */
