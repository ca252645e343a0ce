/*
 * scenegen.hpp -- seeded synthetic scenes in the rt_SCENE struct format.
 *
 * Header-only and written against the struct names of the scene format
 * only, so the same source builds against qr_format.h (the MI355X renderer)
 * and against QuadRay's own core/engine/format.h (the CPU reference driver
 * under oracle/).  All generated parameters are multiples of 1/256, exact in
 * fp32 and fp64 builds alike.
 *
 * Workloads:
 *   1  plain scene: checker ground plane, 3 spheres, a clipped cylinder and
 *      cone, one white light; diffuse + specular, hard shadows
 *   2  Whitted scene: 64 quadrics of 7 kinds in a 3-level transform
 *      hierarchy, a quarter clipped by boolean relations, 4 coloured lights
 *      with attenuation, metal and glass with Fresnel
 *   3  path-traced box: 6 walls with an emissive patch, 16 diffuse, mirror
 *      and glass surfaces
 *   4, 5 as 2 with 16 / 24 surfaces per level-2 array
 *   6  the 64-surface Whitted scene with glass everywhere
 *   7, 8 Whitted scenes of 8 to 512 / 16 to 1024 quadrics (batches, serving)
 *   9  every surface kind, apexes on the camera's axis, gamma on
 *   10 coincident surfaces (which of two equal hit distances wins)
 *   11+ as 2 with (config - 2) * 8 surfaces per level-2 array
 */
#ifndef QR_SCENEGEN_HPP
#define QR_SCENEGEN_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <memory>
#include <string>
#include <vector>

#include "image.h"

#ifndef RT_OPTS_PRUNING
#define RT_OPTS_PRUNING (RT_OPTS_TILING | RT_OPTS_TILING_EXT1 | RT_OPTS_VARRAY | \
                         RT_OPTS_SHADOW | RT_OPTS_SHADOW_EXT1 |                 \
                         RT_OPTS_SHADOW_EXT2 | RT_OPTS_2SIDED |                 \
                         RT_OPTS_2SIDED_EXT1 | RT_OPTS_2SIDED_EXT2)
#endif

namespace qrgen {

/* split-mix 64 generator */
struct Rng
{
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x1234567ull) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    /* uniform in [a, b), quantised to 1/256 */
    double u(double a, double b)
    {
        double x = (double)(next() >> 11) * (1.0 / 9007199254740992.0);
        double v = a + (b - a) * x;
        return std::floor(v * 256.0) / 256.0;
    }
    int i(int n) { return (int)(next() % (uint64_t)n); }
    /* three draws in a fixed order (function arguments are unsequenced) */
    void u3(double *v, double a0, double b0, double a1, double b1,
            double a2, double b2)
    {
        v[0] = u(a0, b0);
        v[1] = u(a1, b1);
        v[2] = u(a2, b2);
    }
};

struct Scene
{
    rt_SCENE root;
    std::deque<rt_OBJECT *> objarrs;        /* owned arrays of objects */
    std::deque<std::unique_ptr<unsigned char[]>> blobs;
    std::vector<rt_OBJECT *> surfaces;      /* surface objects, scene order */
    std::vector<rt_MATERIAL *> id_mats;     /* first-hit id materials */
    std::vector<rt_MATERIAL *> tex_mats;    /* materials with texel arrays */
    int x_res = 0, y_res = 0, fsaa = 0, depth = 4, pt = 0;

    template <class Tp> Tp *alloc(size_t n = 1)
    {
        unsigned char *p = new unsigned char[sizeof(Tp) * n];
        std::memset(p, 0, sizeof(Tp) * n);
        blobs.emplace_back(p);
        return (Tp *)p;
    }
    rt_OBJECT *objs(size_t n)
    {
        rt_OBJECT *p = alloc<rt_OBJECT>(n);
        return p;
    }
};

static inline void set_trm(rt_OBJECT *o, double sx, double sy, double sz,
                           double rx, double ry, double rz,
                           double px, double py, double pz)
{
    o->trm.scl[0] = (rt_real)sx; o->trm.scl[1] = (rt_real)sy; o->trm.scl[2] = (rt_real)sz;
    o->trm.rot[0] = (rt_real)rx; o->trm.rot[1] = (rt_real)ry; o->trm.rot[2] = (rt_real)rz;
    o->trm.pos[0] = (rt_real)px; o->trm.pos[1] = (rt_real)py; o->trm.pos[2] = (rt_real)pz;
}

static inline void set_obj(rt_OBJECT *o, int tag, void *p, int n = 1)
{
    o->obj.tag = tag;
    o->obj.pobj = p;
    o->obj.obj_num = n;
    o->obj.prel = RT_NULL;
    o->obj.rel_num = 0;
    o->obj.pmat_outer = RT_NULL;
    o->obj.pmat_inner = RT_NULL;
    o->f_anim = RT_NULL;
    o->time = 0;
}

static inline rt_MATERIAL *material(Scene &sc, int tag, uint32_t color,
                                    double dff, double spc, double pw,
                                    double rfl, double trn, double rfr,
                                    double ext = 0.0)
{
    rt_MATERIAL *m = sc.alloc<rt_MATERIAL>();
    m->tag = tag;
    m->tex.tag = RT_TEX_PCOLOR;
    m->tex.col.val = color;
    m->tex.ptex = RT_NULL;
    m->lgt[0] = (rt_real)dff; m->lgt[1] = (rt_real)spc; m->lgt[2] = (rt_real)pw;
    m->prp[0] = (rt_real)rfl; m->prp[1] = (rt_real)trn; m->prp[2] = (rt_real)rfr;
    m->prp[3] = (rt_real)ext;
    return m;
}

static inline rt_MATERIAL *checker(Scene &sc, uint32_t c0, uint32_t c1)
{
    rt_ui32 *tx = sc.alloc<rt_ui32>(4);
    tx[0] = c0; tx[1] = c1; tx[2] = c1; tx[3] = c0;
    rt_MATERIAL *m = material(sc, RT_MAT_PLAIN, 0, 1.0, 0.3, 16.0, 0.0, 0.0, 1.0);
    m->tex.col.val = 0;
    m->tex.ptex = tx;
    m->tex.x_dim = 2;
    m->tex.y_dim = 2;
    sc.tex_mats.push_back(m);
    return m;
}

static inline void side(rt_SIDE *sd, rt_MATERIAL *m, double scl = 1.0)
{
    sd->scl[0] = (rt_real)scl; sd->scl[1] = (rt_real)scl;
    sd->rot = 0;
    sd->pos[0] = 0; sd->pos[1] = 0;
    sd->pmat = m;
}

static inline void minmax(rt_SURFACE *s, double i0, double i1, double j0,
                          double j1, double k0, double k1)
{
    s->min[0] = (rt_real)i0; s->max[0] = (rt_real)i1;
    s->min[1] = (rt_real)j0; s->max[1] = (rt_real)j1;
    s->min[2] = (rt_real)k0; s->max[2] = (rt_real)k1;
}

static const double INF = (double)RT_INF;

/* a surface of kind 0..6 (plane, cylinder, sphere, cone, paraboloid,
 * one-sheet hyperboloid, two-sheet hyperboloid) or 7..11 (double cone as a
 * hyperboloid with hyp 0, paracylinder, hypercylinder, two crossing planes
 * as a hypercylinder with hyp 0, hyperparaboloid) with bounded extent */
static inline void make_surface(Scene &sc, rt_OBJECT *o, int kind, double size,
                                rt_MATERIAL *mo, rt_MATERIAL *mi)
{
    double h = size;
    switch (kind)
    {
        case 0:
        {
            rt_PLANE *p = sc.alloc<rt_PLANE>();
            minmax(&p->srf, -h, +h, -h, +h, -INF, +INF);
            side(&p->srf.side_outer, mo); side(&p->srf.side_inner, mi);
            set_obj(o, RT_TAG_PLANE, p);
        }
        break;
        case 1:
        {
            rt_CYLINDER *p = sc.alloc<rt_CYLINDER>();
            minmax(&p->srf, -INF, +INF, -INF, +INF, -h, +h);
            side(&p->srf.side_outer, mo); side(&p->srf.side_inner, mi);
            p->rad = (rt_real)(std::floor(size * 0.5 * 256.0) / 256.0);
            set_obj(o, RT_TAG_CYLINDER, p);
        }
        break;
        case 2:
        {
            rt_SPHERE *p = sc.alloc<rt_SPHERE>();
            minmax(&p->srf, -INF, +INF, -INF, +INF, -INF, +INF);
            side(&p->srf.side_outer, mo); side(&p->srf.side_inner, mi);
            p->rad = (rt_real)(std::floor(size * 0.75 * 256.0) / 256.0);
            set_obj(o, RT_TAG_SPHERE, p);
        }
        break;
        case 3:
        {
            rt_CONE *p = sc.alloc<rt_CONE>();
            minmax(&p->srf, -INF, +INF, -INF, +INF, -h, +h * 0.5);
            side(&p->srf.side_outer, mo); side(&p->srf.side_inner, mi);
            p->rat = (rt_real)0.5;
            set_obj(o, RT_TAG_CONE, p);
        }
        break;
        case 4:
        {
            rt_PARABOLOID *p = sc.alloc<rt_PARABOLOID>();
            minmax(&p->srf, -INF, +INF, -INF, +INF, -INF, +h);
            side(&p->srf.side_outer, mo); side(&p->srf.side_inner, mi);
            p->par = (rt_real)0.75;
            set_obj(o, RT_TAG_PARABOLOID, p);
        }
        break;
        case 5:
        {
            rt_HYPERBOLOID *p = sc.alloc<rt_HYPERBOLOID>();
            minmax(&p->srf, -INF, +INF, -INF, +INF, -h, +h);
            side(&p->srf.side_outer, mo); side(&p->srf.side_inner, mi);
            p->rat = (rt_real)0.75;
            p->hyp = (rt_real)(std::floor(size * 0.25 * 256.0) / 256.0 + 0.0625);
            set_obj(o, RT_TAG_HYPERBOLOID, p);
        }
        break;
        case 6:
        {
            rt_HYPERBOLOID *p = sc.alloc<rt_HYPERBOLOID>();
            minmax(&p->srf, -INF, +INF, -INF, +INF, -h * 1.5, +h * 1.5);
            side(&p->srf.side_outer, mo); side(&p->srf.side_inner, mi);
            p->rat = (rt_real)1.0;
            p->hyp = (rt_real)-(std::floor(size * 0.25 * 256.0) / 256.0 + 0.0625);
            set_obj(o, RT_TAG_HYPERBOLOID, p);
        }
        break;
        case 7:
        {
            rt_HYPERBOLOID *p = sc.alloc<rt_HYPERBOLOID>();
            minmax(&p->srf, -INF, +INF, -INF, +INF, -h, +h);
            side(&p->srf.side_outer, mo); side(&p->srf.side_inner, mi);
            p->rat = (rt_real)0.75;
            p->hyp = (rt_real)0.0;
            set_obj(o, RT_TAG_HYPERBOLOID, p);
        }
        break;
        case 8:
        {
            rt_PARACYLINDER *p = sc.alloc<rt_PARACYLINDER>();
            minmax(&p->srf, -h, +h, -h, +h, -INF, +h);
            side(&p->srf.side_outer, mo); side(&p->srf.side_inner, mi);
            p->par = (rt_real)0.5;
            set_obj(o, RT_TAG_PARACYLINDER, p);
        }
        break;
        case 9:
        {
            rt_HYPERCYLINDER *p = sc.alloc<rt_HYPERCYLINDER>();
            minmax(&p->srf, -INF, +INF, -h, +h, -h, +h);
            side(&p->srf.side_outer, mo); side(&p->srf.side_inner, mi);
            p->rat = (rt_real)0.75;
            p->hyp = (rt_real)(std::floor(size * 0.25 * 256.0) / 256.0 + 0.0625);
            set_obj(o, RT_TAG_HYPERCYLINDER, p);
        }
        break;
        case 10:
        {
            rt_HYPERCYLINDER *p = sc.alloc<rt_HYPERCYLINDER>();
            minmax(&p->srf, -h, +h, -h, +h, -h, +h);
            side(&p->srf.side_outer, mo); side(&p->srf.side_inner, mi);
            p->rat = (rt_real)1.25;
            p->hyp = (rt_real)0.0;
            set_obj(o, RT_TAG_HYPERCYLINDER, p);
        }
        break;
        default:
        {
            rt_HYPERPARABOLOID *p = sc.alloc<rt_HYPERPARABOLOID>();
            minmax(&p->srf, -h, +h, -h, +h, -INF, +INF);
            side(&p->srf.side_outer, mo); side(&p->srf.side_inner, mi);
            p->pr1 = (rt_real)0.5;
            p->pr2 = (rt_real)1.5;
            set_obj(o, RT_TAG_HYPERPARABOLOID, p);
        }
        break;
    }
    sc.surfaces.push_back(o);
}

static inline rt_CAMERA *camera(Scene &sc, double amb)
{
    rt_CAMERA *c = sc.alloc<rt_CAMERA>();
    c->tag = RT_CAM_PLAIN;
    c->col.val = 0xFFFFFFFF;
    c->lum[0] = (rt_real)amb;
    c->vpt[0] = (rt_real)1.0;
    for (int a = 0; a < 3; a++) { c->dps[a] = (rt_real)0.5; c->drt[a] = (rt_real)1.0; }
    return c;
}

static inline rt_LIGHT *light(Scene &sc, uint32_t color, double amb, double src,
                              double cnt, double lnr, double qdr)
{
    rt_LIGHT *l = sc.alloc<rt_LIGHT>();
    l->tag = RT_LGT_PLAIN;
    l->col.val = color;
    l->lum[0] = (rt_real)amb; l->lum[1] = (rt_real)src;
    l->atn[0] = 0; l->atn[1] = (rt_real)cnt; l->atn[2] = (rt_real)lnr;
    l->atn[3] = (rt_real)qdr;
    return l;
}

/* workload 1 (800x480 plain scene) */
static inline void build_plain(Scene &sc, uint64_t seed)
{
    Rng r(seed);
    double p3[3];
    rt_OBJECT *tree = sc.objs(9);
    int n = 0;

    rt_OBJECT *o = &tree[n++];
    set_trm(o, 1, 1, 1, -105, 0, 0, 0, -14, 7);
    set_obj(o, RT_TAG_CAMERA, camera(sc, 0.1));

    o = &tree[n++];
    r.u3(p3, -4, 4, -6, -2, 6, 9);
    set_trm(o, 1, 1, 1, 0, 0, 0, p3[0], p3[1], p3[2]);
    set_obj(o, RT_TAG_LIGHT, light(sc, 0xFFFFFFFF, 0.05, 1.5, 0.2, 0.05, 0.01));

    rt_MATERIAL *floor = checker(sc, 0xFFD0D0D0, 0xFF404040);
    rt_MATERIAL *under = material(sc, RT_MAT_PLAIN, 0xFF303030, 1.0, 0.0, 1.0, 0, 0, 1);
    o = &tree[n++];
    set_trm(o, 1, 1, 1, 0, 0, 0, 0, 0, 0);
    {
        rt_PLANE *p = sc.alloc<rt_PLANE>();
        minmax(&p->srf, -INF, +INF, -INF, +INF, -INF, +INF);
        side(&p->srf.side_outer, floor, 2.0);
        side(&p->srf.side_inner, under);
        set_obj(o, RT_TAG_PLANE, p);
        sc.surfaces.push_back(o);
    }

    const uint32_t cols[5] = {0xFFC02020, 0xFF20A040, 0xFF2040C0, 0xFFC0A020, 0xFF9030B0};
    for (int k = 0; k < 3; k++)
    {
        rt_MATERIAL *m = material(sc, RT_MAT_PLAIN, cols[k], 0.9, 0.5, 12.0, 0, 0, 1);
        o = &tree[n++];
        r.u3(p3, -5, 5, -5, 5, 0.5, 3);
        set_trm(o, 1, 1, 1, 0, 0, 0, p3[0], p3[1], p3[2]);
        rt_SPHERE *p = sc.alloc<rt_SPHERE>();
        minmax(&p->srf, -INF, +INF, -INF, +INF, -INF, +INF);
        side(&p->srf.side_outer, m); side(&p->srf.side_inner, m);
        p->rad = (rt_real)r.u(0.5, 1.5);
        set_obj(o, RT_TAG_SPHERE, p);
        sc.surfaces.push_back(o);
    }

    rt_MATERIAL *mc = material(sc, RT_MAT_PLAIN, cols[3], 0.9, 0.5, 20.0, 0, 0, 1);
    o = &tree[n++];
    r.u3(p3, -5, 5, -5, 5, 0, 1);
    set_trm(o, 1, 1, 1, 0, 0, 0, p3[0], p3[1], 1.25);
    {
        rt_CYLINDER *p = sc.alloc<rt_CYLINDER>();
        minmax(&p->srf, -INF, +INF, -INF, +INF, -1.25, +1.25);
        side(&p->srf.side_outer, mc); side(&p->srf.side_inner, mc);
        p->rad = (rt_real)r.u(0.5, 1.0);
        set_obj(o, RT_TAG_CYLINDER, p);
        sc.surfaces.push_back(o);
    }

    rt_MATERIAL *mn = material(sc, RT_MAT_PLAIN, cols[4], 0.9, 0.5, 20.0, 0, 0, 1);
    o = &tree[n++];
    r.u3(p3, -5, 5, -5, 5, 0, 1);
    set_trm(o, 1, 1, 1, 0, 0, 0, p3[0], p3[1], 2.5);
    {
        rt_CONE *p = sc.alloc<rt_CONE>();
        minmax(&p->srf, -INF, +INF, -INF, +INF, -2.5, -0.25);
        side(&p->srf.side_outer, mn); side(&p->srf.side_inner, mn);
        p->rat = (rt_real)r.u(0.3, 0.6);
        set_obj(o, RT_TAG_CONE, p);
        sc.surfaces.push_back(o);
    }

    sc.root.root.tag = RT_TAG_ARRAY;
    sc.root.root.pobj = tree;
    sc.root.root.obj_num = n;
    sc.root.opts = RT_OPTS_PT | RT_OPTS_PRUNING;
    sc.x_res = 800; sc.y_res = 480; sc.fsaa = 0; sc.depth = 4;
}

/* workload 2 and larger variants (Whitted, 3-level hierarchy) */
static inline void build_whitted(Scene &sc, uint64_t seed, int per_leaf,
                                 bool all_glass = false, int n_mid = 4)
{
    Rng r(seed);
    double p3[3], a3[3];
    const int n_top = 4;

    /* materials: plain, metal, glass (outer air->glass, inner glass->air) */
    const uint32_t cols[8] = {0xFFE03030, 0xFF30C050, 0xFF3060E0, 0xFFE0C040,
                              0xFFB040C0, 0xFF40C0C0, 0xFFE08040, 0xFFD0D0D0};
    std::vector<rt_MATERIAL *> plain, metal, glass_o, glass_i;
    for (int k = 0; k < 8; k++)
    {
        plain.push_back(material(sc, RT_MAT_PLAIN, cols[k], 0.9, 0.6, r.u(4, 32), 0, 0, 1));
        metal.push_back(material(sc, RT_MAT_METAL, cols[k], 0.8, 0.8, r.u(8, 48),
                                 0.6, 0.0, 1.5, 2.5));
    }
    for (int k = 0; k < 4; k++)
    {
        double n = r.u(1.1, 1.6);
        double rfr = std::floor((1.0 / n) * 256.0) / 256.0;
        glass_o.push_back(material(sc, RT_MAT_PLAIN, 0xFFF0F0FF, 0.2, 0.5, 24,
                                   0.0, 0.8, rfr));
        glass_i.push_back(material(sc, RT_MAT_PLAIN, 0xFFF0F0FF, 0.2, 0.5, 24,
                                   0.0, 0.8, n));
    }

    rt_OBJECT *tree = sc.objs(2 + 4 + n_top);
    int n = 0;
    rt_OBJECT *o = &tree[n++];
    set_trm(o, 1, 1, 1, -112, 0, 18, 14, -40, 18);
    set_obj(o, RT_TAG_CAMERA, camera(sc, 0.08));

    const uint32_t lcol[4] = {0xFFFFE0C0, 0xFFC0E0FF, 0xFFFFC0E0, 0xFFE0FFC0};
    for (int k = 0; k < 4; k++)
    {
        o = &tree[n++];
        r.u3(p3, -18, 18, -22, 6, 14, 24);
        set_trm(o, 1, 1, 1, 0, 0, 0, p3[0], p3[1], p3[2]);
        set_obj(o, RT_TAG_LIGHT, light(sc, lcol[k], 0.02, 1.4, 0.3, 0.02, 0.001));
    }

    /* ground */
    o = &tree[n++];
    set_trm(o, 1, 1, 1, 0, 0, 0, 0, 0, -6);
    {
        rt_PLANE *p = sc.alloc<rt_PLANE>();
        minmax(&p->srf, -40, +40, -40, +40, -INF, +INF);
        side(&p->srf.side_outer, checker(sc, 0xFFB0B0B0, 0xFF505060), 4.0);
        side(&p->srf.side_inner, plain[7]);
        set_obj(o, RT_TAG_PLANE, p);
        sc.surfaces.push_back(o);
    }

    for (int a = 0; a < n_top; a++)
    {
        rt_OBJECT *top = &tree[n++];
        r.u3(a3, -40, 40, -40, 40, -180, 180);
        r.u3(p3, 5, 9, 4, 8, 0, 3);
        set_trm(top, 1, 1, 1, a3[0], a3[1], a3[2],
                (a & 1 ? 1 : -1) * p3[0], (a & 2 ? 1 : -1) * p3[1], p3[2]);
        rt_OBJECT *mids = sc.objs(n_mid);
        set_obj(top, RT_TAG_ARRAY, mids, n_mid);

        for (int b = 0; b < n_mid; b++)
        {
            rt_OBJECT *mid = &mids[b];
            double s = r.u(0.75, 1.25);
            if (b & 1)
            {
                /* scale-only sub-array */
                double ry = 90.0 * r.i(4);
                r.u3(p3, -4, 4, -4, 4, -2, 2);
                set_trm(mid, s, s, s, 0, ry, 0, p3[0], p3[1], p3[2]);
            }
            else
            {
                r.u3(a3, -60, 60, -60, 60, -180, 180);
                r.u3(p3, -4, 4, -4, 4, -2, 2);
                set_trm(mid, s, s, s, a3[0], a3[1], a3[2], p3[0], p3[1], p3[2]);
            }
            rt_OBJECT *leaf = sc.objs(per_leaf);
            set_obj(mid, RT_TAG_ARRAY, leaf, per_leaf);

            for (int c = 0; c < per_leaf; c++)
            {
                rt_OBJECT *so = &leaf[c];
                int kind = r.i(7);
                double size = r.u(0.6, 1.6);
                double sx = r.u(0.5, 2.0), sy = r.u(0.5, 2.0), sz = r.u(0.5, 2.0);
                if (kind == 2 || kind == 0) { sy = sx; }
                if (c & 1)
                {
                    r.u3(a3, -180, 180, -180, 180, -180, 180);
                    r.u3(p3, -3, 3, -3, 3, -2, 2);
                    set_trm(so, sx, sy, sz, a3[0], a3[1], a3[2], p3[0], p3[1], p3[2]);
                }
                else
                {
                    double rx = 90.0 * r.i(4);
                    double ry = 90.0 * r.i(4);
                    r.u3(p3, -3, 3, -3, 3, -2, 2);
                    set_trm(so, sx, sy, sz, rx, ry, 0, p3[0], p3[1], p3[2]);
                }
                int mk = r.i(10);
                if (all_glass) mk = 8 + (mk & 1);
                rt_MATERIAL *mo, *mi;
                if (mk < 5)       { mo = plain[r.i(8)]; mi = plain[r.i(8)]; }
                else if (mk < 8)  { mo = metal[r.i(8)]; mi = plain[r.i(8)]; }
                else              { int g = r.i(4); mo = glass_o[g]; mi = glass_i[g]; }
                if (kind == 0 && mk >= 8) { mo = metal[r.i(8)]; }
                make_surface(sc, so, kind, size, mo, mi);
            }

            /* boolean clipping: a quarter of the surfaces get relations */
            int nrel = per_leaf / 4 > 0 ? per_leaf / 4 : 1;
            rt_RELATION *rel = sc.alloc<rt_RELATION>(4 * nrel);
            int nr = 0;
            for (int k = 0; k < nrel; k++)
            {
                int s0 = r.i(per_leaf), s1 = r.i(per_leaf);
                if (s1 == s0) s1 = (s0 + 1) % per_leaf;
                int kind = r.i(3);
                if (kind == 0)
                {
                    rel[nr++] = {s0, RT_REL_MINUS_OUTER, s1};
                }
                else if (kind == 1)
                {
                    rel[nr++] = {s0, RT_REL_MINUS_INNER, s1};
                }
                else
                {
                    int s2 = (s1 + 1) % per_leaf;
                    if (s2 == s0) s2 = (s2 + 1) % per_leaf;
                    rel[nr++] = {-1, RT_REL_MINUS_INNER, s1};
                    rel[nr++] = {-1, RT_REL_MINUS_OUTER, s2};
                    rel[nr++] = {s0, RT_REL_MINUS_ACCUM, -1};
                }
            }
            mid->obj.prel = rel;
            mid->obj.rel_num = nr;
        }
    }

    sc.root.root.tag = RT_TAG_ARRAY;
    sc.root.root.pobj = tree;
    sc.root.root.obj_num = n;
    sc.root.opts = RT_OPTS_PT | RT_OPTS_FRESNEL | RT_OPTS_PRUNING;
    sc.x_res = 1920; sc.y_res = 1080; sc.fsaa = 2; sc.depth = 4;
}

/* workload 3 (path-traced box) */
static inline void build_ptbox(Scene &sc, uint64_t seed)
{
    Rng r(seed);
    double p3[3];
    rt_MATERIAL *white = material(sc, RT_MAT_PLAIN, 0xFFBFBFBF, 1.0, 0.0, 1.0, 0, 0, 1);
    rt_MATERIAL *red = material(sc, RT_MAT_PLAIN, 0xFFBF4040, 1.0, 0.0, 1.0, 0, 0, 1);
    rt_MATERIAL *blue = material(sc, RT_MAT_PLAIN, 0xFF4040BF, 1.0, 0.0, 1.0, 0, 0, 1);
    rt_MATERIAL *lamp = material(sc, RT_MAT_LIGHT, 0xFFFFFFFF, 1.0, 0.0, 1.0, 0, 0, 1);
    rt_MATERIAL *mirror = material(sc, RT_MAT_METAL, 0xFFF0F0F0, 0.0, 0.0, 1.0, 1.0, 0.0, 1.5, 2.0);
    rt_MATERIAL *glass_o = material(sc, RT_MAT_PLAIN, 0xFFFFFFFF, 0.0, 0.0, 1.0, 0.0, 1.0, 0.6640625);
    rt_MATERIAL *glass_i = material(sc, RT_MAT_PLAIN, 0xFFFFFFFF, 0.0, 0.0, 1.0, 0.0, 1.0, 1.5);

    rt_OBJECT *tree = sc.objs(2 + 6 + 16);
    int n = 0;
    rt_OBJECT *o = &tree[n++];
    set_trm(o, 1, 1, 1, -90, 0, 0, 0, -9.5, 5);
    set_obj(o, RT_TAG_CAMERA, camera(sc, 0.0));

    /* lamp array: point light plus emissive patch */
    o = &tree[n++];
    set_trm(o, 1, 1, 1, 0, 0, 0, 0, 0, 9.9375);
    {
        rt_OBJECT *lamp_objs = sc.objs(2);
        set_trm(&lamp_objs[0], 1, 1, 1, 0, 0, 0, 0, 0, -0.5);
        set_obj(&lamp_objs[0], RT_TAG_LIGHT, light(sc, 0xFFFFFFFF, 0.0, 0.12, 0, 0, 0));
        set_trm(&lamp_objs[1], 1, 1, 1, 180, 0, 0, 0, 0, 0);
        rt_PLANE *p = sc.alloc<rt_PLANE>();
        minmax(&p->srf, -1.5, 1.5, -1.5, 1.5, -INF, +INF);
        side(&p->srf.side_outer, lamp); side(&p->srf.side_inner, lamp);
        set_obj(&lamp_objs[1], RT_TAG_PLANE, p);
        sc.surfaces.push_back(&lamp_objs[1]);
        set_obj(o, RT_TAG_ARRAY, lamp_objs, 2);
    }

    /* walls: floor, ceiling, left, right, back, front */
    struct Wall { double rx, ry, px, py, pz; rt_MATERIAL *m; };
    const Wall walls[6] = {
        {0, 0, 0, 0, 0, white}, {180, 0, 0, 0, 10, white},
        {0, 90, -5, 0, 5, red}, {0, -90, 5, 0, 5, blue},
        {90, 0, 0, 5, 5, white}, {-90, 0, 0, -10, 5, white},
    };
    for (int k = 0; k < 6; k++)
    {
        o = &tree[n++];
        set_trm(o, 1, 1, 1, walls[k].rx, walls[k].ry, 0, walls[k].px, walls[k].py, walls[k].pz);
        rt_PLANE *p = sc.alloc<rt_PLANE>();
        minmax(&p->srf, -10, 10, -10, 10, -INF, +INF);
        side(&p->srf.side_outer, walls[k].m); side(&p->srf.side_inner, walls[k].m);
        set_obj(o, RT_TAG_PLANE, p);
        sc.surfaces.push_back(o);
    }

    for (int k = 0; k < 16; k++)
    {
        o = &tree[n++];
        int kind = k < 10 ? 2 : (1 + r.i(6));
        double size = r.u(0.5, 1.0);
        double rx = kind == 2 ? 0 : 90.0 * r.i(4);
        r.u3(p3, -4, 4, -3, 4, 0.75, 6);
        set_trm(o, 1, 1, 1, rx, 0, 0, p3[0], p3[1], p3[2]);
        int mk = r.i(3);
        rt_MATERIAL *mo = mk == 0 ? white : mk == 1 ? mirror : glass_o;
        rt_MATERIAL *mi = mk == 2 ? glass_i : mo;
        make_surface(sc, o, kind, size, mo, mi);
    }

    sc.root.root.tag = RT_TAG_ARRAY;
    sc.root.root.pobj = tree;
    sc.root.root.obj_num = n;
    sc.root.opts = RT_OPTS_PT | RT_OPTS_FRESNEL | RT_OPTS_PRUNING;
    sc.x_res = 1280; sc.y_res = 720; sc.fsaa = 0; sc.depth = 8; sc.pt = 64;
}

/* workload 9 (every surface kind): the twelve kinds of make_surface with
 * seeded scales and rotations (multiples of 90 degrees or arbitrary) at the
 * top level and again inside a rotated array (a transform node) where two of
 * them clip others; a stack of apexes (cone, double cone, crossing planes)
 * on the axis of a camera looking straight down, so that at an odd
 * resolution without AA the centre pixel's ray meets each apex; metal and
 * glass with Fresnel, gamma correction on */
static inline void build_kinds(Scene &sc, uint64_t seed)
{
    Rng r(seed);
    double p3[3], a3[3];
    const uint32_t cols[8] = {0xFFE03030, 0xFF30C050, 0xFF3060E0, 0xFFE0C040,
                              0xFFB040C0, 0xFF40C0C0, 0xFFE08040, 0xFFD0D0D0};
    std::vector<rt_MATERIAL *> plain, metal;
    for (int k = 0; k < 8; k++)
    {
        plain.push_back(material(sc, RT_MAT_PLAIN, cols[k], 0.9, 0.5, r.u(4, 32), 0, 0, 1));
        metal.push_back(material(sc, RT_MAT_METAL, cols[k], 0.7, 0.7, r.u(8, 48),
                                 0.5, 0.0, 1.5, 2.5));
    }
    rt_MATERIAL *glass_o = material(sc, RT_MAT_PLAIN, 0xFFF0F0FF, 0.2, 0.5, 24,
                                    0.0, 0.8, 0.6640625);
    rt_MATERIAL *glass_i = material(sc, RT_MAT_PLAIN, 0xFFF0F0FF, 0.2, 0.5, 24,
                                    0.0, 0.8, 1.5);

    rt_OBJECT *tree = sc.objs(3 + 1 + 12 + 1 + 1);
    int n = 0;
    rt_OBJECT *o = &tree[n++];
    set_trm(o, 1, 1, 1, 180, 0, 0, 0, 0, 24);
    set_obj(o, RT_TAG_CAMERA, camera(sc, 0.1));
    o = &tree[n++];
    set_trm(o, 1, 1, 1, 0, 0, 0, -9, -7, 21);
    set_obj(o, RT_TAG_LIGHT, light(sc, 0xFFFFF0E0, 0.05, 1.3, 0.2, 0.02, 0.001));
    o = &tree[n++];
    set_trm(o, 1, 1, 1, 0, 0, 0, 8, 6, 15);
    set_obj(o, RT_TAG_LIGHT, light(sc, 0xFFC0D0FF, 0.03, 1.0, 0.3, 0.01, 0.002));

    /* checker ground */
    o = &tree[n++];
    set_trm(o, 1, 1, 1, 0, 0, 0, 0, 0, -2);
    {
        rt_PLANE *p = sc.alloc<rt_PLANE>();
        minmax(&p->srf, -30, +30, -30, +30, -INF, +INF);
        side(&p->srf.side_outer, checker(sc, 0xFFB0B0B0, 0xFF505060), 2.0);
        side(&p->srf.side_inner, plain[7]);
        set_obj(o, RT_TAG_PLANE, p);
        sc.surfaces.push_back(o);
    }

    /* the twelve kinds on a ring around the camera axis */
    for (int k = 0; k < 12; k++)
    {
        o = &tree[n++];
        const double ang = k * (2.0 * 3.14159265358979 / 12.0);
        const double px = std::floor(9.0 * std::cos(ang) * 256.0) / 256.0;
        const double py = std::floor(7.0 * std::sin(ang) * 256.0) / 256.0;
        double sx = r.u(0.6, 1.6), sy = r.u(0.6, 1.6), sz = r.u(0.6, 1.6);
        if (k & 1)
        {
            r.u3(a3, -180, 180, -180, 180, -180, 180);
        }
        else
        {
            a3[0] = 90.0 * r.i(4); a3[1] = 90.0 * r.i(4); a3[2] = 90.0 * r.i(4);
        }
        set_trm(o, sx, sy, sz, a3[0], a3[1], a3[2], px, py, r.u(0, 2));
        int mk = r.i(6);
        rt_MATERIAL *mo = mk < 3 ? plain[r.i(8)] : mk < 5 ? metal[r.i(8)] : glass_o;
        rt_MATERIAL *mi = mk == 5 ? glass_i : plain[r.i(8)];
        make_surface(sc, o, k, r.u(1.0, 1.75), mo, mi);
    }

    /* a rotated array holding the five less common kinds; two clip others */
    o = &tree[n++];
    r.u3(a3, -30, 30, -30, 30, -180, 180);
    set_trm(o, 1, 1, 1, a3[0], a3[1], a3[2], 0, -1, 4);
    {
        rt_OBJECT *kids = sc.objs(5);
        for (int k = 0; k < 5; k++)
        {
            r.u3(p3, -4, 4, -3, 3, -1, 1);
            set_trm(&kids[k], 1, 1, 1, 0, 90.0 * r.i(4), 0, p3[0], p3[1], p3[2]);
            rt_MATERIAL *mo = k == 2 ? glass_o : (k & 1 ? metal[r.i(8)] : plain[r.i(8)]);
            rt_MATERIAL *mi = k == 2 ? glass_i : plain[r.i(8)];
            make_surface(sc, &kids[k], 7 + k, r.u(1.0, 1.5), mo, mi);
        }
        set_obj(o, RT_TAG_ARRAY, kids, 5);
        rt_RELATION *rel = sc.alloc<rt_RELATION>(2);
        rel[0] = {0, RT_REL_MINUS_OUTER, 1};
        rel[1] = {3, RT_REL_MINUS_INNER, 4};
        o->obj.prel = rel;
        o->obj.rel_num = 2;
    }

    /* apexes on the camera's axis, the farthest first in the list */
    o = &tree[n++];
    set_trm(o, 1, 1, 1, 0, 0, 0, 0, 0, 0);
    {
        rt_OBJECT *kids = sc.objs(3);
        const int kinds[3] = {3, 10, 7};
        const double z[3] = {-1.0, 3.0, 7.0};
        for (int k = 0; k < 3; k++)
        {
            set_trm(&kids[k], 1, 1, 1, 0, 0, 0, 0, 0, z[k]);
            make_surface(sc, &kids[k], kinds[k], 1.0, metal[k + 2], plain[k + 4]);
        }
        set_obj(o, RT_TAG_ARRAY, kids, 3);
    }

    sc.root.root.tag = RT_TAG_ARRAY;
    sc.root.root.pobj = tree;
    sc.root.root.obj_num = n;
    sc.root.opts = RT_OPTS_PT | RT_OPTS_FRESNEL | RT_OPTS_GAMMA | RT_OPTS_PRUNING;
    sc.x_res = 321; sc.y_res = 181; sc.fsaa = 0; sc.depth = 6;
}

/* workload 10 (coincident surfaces): pairs of identical surfaces with
 * different materials, at the top level of the tree and inside a rotated
 * array (one transform node), seen directly, in a mirror wall and from a
 * diffuse floor (path tracer), so which of two equal hit distances wins is
 * visible wherever they are seen */
static inline void build_ties(Scene &sc, uint64_t seed)
{
    Rng r(seed);
    double p3[3];
    const uint32_t cols[8] = {0xFFE03030, 0xFF30C050, 0xFF3060E0, 0xFFE0C040,
                              0xFFB040C0, 0xFF40C0C0, 0xFFE08040, 0xFFD0D0D0};
    std::vector<rt_MATERIAL *> m;
    for (int k = 0; k < 8; k++)
    {
        m.push_back(material(sc, RT_MAT_PLAIN, cols[k], 0.9, 0.4, 12.0, 0, 0, 1));
    }
    rt_MATERIAL *mirror = material(sc, RT_MAT_METAL, 0xFFF0F0F0, 0.2, 0.2, 8.0,
                                   0.8, 0.0, 1.5, 2.0);
    rt_MATERIAL *glass_o = material(sc, RT_MAT_PLAIN, 0xFFF0F0FF, 0.1, 0.4, 24,
                                    0.0, 0.85, 0.6640625);
    rt_MATERIAL *glass_i = material(sc, RT_MAT_PLAIN, 0xFFF0F0FF, 0.1, 0.4, 24,
                                    0.0, 0.85, 1.5);

    rt_OBJECT *tree = sc.objs(14);
    int n = 0;
    rt_OBJECT *o = &tree[n++];
    set_trm(o, 1, 1, 1, -100, 0, 20, 7, -16, 6);
    set_obj(o, RT_TAG_CAMERA, camera(sc, 0.1));
    o = &tree[n++];
    r.u3(p3, -6, 6, -8, 0, 8, 12);
    set_trm(o, 1, 1, 1, 0, 0, 0, p3[0], p3[1], p3[2]);
    set_obj(o, RT_TAG_LIGHT, light(sc, 0xFFFFFFFF, 0.05, 1.4, 0.2, 0.02, 0.002));

    /* diffuse floor, and a mirror wall behind the pairs */
    o = &tree[n++];
    set_trm(o, 1, 1, 1, 0, 0, 0, 0, 0, 0);
    {
        rt_PLANE *p = sc.alloc<rt_PLANE>();
        minmax(&p->srf, -12, +12, -12, +12, -INF, +INF);
        side(&p->srf.side_outer, m[7]);
        side(&p->srf.side_inner, m[6]);
        set_obj(o, RT_TAG_PLANE, p);
        sc.surfaces.push_back(o);
    }
    o = &tree[n++];
    set_trm(o, 1, 1, 1, 80, 0, 10, 0, 9, 4);
    {
        rt_PLANE *p = sc.alloc<rt_PLANE>();
        minmax(&p->srf, -10, +10, -5, +5, -INF, +INF);
        side(&p->srf.side_outer, mirror);
        side(&p->srf.side_inner, m[5]);
        set_obj(o, RT_TAG_PLANE, p);
        sc.surfaces.push_back(o);
    }
    /* an emissive panel overhead: the path tracer's only light */
    o = &tree[n++];
    set_trm(o, 1, 1, 1, 180, 0, 0, 0, 0, 12);
    {
        rt_MATERIAL *lamp = material(sc, RT_MAT_LIGHT, 0xFFFFFFFF, 1.0, 0.0, 1.0, 0, 0, 1);
        rt_PLANE *p = sc.alloc<rt_PLANE>();
        minmax(&p->srf, -6, +6, -6, +6, -INF, +INF);
        side(&p->srf.side_outer, lamp);
        side(&p->srf.side_inner, lamp);
        set_obj(o, RT_TAG_PLANE, p);
        sc.surfaces.push_back(o);
    }

    /* top-level pairs: spheres, cylinders, walls, glass spheres */
    struct Pair { int kind; double x, y, z, rx, size; int glass; };
    const Pair pairs[4] = {
        {2, -3.0, 1.0, 2.0, 0, 1.5, 0}, {1, 3.0, 2.0, 1.5, 90, 1.25, 0},
        {0, 0.0, 6.0, 3.0, 90, 3.0, 0}, {2, 1.0, -3.0, 1.25, 0, 1.0, 1},
    };
    for (int k = 0; k < 4; k++)
    {
        for (int twin = 0; twin < 2; twin++)
        {
            o = &tree[n++];
            set_trm(o, 1, 1, 1, pairs[k].rx, 0, 0, pairs[k].x, pairs[k].y, pairs[k].z);
            rt_MATERIAL *mo = pairs[k].glass ? glass_o : m[(2 * k + twin) & 7];
            rt_MATERIAL *mi = pairs[k].glass ? glass_i : m[(2 * k + twin + 3) & 7];
            make_surface(sc, o, pairs[k].kind, pairs[k].size, mo, mi);
        }
    }

    /* a rotated array (a transform node) holding two more pairs */
    o = &tree[n++];
    set_trm(o, 1, 1, 1, 0, 0, 35, -4, -4, 1.5);
    {
        rt_OBJECT *kids = sc.objs(4);
        for (int k = 0; k < 4; k++)
        {
            int kind = k < 2 ? 3 : 5;
            set_trm(&kids[k], 1, 1, 1, 0, 0, 0, k < 2 ? -1.5 : 1.5, 0, 0);
            make_surface(sc, &kids[k], kind, 1.0, m[(k + 1) & 7], m[(k + 5) & 7]);
        }
        set_obj(o, RT_TAG_ARRAY, kids, 4);
    }

    sc.root.root.tag = RT_TAG_ARRAY;
    sc.root.root.pobj = tree;
    sc.root.root.obj_num = n;
    sc.root.opts = RT_OPTS_PT | RT_OPTS_FRESNEL | RT_OPTS_PRUNING;
    sc.x_res = 320; sc.y_res = 180; sc.fsaa = 0; sc.depth = 4;
}

static inline Scene *build(int config, uint64_t seed)
{
    Scene *sc = new Scene();
    std::memset(&sc->root, 0, sizeof(sc->root));
    if (config == 1) build_plain(*sc, seed);
    else if (config == 3) build_ptbox(*sc, seed);
    else if (config == 2) build_whitted(*sc, seed, 4);
    else if (config == 6) build_whitted(*sc, seed, 4, true);
    else if (config == 9) build_kinds(*sc, seed);
    else if (config == 10) build_ties(*sc, seed);
    else if (config == 7 || config == 8)
    {
        /* batch and serving scenes: 8 to 512 (config 7) or 16 to 1024
         * (config 8) quadrics, log-uniform in the seed */
        Rng r(seed ^ 0x5bd1e995ull);
        const double lo = config == 7 ? 8.0 : 16.0;
        const double n = lo * std::pow(64.0, r.u(0.0, 1.0));
        if (n < 12.0) build_whitted(*sc, seed, 1, false, 2);
        else if (n < 24.0) build_whitted(*sc, seed, 1);
        else build_whitted(*sc, seed, std::min(64, (int)std::lround(n / 16.0)));
    }
    else build_whitted(*sc, seed, 8 * (config - 2));
    return sc;
}

/* replace every surface's materials by LIGHT materials whose colour encodes
 * (surface id << 1 | side) + 1, so an rgb frame is a first-hit id buffer */
static inline void override_ids(Scene &sc)
{
    for (size_t k = 0; k < sc.surfaces.size(); k++)
    {
        for (int side = 0; side < 2; side++)
        {
            uint32_t v = (uint32_t)(((k << 1) | side) + 1);
            rt_MATERIAL *m = material(sc, RT_MAT_LIGHT, 0xFF000000u | v,
                                      1.0, 0.0, 1.0, 0, 0, 1);
            if (side == 0) sc.surfaces[k]->obj.pmat_outer = m;
            else           sc.surfaces[k]->obj.pmat_inner = m;
        }
    }
}

/* animators (rt_OBJECT::f_anim) of animated workloads: functions of the
 * time alone, so a frame depends on its time and not on the times updated
 * before it */
static rt_void anim_spin(rt_time time, rt_time, rt_TRANSFORM3D *trm, rt_pntr)
{
    trm->rot[RT_Z] = (rt_real)std::fmod((double)time * 0.036, 360.0);
}

static rt_void anim_bob(rt_time time, rt_time, rt_TRANSFORM3D *trm, rt_pntr)
{
    trm->pos[RT_Z] = (rt_real)(1.5 + 0.75 * std::sin((double)time * 0.002));
}

/* the root's child arrays turned about z (even ones) or moved up and down
 * (odd ones) with time */
static inline void animate(Scene &sc)
{
    rt_OBJECT *kids = (rt_OBJECT *)sc.root.root.pobj;
    for (int k = 0, a = 0; k < sc.root.root.obj_num; k++)
    {
        if (kids[k].obj.tag != RT_TAG_ARRAY) continue;
        kids[k].f_anim = (a++ & 1) ? anim_bob : anim_spin;
    }
}

/* every texel array as a BMP file under the data root's data/textures/,
 * named <prefix><k>.bmp, and its material naming that file instead (a
 * texture to load, rt_TEX with x_dim = y_dim = 0); false if a file cannot
 * be written */
static inline bool file_textures(Scene &sc, const std::string &prefix)
{
    for (size_t k = 0; k < sc.tex_mats.size(); k++)
    {
        rt_TEX *tx = &sc.tex_mats[k]->tex;
        const std::string name = prefix + std::to_string(k) + ".bmp";
        if (!qr::bmp_save(qr::data_root() + "data/textures/" + name,
                          (const uint32_t *)tx->ptex, tx->x_dim, tx->y_dim, tx->x_dim))
        {
            return false;
        }
        char *str = sc.alloc<char>(name.size() + 1);
        std::memcpy(str, name.c_str(), name.size() + 1);
        tx->ptex = str;
        tx->x_dim = 0;
        tx->y_dim = 0;
    }
    return true;
}

} /* namespace qrgen */

#endif /* QR_SCENEGEN_HPP */
