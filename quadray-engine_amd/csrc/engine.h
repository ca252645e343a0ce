/*
 * engine.h -- host side scene manager of the MI355X renderer.
 *
 * Builds an object hierarchy from an rt_SCENE tree and runs the per-frame
 * update stage (transform hierarchy, axis mapping, quadric coefficients,
 * clip boxes, custom clip lists, surface/light lists), emitting the
 * structure-of-arrays tables of qr_soa.h plus the per-frame camera
 * description.  Semantics follow QuadRay's scene manager
 * (core/engine/engine.cpp, core/engine/object.cpp); the code is organised
 * around flat vectors that are uploaded to the GPU as they are.
 */
#ifndef QR_ENGINE_H
#define QR_ENGINE_H

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "qr_format.h"
#include "qr_soa.h"
#include "fields.h"
#include "refit.h"

namespace qr {

/* error raised by the host engine; carries the same messages as QuadRay's
 * rt_Exception (core/engine/engine.cpp:2842-2958, object.cpp:72, 869, 4044) */
struct Error : public std::runtime_error
{
    explicit Error(const std::string &m) : std::runtime_error(m) {}
};

/* transform flags (QuadRay core/engine/object.h:74-83) */
enum { FLAG_SCL = 1, FLAG_ROT = 2, FLAG_OBJ = 4 };

struct Object;

/* one material instance per surface side (QuadRay rt_Material) */
struct Material
{
    const rt_MATERIAL *mat = nullptr;
    const rt_SIDE     *sd = nullptr;
    uint32_t props = 0;
    int      map[2] = {0, 1};
    float    scl[2] = {1.0f, 1.0f};
    qr_material rec;          /* record written to the tables */
    std::vector<uint32_t> texels;
    int      x_dim = 1, y_dim = 1;
    const void *tex_src = nullptr;  /* texel array or loaded file the texels
                                       came from (null: a plain colour); one
                                       copy of each goes into the texel pool */
};

/* relation template element: marker (target null) or clipper target */
struct RelElem
{
    int     data;             /* relation sign or accumulator marker */
    Object *target;           /* surface or array, null for a marker */
};

struct Object
{
    enum Kind { ARRAY, SURFACE, CAMERA, LIGHT };

    Kind        kind;
    int         tag;
    Object     *parent = nullptr;
    rt_OBJECT  *obj = nullptr;
    rt_TRANSFORM3D *trm = nullptr;
    rt_TRANSFORM3D  otm;      /* original transform, restored on teardown */

    rt_mat4     mtx;          /* object matrix (relative to trnode if any) */
    rt_mat4     inv;
    const float (*pmtx)[4] = nullptr; /* matrix handed down by the parent */

    int         obj_changed = 0;
    int         obj_has_trm = 0;
    int         mtx_has_trm = 0;
    Object     *trnode = nullptr;

    int         map[4] = {0, 1, 2, 3};
    int         sgn[4] = {1, 1, 1, 1};
    float       scl[4] = {1.0f, 1.0f, 1.0f, 1.0f};

    /* array */
    std::vector<Object *> children;
    rt_mat4     scm;          /* scalers handed to children of a trnode */
    const float (*cmtx)[4] = nullptr; /* matrix handed to children */
    float       col_hdr[4] = {0, 0, 0, 0}; /* light accumulated in array */

    /* surface */
    rt_SURFACE *srf = nullptr;
    Material    outer, inner;
    float       sci[4], scj[4];
    float       bmin[4], bmax[4], cmin[4], cmax[4];
    std::vector<RelElem> rel; /* relations template, as built */
    int         srf_changed = 0;
    int         id = -1;      /* user-visible surface id, construction order */

    /* camera */
    rt_CAMERA  *cam = nullptr;
    float       pov = 1.0f;
    float       hor_sin = 0.0f;   /* of rot[RT_Z] at the last update */
    float       hor_cos = 1.0f;

    /* light */
    rt_LIGHT   *lgt = nullptr;
    qr_light    lrec;

    int         rec = -1;     /* record index in the surface table */

    float *pos() { return mtx[3]; }
};

/* per-surface custom clip list element, in backend order */
struct ClipElem
{
    int     kind;             /* QR_CLIP_* */
    Object *obj;
    int     rel;
    int     last;             /* trnode: position of its last element */
};

/* surface list element */
struct ListElem
{
    Object *obj;              /* surface, or trnode array */
    int     last;             /* trnode: position of its last element */
};

class Engine
{
  public:
    Engine(rt_SCENE *scn, int x_res, int y_res, int x_row);
    ~Engine();

    /* run the update stage for "time" and refresh the tables */
    void update(rt_time time);
    /* camera parameters for the current camera and the scene's frame size,
     * or for x_res x y_res when both are positive; "pose" (12 floats: the
     * right, down and view axes, then the position) replaces the camera's
     * own basis when not null */
    void frame_desc(qr_frame_desc *fd, int fsaa, int depth, int x_res = 0,
                    int y_res = 0, const float *pose = nullptr) const;
    /* the current camera's basis and position in that layout */
    void camera_pose(float *pose) const;

    const qr_scene_tables &tables() const { return tab_; }

    /* The device update stage's inputs.  update_transforms(time) runs the
     * host part of an update alone: animators, hierarchical matrices,
     * cameras and lights, and refreshes every record's inputs (geoms()).
     * It does so only when the structure the tables were last built with
     * still holds (each object's transform node and transform flags, the
     * options); then the records, clip boxes and culling inputs are left
     * to the device (fields.h in update.hip) and the host tables wait for
     * materialize().  Otherwise it runs the whole update on the host and
     * returns false. */
    bool update_transforms(rt_time time);
    /* bring the host tables to the last update_transforms (a no-op after a
     * whole update) */
    void materialize();
    bool host_tables_current() const { return !stale_; }
    const std::vector<qr_geom> &geoms() const { return t_geom_; }
    /* per element of the surface list: the world and local box flags the
     * current programs were built for */
    const std::vector<int32_t> &program_key() const { return topo_key_; }
    const std::vector<qr_light> &light_records() const { return t_lgt_; }

    /* the culling programs' state after the last update: the per-element
     * inputs of their boxes, whether the update refitted the previous
     * programs (same elements, same boxed steps) rather than building new
     * ones, the build those programs come from (a counter), and the
     * reference point of their world boxes' allowance */
    const std::vector<qr_refit_elem> &refit_inputs() const { return t_refit_; }
    bool     refitted() const { return refit_; }
    uint64_t program_build() const { return topo_id_; }
    const float *cull_ref() const { return cull_ref_; }

    int  get_opts() const { return opts_; }
    int  set_opts(int opts);
    int  next_cam();
    /* move or turn the current camera as "action" (QRHIP_CAMERA_*) for the
     * time since its last update; takes effect at the next update */
    void update_action(rt_time time, int action);
    int  get_cam_idx() const { return cam_idx_; }
    int  num_surfaces() const { return (int)surfaces_.size(); }
    int  num_lights() const { return (int)lights_.size(); }

    int  x_res() const { return x_res_; }
    int  y_res() const { return y_res_; }
    int  x_row() const { return x_row_; }

  private:
    Object *make_array(Object *parent, rt_OBJECT *obj);
    Object *make_object(Object *parent, rt_OBJECT *obj);
    void    init_material(Material &m, const rt_SIDE *sd, const rt_MATERIAL *mat);
    void    build_relations(Object *arr, rt_OBJECT *obj);
    void    add_relation(Object *o, const std::vector<RelElem> &lst);

    void    update_object(Object *o, rt_time time, int flags, Object *trnode,
                          const float (*mtx)[4]);
    void    update_status(Object *o, rt_time time, int flags, Object *trnode);
    void    update_matrix(Object *o, const float (*mtx)[4]);
    void    update_fields(Object *o);
    void    build_clip_list(Object *s, std::vector<ClipElem> &out);
    void    make_geom(Object *o, qr_geom &g) const;
    void    geom_transform(Object *o, qr_geom &g) const;
    void    finish_update();
    void    build_surface_list();
    void    emit_tables(const std::vector<std::vector<ClipElem>> &clips);
    void    elem_inputs(std::vector<int32_t> &key);
    void    build_cull_program();
    void    refit_cull_program();

    rt_SCENE  *scn_;
    rt_OBJECT  rootobj_;
    int        opts_;
    int        x_res_, y_res_, x_row_;
    float      factor_, aspect_;

    std::vector<Object *> objects_;   /* ownership */
    Object    *root_ = nullptr;
    std::vector<Object *> arrays_;    /* construction order */
    std::vector<Object *> surfaces_;  /* construction order */
    std::vector<Object *> cameras_;   /* construction order */
    std::vector<Object *> lights_;    /* construction order */
    Object    *cam_ = nullptr;
    int        cam_idx_ = 0;

    std::vector<ListElem> slist_;

    /* tables */
    std::vector<qr_geom>     t_geom_;     /* per record: its inputs (fields.h) */
    struct Box4 { float v[4]; };
    std::vector<Box4>        t_box_;      /* per record: bmin, bmax, cmin, cmax */
    std::vector<qr_surface>  t_surf_;
    std::vector<qr_material> t_mat_;
    std::vector<qr_light>    t_lgt_;
    std::vector<qr_clip>     t_clip_;
    std::vector<qr_elem>     t_elem_;
    std::vector<uint32_t>    t_texel_;
    struct Texture
    {
        std::vector<uint32_t> texels;
        int x_dim = 0, y_dim = 0;
    };
    std::map<std::string, Texture> textures_;   /* texture files by name */

    std::vector<qr_cull_step> t_step_;      /* nearest-hit program (list order) */
    std::vector<qr_cull_step> t_step_any_;  /* any-hit program (spatial order) */
    std::vector<qr_refit_elem> t_refit_;    /* per-element box inputs */
    std::vector<int32_t>     topo_key_;     /* what the programs were built for */
    uint64_t                 topo_id_ = 0;
    bool                     refit_ = false;
    float                    cull_ref_[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    qr_scene_tables          tab_;

    /* the structure the tables were built with: per record object its
     * transform node and transform flags, and the options */
    struct Sig { Object *o; Object *trnode; int trm; };
    std::vector<Sig>         sig_;
    int                      sig_opts_ = -1;
    bool                     stale_ = false;    /* host tables behind */
};

/* 4x4 helpers with QuadRay's operation order (core/engine/rtgeom.cpp:60-200) */
void mat_mul_mat(rt_mat4 mp, const float (*m1)[4], const float (*m2)[4]);
void mat_from_transform(rt_mat4 mp, const rt_TRANSFORM3D *t, bool with_scale);
void mat_inverse(rt_mat4 mp, const float (*m1)[4]);
extern const rt_mat4 iden4;

} /* namespace qr */

#endif /* QR_ENGINE_H */
