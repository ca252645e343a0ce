/*
 * qr_soa.h -- structure-of-arrays scene tables: the data that crosses the
 * host/device boundary of the MI355X renderer.
 *
 * The host engine (quadray-engine_amd/csrc/engine.cpp) walks an rt_SCENE
 * tree, runs the update stage (transforms, axis maps, quadric coefficients,
 * clip boxes, clip and surface lists) and fills one qr_scene_tables.  The HIP
 * kernels read the tables from HBM and stage them in LDS; the CPU oracle reads
 * the same tables.  Plain C, fixed-width fields, no pointers inside records:
 * every cross reference is an index, so the tables can be copied to device
 * memory verbatim.
 *
 * Field meanings follow QuadRay's backend structures rt_SIMD_SURFACE,
 * rt_SIMD_MATERIAL, rt_SIMD_LIGHT and rt_SIMD_CAMERA
 * (core/tracer/tracer.h:869-1292), stored once per object instead of being
 * replicated across SIMD lanes.
 */
#ifndef QR_SOA_H
#define QR_SOA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* material property bits (QuadRay core/tracer/tracer.h:66-77) */
#define QR_PROP_LIGHT       0x00000010u
#define QR_PROP_METAL       0x00000020u
#define QR_PROP_GAMMA       0x00000040u
#define QR_PROP_FRESNEL     0x00000080u
#define QR_PROP_NORMAL      0x00000100u
#define QR_PROP_OPAQUE      0x00000200u
#define QR_PROP_TRANSP      0x00000400u
#define QR_PROP_TEXTURE     0x00000800u
#define QR_PROP_REFLECT     0x00001000u
#define QR_PROP_REFRACT     0x00002000u
#define QR_PROP_DIFFUSE     0x00004000u
#define QR_PROP_SPECULAR    0x00008000u

/* solver kinds (QuadRay update0, core/tracer/tracer.cpp:8115-8150) */
#define QR_SOLVE_NONE       0   /* transform node: only feeds the cache */
#define QR_SOLVE_PLANE      1
#define QR_SOLVE_QUADRIC    2
#define QR_SOLVE_TWOPLANE   3

/* clip-list element kinds */
#define QR_CLIP_SURFACE     0   /* clipper surface, rel = -1 inner / +1 outer */
#define QR_CLIP_TRNODE      1   /* transform node, last = index of its last element */
#define QR_CLIP_ENTER       2   /* accumulator segment start */
#define QR_CLIP_LEAVE       3   /* accumulator segment end */

/*
 * One record per list target: every surface and every array that acts as a
 * transform node ("trnode": an array with a non-trivial rotation whose
 * children have trivial transforms relative to it).
 */
typedef struct qr_surface
{
    float    pos[3];        /* position (trnode-local for cached surfaces) */
    float    cmin[3];       /* axis clip values, axis-mapped local frame */
    float    cmax[3];
    float    tc[3][3];      /* inverse transform rows: tc[i][x] = tci_x ... */
    float    sci[4];        /* quadric scalers x, y, z, w */
    float    scj[3];        /* quadric linear terms (already halved) */
    float    d_eps;         /* near-zero discriminant threshold */
    float    t_eps;         /* root split threshold */
    int32_t  amap[3];       /* local I, J, K -> world-aligned axis 0..2 */
    int32_t  asgn[3];       /* 1 if that local axis is negated */
    int32_t  shift;         /* 1: local coordinates come from the transformed
                               (IJK) fields, 0: from world-aligned XYZ */
    int32_t  trm;           /* transform flags: 0 none, 1 diagonal scale,
                               2 rotation, 3 rotation + scale */
    int32_t  tag;           /* RT_TAG_* or -1 for a trnode array */
    int32_t  solver;        /* QR_SOLVE_* */
    int32_t  nrm_kind;      /* 1 plane, 2 quadric with linear terms, 3 without */
    int32_t  clp_kind;      /* same encoding, used when acting as a clipper */
    int32_t  apex;          /* conic apex snap: 0 off, 1 cone-like, 2 two-plane */
    int32_t  trnode;        /* index of the transform node record, or -1 */
    int32_t  mat[2];        /* material index, outer / inner side */
    uint32_t props[2];      /* material property bits, outer / inner side */
    int32_t  clip_beg;      /* custom clip list: first element in clip table */
    int32_t  clip_cnt;      /* number of elements */
    int32_t  cmin_on;       /* bit a set: axis a has a min clip */
    int32_t  cmax_on;       /* bit a set: axis a has a max clip */
    int32_t  id;            /* user-visible surface id (scene order), -1 trnode */
    int32_t  pad[3];
} qr_surface;

typedef struct qr_clip
{
    int32_t  kind;          /* QR_CLIP_* */
    int32_t  surf;          /* index into surfaces (SURFACE, TRNODE) */
    int32_t  rel;           /* relation sign for SURFACE, last element for TRNODE */
    int32_t  pad;
} qr_clip;

typedef struct qr_material
{
    float    xscal, yscal;  /* texture transform */
    float    xoffs, yoffs;
    uint32_t xmask, ymask;  /* texel index masks (power-of-two textures) */
    uint32_t yshft;         /* log2 of texture width */
    int32_t  tex_off;       /* first texel in the texel pool */
    int32_t  tmap[2];       /* texture x/y take local u (0) or v (1) */
    float    l_dff;         /* diffuse weight */
    float    l_spc;         /* specular weight */
    uint32_t l_pow;         /* specular power, fixed point 28.4 */
    float    c_rfl;         /* reflectivity */
    float    c_trn;         /* transparency */
    float    c_rfr;         /* relative refractive index */
    float    rfr_2;         /* c_rfr squared */
    float    c_rcp;         /* 1 / prp[2] */
    float    ext_2;         /* metal extinction squared */
    float    col[3];        /* path-tracer emission (LIGHT materials) */
    float    e_src;
    int32_t  pad[2];
} qr_material;

typedef struct qr_light
{
    float    pos[3];
    float    col[3];        /* color * source intensity */
    float    l_src;
    float    a_qdr, a_lnr, a_cnt, a_rng;
    float    t_max;
} qr_light;

/* list element of the surface list walked by every ray */
typedef struct qr_elem
{
    int32_t  surf;          /* index into surfaces */
    int32_t  last;          /* trnode element: index of the last element of its
                               group in this list; otherwise -1 */
} qr_elem;

/* culling programs over the surface list.  The nearest-hit program visits
 * list elements in list order; the any-hit program may visit the top-level
 * units (a transform-node group or a lone surface) in any order, since a
 * shadow test's result does not depend on it.  A box step lets a wave jump to
 * "skip" when no ray of the wave can meet the box, so skipping changes no
 * result.  A box bounds every hit
 * the tracer can accept in its range (with a margin for rounding of hit
 * points); a ray tests it grown by QR_CULL_FUZZ times the L1 distance from the
 * ray's origin to the box centre, which covers the false roots that fp32
 * cancellation in b*b - a*c lets far-away rays accept near a surface.  Built
 * by the update stage; no counterpart in QuadRay, whose tile lists
 * (core/engine/engine.cpp:1985-2160) play this role on the CPU. */
#define QR_CULL_FUZZ 0.003f

typedef struct qr_cull_step
{
    float    c[3], h[3];    /* box centre and half extents: world box for a
                               box step; for an element step with skip -2
                               the surface's own box in its tracing frame,
                               tested after its transform */
    int32_t  elem;          /* list element to walk, -1 for a box step */
    int32_t  skip;          /* box step: first step after its range;
                               element step: -1, or -2 with a local box */
} qr_cull_step;

typedef struct qr_scene_tables
{
    int32_t      n_surf;    /* records in surfaces (surfaces + trnodes) */
    int32_t      n_mat;
    int32_t      n_lgt;
    int32_t      n_clip;
    int32_t      n_elem;    /* elements in the surface list */
    int32_t      n_texel;
    int32_t      n_user;    /* user-visible surfaces (id 0 .. n_user-1) */
    int32_t      pad;
    qr_surface  *surfaces;
    qr_material *materials;
    qr_light    *lights;
    qr_clip     *clips;
    qr_elem     *elems;
    uint32_t    *texels;    /* 0xAARRGGBB, A ignored */
    int32_t      n_step;    /* culling program of nearest-hit walks, units in
                               list order; 0: walk the whole list */
    int32_t      pad2;
    qr_cull_step *steps;
    int32_t      n_step_any;    /* culling program of any-hit (shadow) walks,
                                   units in any order, each group's elements
                                   in list order; 0: use "steps" */
    int32_t      pad3;
    qr_cull_step *steps_any;
} qr_scene_tables;

/* per-frame camera and frame parameters (QuadRay rt_Scene::render and
 * render_slice, core/engine/engine.cpp:3078-3098, 3302-3318, 3514-3697) */
typedef struct qr_frame_desc
{
    float    org[3];        /* camera position */
    float    dir[3];        /* ray to the top-left pixel centre */
    float    hor[3];        /* one pixel step to the right */
    float    ver[3];        /* one pixel step down */
    float    amb[4];        /* accumulated ambient color, intensity */
    float    t_min;         /* camera pov: near clip of primary rays */
    int32_t  x_res, y_res;  /* pixels */
    int32_t  x_row;         /* frame stride in pixels (may be negative) */
    int32_t  fsaa;          /* 0 none, 1 2X, 2 4X */
    int32_t  depth;         /* secondary ray depth */
    uint32_t gamma;         /* QR_PROP_GAMMA when gamma correction is on */
    int32_t  pt_on;         /* path tracer samples per render call, 0 off */
    float    pts_c;         /* path tracer samples accumulated so far */
    int32_t  pt_split;      /* path tracer: Fresnel split when depth <= this */
    int32_t  pt_roul;       /* path tracer: Russian roulette when depth <= this */
    int32_t  pad;
    float    aa_h[4];       /* antialiasing sub-sample offsets in pixels */
    float    aa_v[4];
} qr_frame_desc;

#ifdef __cplusplus
}
#endif

#endif /* QR_SOA_H */
