/*
 * qr_format.h -- scene description format consumed by the MI355X renderer.
 *
 * A scene is a tree of static C structs (arrays of objects, cameras, lights,
 * surfaces, materials, relations).  The structs, their field order and the
 * initializer macros are kept identical to QuadRay's scene format
 * (core/engine/format.h:40-764) so that scene sources written for QuadRay
 * compile unchanged against this header and can be handed to
 * qrhip_scene_create() / qr::Scene.  Only the data layout is shared; the code
 * that interprets it lives in quadray-engine_amd/csrc.
 *
 * Every surface is described in its local IJK frame, K being the axis of
 * symmetry (the normal for planes).  rt_real is 32-bit float.
 */
#ifndef QR_FORMAT_H
#define QR_FORMAT_H

#include <float.h>
#include <stdint.h>

/* ----------------------------------------------------------------------------
 * base types (QuadRay core/config/rtbase.h:574-689 for the fp32 build)
 * ------------------------------------------------------------------------- */

#ifndef QR_BASE_TYPES
#define QR_BASE_TYPES

typedef void                rt_void;
typedef void               *rt_pntr;
typedef const char         *rt_pstr;
typedef signed char         rt_si08;
typedef unsigned char       rt_ui08;
typedef int                 rt_si32;
typedef unsigned int        rt_ui32;
typedef long long           rt_si64;
typedef unsigned long long  rt_ui64;
typedef rt_si08             rt_bool;
typedef rt_si64             rt_time;
typedef intptr_t            rt_cell;
typedef uintptr_t           rt_word;
typedef float               rt_real;
typedef float               rt_fp32;
typedef rt_real             rt_vec2[2];
typedef rt_real             rt_vec3[3];
typedef rt_real             rt_vec4[4];
typedef rt_real             rt_mat4[4][4];

#define RT_NULL             0
#define RT_TRUE             1
#define RT_FALSE            0
#define RT_INF              (FLT_MAX)
#define RT_ARR_SIZE(a)      ((rt_si32)(sizeof(a) / sizeof(a[0])))

/* axis names */
#define RT_X    0
#define RT_Y    1
#define RT_Z    2
#define RT_W    3
#define RT_I    0
#define RT_J    1
#define RT_K    2
#define RT_L    3
#define RT_U    0
#define RT_V    1
#define RT_R    0
#define RT_G    1
#define RT_B    2
#define RT_A    3

#endif /* QR_BASE_TYPES */

/* ----------------------------------------------------------------------------
 * runtime option bits (QuadRay core/engine/format.h:40-110); a bit set in
 * rt_SCENE.opts turns the corresponding option off for that scene
 * ------------------------------------------------------------------------- */

#define RT_OPTS_THREAD          (1 << 0)
#define RT_OPTS_TILING          (1 << 1)
#define RT_OPTS_TILING_EXT1     (1 << 2)
#define RT_OPTS_FSCALE          (1 << 3)
#define RT_OPTS_TARRAY          (1 << 4)
#define RT_OPTS_VARRAY          (1 << 5)
#define RT_OPTS_ADJUST          (1 << 7)
#define RT_OPTS_UPDATE          (1 << 8)
#define RT_OPTS_RENDER          (1 << 9)
#define RT_OPTS_SHADOW          (1 << 10)
#define RT_OPTS_SHADOW_EXT1     (1 << 11)
#define RT_OPTS_SHADOW_EXT2     (1 << 12)
#define RT_OPTS_2SIDED          (1 << 13)
#define RT_OPTS_2SIDED_EXT1     (1 << 14)
#define RT_OPTS_2SIDED_EXT2     (1 << 15)
#define RT_OPTS_INSERT          (0 << 16)
#define RT_OPTS_INSERT_EXT1     (0 << 17)
#define RT_OPTS_INSERT_EXT2     (0 << 18)
#define RT_OPTS_REMOVE          (0 << 19)
#define RT_OPTS_GAMMA           (1 << 20) /* set: gamma correction off */
#define RT_OPTS_FRESNEL         (1 << 21) /* set: Fresnel reflectance off */
#define RT_OPTS_BUFFERS         (0 << 24)
#define RT_OPTS_PT              (1 << 25) /* set: path tracer prohibited */
#define RT_OPTS_UPDATE_EXT0     (1 << 26)
#define RT_OPTS_UPDATE_EXT1     (1 << 27)
#define RT_OPTS_UPDATE_EXT2     (1 << 28)
#define RT_OPTS_UPDATE_EXT3     (1 << 29)
#define RT_OPTS_RENDER_EXT0     (1 << 30)
#define RT_OPTS_RENDER_EXT1     (1 << 31)

#define RT_OPTS_NONE    (RT_OPTS_GAMMA | RT_OPTS_FRESNEL | RT_OPTS_BUFFERS |    \
                         RT_OPTS_PT)

#define RT_OPTS_FULL    (RT_OPTS_THREAD | RT_OPTS_TILING | RT_OPTS_TILING_EXT1 |\
                         RT_OPTS_FSCALE | RT_OPTS_TARRAY | RT_OPTS_VARRAY |     \
                         RT_OPTS_ADJUST | RT_OPTS_UPDATE | RT_OPTS_RENDER |     \
                         RT_OPTS_SHADOW | RT_OPTS_SHADOW_EXT1 |                 \
                         RT_OPTS_SHADOW_EXT2 | RT_OPTS_2SIDED |                 \
                         RT_OPTS_2SIDED_EXT1 | RT_OPTS_2SIDED_EXT2 |            \
                         RT_OPTS_INSERT | RT_OPTS_INSERT_EXT1 |                 \
                         RT_OPTS_INSERT_EXT2 | RT_OPTS_REMOVE | RT_OPTS_GAMMA | \
                         RT_OPTS_FRESNEL | RT_OPTS_BUFFERS | RT_OPTS_PT)

/* CPU list-pruning heuristics of QuadRay (per-tile lists, bounding-volume
 * skips, per-side and per-light list pruning).  The GPU path always walks
 * complete lists; a scene that sets these bits in rt_SCENE.opts makes
 * QuadRay's CPU renderer walk complete lists too. */
#define RT_OPTS_PRUNING (RT_OPTS_TILING | RT_OPTS_TILING_EXT1 | RT_OPTS_VARRAY | \
                         RT_OPTS_SHADOW | RT_OPTS_SHADOW_EXT1 |                 \
                         RT_OPTS_SHADOW_EXT2 | RT_OPTS_2SIDED |                 \
                         RT_OPTS_2SIDED_EXT1 | RT_OPTS_2SIDED_EXT2)

/* ----------------------------------------------------------------------------
 * object tags (QuadRay core/engine/format.h:120-137)
 * ------------------------------------------------------------------------- */

#define RT_TAG_ARRAY                       -1
#define RT_TAG_PLANE                        0
#define RT_TAG_CYLINDER                     1
#define RT_TAG_SPHERE                       2
#define RT_TAG_CONE                         3
#define RT_TAG_PARABOLOID                   4
#define RT_TAG_HYPERBOLOID                  5
#define RT_TAG_PARACYLINDER                 6
#define RT_TAG_HYPERCYLINDER                7
#define RT_TAG_HYPERPARABOLOID              8
#define RT_TAG_SURFACE_MAX                  9
#define RT_TAG_CAMERA                       100
#define RT_TAG_LIGHT                        101
#define RT_TAG_MAX                          102

#define RT_IS_CAMERA(o)     ((o)->tag == RT_TAG_CAMERA)
#define RT_IS_LIGHT(o)      ((o)->tag == RT_TAG_LIGHT)
#define RT_IS_ARRAY(o)      ((o)->tag == RT_TAG_ARRAY)
#define RT_IS_SURFACE(o)    ((o)->tag  > RT_TAG_ARRAY && (o)->tag < RT_TAG_SURFACE_MAX)
#define RT_IS_PLANE(o)      ((o)->tag == RT_TAG_PLANE)

/* ----------------------------------------------------------------------------
 * relations (custom clippers, boolean accumulators)
 * ------------------------------------------------------------------------- */

#define RT_REL_MINUS_INNER                 -1 /* subtract srf inner subspace */
#define RT_REL_MINUS_OUTER                 +1 /* subtract srf outer subspace */
#define RT_REL_MINUS_ACCUM                  2 /* subtract accum subspace */
#define RT_REL_INDEX_ARRAY                  3 /* next index in sub-array */
#define RT_REL_BOUND_ARRAY                  4
#define RT_REL_UNTIE_ARRAY                  5
#define RT_REL_BOUND_INDEX                  6
#define RT_REL_UNTIE_INDEX                  7

struct rt_RELATION
{
    rt_si32             obj1;
    rt_si32             rel;
    rt_si32             obj2;
};

/* ----------------------------------------------------------------------------
 * transforms, colors, textures, materials
 * ------------------------------------------------------------------------- */

struct rt_TRANSFORM2D
{
    rt_vec2             scl;
    rt_real             rot;
    rt_vec2             pos;
};

struct rt_TRANSFORM3D
{
    rt_vec3             scl;
    rt_vec3             rot;
    rt_vec3             pos;
};

#define RT_COL(val)                 { val, {0.0f} }
#define RT_COL_HDR(r, g, b, a)      { 0x0, {r, g, b, a} }

struct rt_COL
{
    rt_ui32             val;
    rt_vec4             hdr;
};

#define RT_TEX_PCOLOR                       0
#define RT_TEX_HDR_PCOLOR                   10

#define RT_TEX(tag, val)                                                    \
    { RT_TEX_##tag, RT_COL(val), RT_NULL, 0, RT_NULL, 0, 0, 0 }

#define RT_TEX_LOAD(tag, pstr)                                              \
    { RT_TEX_##tag, RT_COL(0x0), (rt_pntr)pstr, 0, RT_NULL, 0, 0, 0 }

#define RT_TEX_BIND(tag, ptex)                                              \
    { RT_TEX_##tag, RT_COL(0x0), *ptex, 0, RT_NULL, 0,                     \
      RT_ARR_SIZE(**ptex), RT_ARR_SIZE(*ptex) }

#define RT_TEX_ARRAY(parr)                                                  \
    { RT_TAG_ARRAY, RT_COL(0x0), *parr, RT_ARR_SIZE(*parr), RT_NULL, 0, 0, 0 }

#define RT_TEX_ARRAY_REL(parr, prel)                                        \
    { RT_TAG_ARRAY, RT_COL(0x0), *parr, RT_ARR_SIZE(*parr),                 \
      *prel, RT_ARR_SIZE(*prel), 0, 0 }

struct rt_TEX
{
    rt_si32             tag;
    rt_COL              col;
    rt_void            *ptex;
    rt_si32             tex_num;
    rt_RELATION        *prel;
    rt_si32             rel_num;
    rt_si32             x_dim;
    rt_si32             y_dim;
};

struct rt_TEXTURE
{
    rt_vec2             scl;
    rt_real             rot;
    rt_vec2             pos;
    rt_TEX              tex;
    rt_real             wgt;
};

#define RT_MAT_PLAIN                        0
#define RT_MAT_LIGHT                        1
#define RT_MAT_METAL                        2
#define RT_MAT(tag)                         RT_MAT_##tag

struct rt_MATERIAL
{
    rt_si32             tag;
    rt_TEX              tex;
    rt_real             lgt[4];     /* diffuse, specular, power */
    rt_real             prp[4];     /* reflect, transmit, refract, extinct */
};

struct rt_SIDE
{
    rt_vec2             scl;
    rt_real             rot;
    rt_vec2             pos;
    rt_MATERIAL        *pmat;
};

/* ----------------------------------------------------------------------------
 * objects
 * ------------------------------------------------------------------------- */

struct rt_OBJ
{
    rt_si32             tag;
    rt_void            *pobj;
    rt_si32             obj_num;
    rt_RELATION        *prel;
    rt_si32             rel_num;
    rt_MATERIAL        *pmat_outer;     /* per-object material override */
    rt_MATERIAL        *pmat_inner;
};

typedef rt_void (*rt_FUNC_ANIM3D)(rt_time time, rt_time last_time,
                                  rt_TRANSFORM3D *trm, rt_pntr pobj);

struct rt_OBJECT
{
    rt_TRANSFORM3D      trm;
    rt_OBJ              obj;
    rt_FUNC_ANIM3D      f_anim;
    rt_time             time;
};

#define RT_CAM_PLAIN                        0
#define RT_CAM(tag)                         RT_CAM_##tag

struct rt_CAMERA
{
    rt_si32             tag;
    rt_COL              col;    /* ambient color */
    rt_real             lum[1]; /* ambient intensity */
    rt_real             vpt[1]; /* distance from the screen (pov) */
    rt_vec3             dps;
    rt_vec3             drt;
};

#define RT_LGT_PLAIN                        0
#define RT_LGT(tag)                         RT_LGT_##tag

struct rt_LIGHT
{
    rt_si32             tag;
    rt_COL              col;
    rt_real             lum[2]; /* ambient, source */
    rt_real             atn[4]; /* range, constant, linear, quadratic */
};

struct rt_SURFACE
{
    rt_vec3             min;
    rt_vec3             max;
    rt_SIDE             side_outer;
    rt_SIDE             side_inner;
};

struct rt_PLANE             { rt_SURFACE srf; };
struct rt_CYLINDER          { rt_SURFACE srf; rt_real rad; };
struct rt_SPHERE            { rt_SURFACE srf; rt_real rad; };
struct rt_CONE              { rt_SURFACE srf; rt_real rat; };
struct rt_PARABOLOID        { rt_SURFACE srf; rt_real par; };
struct rt_HYPERBOLOID       { rt_SURFACE srf; rt_real rat; rt_real hyp; };
struct rt_PARACYLINDER      { rt_SURFACE srf; rt_real par; };
struct rt_HYPERCYLINDER     { rt_SURFACE srf; rt_real rat; rt_real hyp; };
struct rt_HYPERPARABOLOID   { rt_SURFACE srf; rt_real pr1; rt_real pr2; };

struct rt_SCENE
{
    rt_OBJ              root;
    rt_ui32             opts; /* option bits turned off for this scene */
    rt_pntr             lock;
};

/* typed tag helpers: a wrong struct type in a scene initializer is a compile
 * error, as in QuadRay's format */
static inline rt_si32 AR_(rt_OBJECT *)               { return RT_TAG_ARRAY; }
static inline rt_si32 RL_(rt_OBJECT *, rt_RELATION *) { return RT_TAG_ARRAY; }
static inline rt_si32 CM_(rt_CAMERA *)               { return RT_TAG_CAMERA; }
static inline rt_si32 LT_(rt_LIGHT *)                { return RT_TAG_LIGHT; }
static inline rt_si32 PL_(rt_PLANE *)                { return RT_TAG_PLANE; }
static inline rt_si32 CL_(rt_CYLINDER *)             { return RT_TAG_CYLINDER; }
static inline rt_si32 SP_(rt_SPHERE *)               { return RT_TAG_SPHERE; }
static inline rt_si32 CN_(rt_CONE *)                 { return RT_TAG_CONE; }
static inline rt_si32 PB_(rt_PARABOLOID *)           { return RT_TAG_PARABOLOID; }
static inline rt_si32 HB_(rt_HYPERBOLOID *)          { return RT_TAG_HYPERBOLOID; }
static inline rt_si32 PC_(rt_PARACYLINDER *)         { return RT_TAG_PARACYLINDER; }
static inline rt_si32 HC_(rt_HYPERCYLINDER *)        { return RT_TAG_HYPERCYLINDER; }
static inline rt_si32 HP_(rt_HYPERPARABOLOID *)      { return RT_TAG_HYPERPARABOLOID; }

#define QR_OBJ1(fn, pobj)                                                   \
    { fn(pobj), pobj, 1, RT_NULL, 0, RT_NULL, RT_NULL }
#define QR_OBJ1_MAT(fn, pobj, pmo, pmi)                                     \
    { fn(pobj), pobj, 1, RT_NULL, 0, pmo, pmi }

#define RT_OBJ_ARRAY(parr)                                                  \
    { AR_(*parr), *parr, RT_ARR_SIZE(*parr), RT_NULL, 0, RT_NULL, RT_NULL }
#define RT_OBJ_ARRAY_REL(parr, prel)                                        \
    { RL_(*parr, *prel), *parr, RT_ARR_SIZE(*parr),                         \
      *prel, RT_ARR_SIZE(*prel), RT_NULL, RT_NULL }
#define RT_OBJ_CAMERA(pobj)                 QR_OBJ1(CM_, pobj)
#define RT_OBJ_LIGHT(pobj)                  QR_OBJ1(LT_, pobj)
#define RT_OBJ_PLANE(pobj)                  QR_OBJ1(PL_, pobj)
#define RT_OBJ_CYLINDER(pobj)               QR_OBJ1(CL_, pobj)
#define RT_OBJ_SPHERE(pobj)                 QR_OBJ1(SP_, pobj)
#define RT_OBJ_CONE(pobj)                   QR_OBJ1(CN_, pobj)
#define RT_OBJ_PARABOLOID(pobj)             QR_OBJ1(PB_, pobj)
#define RT_OBJ_HYPERBOLOID(pobj)            QR_OBJ1(HB_, pobj)
#define RT_OBJ_PARACYLINDER(pobj)           QR_OBJ1(PC_, pobj)
#define RT_OBJ_HYPERCYLINDER(pobj)          QR_OBJ1(HC_, pobj)
#define RT_OBJ_HYPERPARABOLOID(pobj)        QR_OBJ1(HP_, pobj)
#define RT_OBJ_PLANE_MAT(p, o, i)           QR_OBJ1_MAT(PL_, p, o, i)
#define RT_OBJ_CYLINDER_MAT(p, o, i)        QR_OBJ1_MAT(CL_, p, o, i)
#define RT_OBJ_SPHERE_MAT(p, o, i)          QR_OBJ1_MAT(SP_, p, o, i)
#define RT_OBJ_CONE_MAT(p, o, i)            QR_OBJ1_MAT(CN_, p, o, i)
#define RT_OBJ_PARABOLOID_MAT(p, o, i)      QR_OBJ1_MAT(PB_, p, o, i)
#define RT_OBJ_HYPERBOLOID_MAT(p, o, i)     QR_OBJ1_MAT(HB_, p, o, i)
#define RT_OBJ_PARACYLINDER_MAT(p, o, i)    QR_OBJ1_MAT(PC_, p, o, i)
#define RT_OBJ_HYPERCYLINDER_MAT(p, o, i)   QR_OBJ1_MAT(HC_, p, o, i)
#define RT_OBJ_HYPERPARABOLOID_MAT(p, o, i) QR_OBJ1_MAT(HP_, p, o, i)

#endif /* QR_FORMAT_H */
