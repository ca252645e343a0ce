/*
 * qr_oracle.c -- scalar CPU restatement of QuadRay's Whitted ray tracer.
 *
 * TEST INFRASTRUCTURE: this file is the checker for the MI355X kernels.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * One ray at a time, fp32, operations in QuadRay's order (no contraction:
 * build with -ffp-contract=off).  QuadRay processes SIMD packets and shades
 * every candidate hit that passes the depth test, overwriting earlier ones;
 * per ray that equals shading the nearest candidate (strict depth test,
 * first in list order wins ties), which is what this restatement does.
 * Section references are to core/tracer/tracer.cpp.
 */
#include "qr_oracle.h"

#include <float.h>
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define MAX_DEPTH_STACK 32

typedef struct ray_ctx
{
    float org[3];       /* ORG */
    float ray[3];       /* RAY_X/Y/Z */
    float t_min;        /* T_MIN */
    float t_max;        /* initial T_BUF */
    int   flags;        /* PARAM(FLG): side | pass | shadow (low bits) */
    int   origin;       /* PARAM(OBJ): originating surface record, -1 none */
    float oloc[3];      /* originating surface's local hit (previous NRM_I/J/K) */
    int   depth;        /* inf_DEPTH */
    int   shadow;       /* shadow ray: any shadowing hit ends the walk */
    int   primary;      /* record near-zero discriminant flags */
} ray_ctx;

typedef struct hit_rec
{
    int   srf;          /* surface record, -1 none */
    int   side;         /* 0 outer, 1 inner */
    float t;
    float hit[3];       /* HIT_X/Y/Z */
    float loc[3];       /* local hit in the surface's frame (NEW + shift) */
    float tex_uv[2];    /* TEX_U, TEX_V (planes) */
} hit_rec;

typedef struct oracle
{
    const qr_scene_tables *tab;
    const qr_frame_desc   *fd;
    uint64_t ops;
    uint64_t rays;
    int      near_zero;  /* primary ray had a near-zero discriminant */
    uint64_t n_dmask, n_twoplane, n_apex;   /* branch coverage counters */
    int      hit_id;     /* first hit of the primary ray, (id << 1 | side) */
    const float *boxes;  /* culling-box audit (not part of QuadRay's walk) */
    uint64_t box_checked, box_outside;
    float    box_excess;  /* largest (distance outside the box) / (L1 origin
                             distance to its centre) of an accepted hit */
    /* op accounting under per-ray culling (not part of QuadRay's walk) */
    int      *anc_off;   /* per element: [anc_off[e], anc_off[e+1]) in anc */
    int      *anc;       /* box steps starting at each element */
    int      *box_last;  /* last element enclosed by each box step */
    int      *elem_step; /* step of each list element */
    /* any-hit program (shadow rays): every box step enclosing each element,
     * outermost first, and each box's result for the current ray */
    int      *any_off;
    int      *any_anc;
    signed char *any_res;
    uint64_t  ops_skip, ops_box;
} oracle;

/* box steps of the culling program grouped by the first list element each
 * encloses (outermost first), with the last element each encloses */
static int cull_setup(oracle *o)
{
    const qr_scene_tables *tab = o->tab;
    if (tab->n_step <= 0 || tab->steps == NULL) return 0;
    int n = tab->n_elem, ns = tab->n_step;
    int *first = malloc(sizeof(int) * (size_t)ns);
    o->anc_off = calloc((size_t)n + 1, sizeof(int));
    o->anc = malloc(sizeof(int) * ((size_t)ns + 1));
    o->box_last = malloc(sizeof(int) * ((size_t)ns + 1));
    o->elem_step = malloc(sizeof(int) * ((size_t)n + 1));
    if (!first || !o->anc_off || !o->anc || !o->box_last || !o->elem_step)
    {
        free(first);
        return -1;
    }
    for (int p = 0; p < ns; p++)
    {
        if (tab->steps[p].elem >= 0 && tab->steps[p].elem < n)
            o->elem_step[tab->steps[p].elem] = p;
    }
    for (int p = 0; p < ns; p++)
    {
        const qr_cull_step *st = &tab->steps[p];
        first[p] = -1;
        o->box_last[p] = -1;
        if (st->elem >= 0) continue;
        for (int q = p + 1; q < st->skip && q < ns; q++)
        {
            int e = tab->steps[q].elem;
            if (e < 0) continue;
            if (first[p] < 0) first[p] = e;
            o->box_last[p] = e;
        }
        if (first[p] >= 0) o->anc_off[first[p] + 1]++;
    }
    for (int e = 0; e < n; e++) o->anc_off[e + 1] += o->anc_off[e];
    int *fill = calloc((size_t)n + 1, sizeof(int));
    if (!fill) { free(first); return -1; }
    for (int p = 0; p < ns; p++)
    {
        if (first[p] >= 0) o->anc[o->anc_off[first[p]] + fill[first[p]]++] = p;
    }
    free(fill);
    free(first);
    return 0;
}

/* box steps of the any-hit program enclosing each element */
static int cull_setup_any(oracle *o)
{
    const qr_scene_tables *tab = o->tab;
    if (tab->n_step_any <= 0 || tab->steps_any == NULL) return 0;
    const int n = tab->n_elem, ns = tab->n_step_any;
    o->any_off = calloc((size_t)n + 1, sizeof(int));
    o->any_res = malloc((size_t)ns);
    if (!o->any_off || !o->any_res) return -1;
    for (int pass = 0; pass < 2; pass++)
    {
        int *fill = pass == 1 ? calloc((size_t)n, sizeof(int)) : NULL;
        if (pass == 1 && !fill) return -1;
        for (int p = 0; p < ns; p++)
        {
            const qr_cull_step *st = &tab->steps_any[p];
            if (st->elem >= 0) continue;
            for (int q = p + 1; q < st->skip && q < ns; q++)
            {
                const int e = tab->steps_any[q].elem;
                if (e < 0 || e >= n) continue;
                if (pass == 0) o->any_off[e + 1]++;
                else o->any_anc[o->any_off[e] + fill[e]++] = p;
            }
        }
        if (pass == 0)
        {
            for (int e = 0; e < n; e++) o->any_off[e + 1] += o->any_off[e];
            o->any_anc = malloc(sizeof(int) * ((size_t)o->any_off[n] + 1));
            if (!o->any_anc) return -1;
        }
        free(fill);
    }
    return 0;
}

static void cull_free(oracle *o)
{
    free(o->anc_off); free(o->anc); free(o->box_last); free(o->elem_step);
    free(o->any_off); free(o->any_anc); free(o->any_res);
}

/* reciprocal for the slab tests, clamped to +-1e30 so that a zero component
 * gives finite slab distances */
static float slab_rcp(float d)
{
    return fminf(fmaxf(1.0f / d, -1e30f), 1e30f);
}

/* the renderer's box test (accounting only): segment (t_min, t_buf) of the
 * ray from org against the box grown by QR_CULL_FUZZ times the L1 distance
 * from org to its centre; the entry must come no later than the exit */
static int box_meets(const qr_cull_step *st, const float *org, const float *inv,
                     float t_min, float t_buf)
{
    float d[3], l1 = 0.0f;
    for (int a = 0; a < 3; a++)
    {
        d[a] = st->c[a] - org[a];
        l1 += fabsf(d[a]);
    }
    const float m = QR_CULL_FUZZ * l1;
    float tn = -INFINITY, tf = INFINITY;
    for (int a = 0; a < 3; a++)
    {
        float h = st->h[a] + m;
        float t0 = (d[a] - h) * inv[a];
        float t1 = (d[a] + h) * inv[a];
        tn = fmaxf(tn, fminf(t0, t1));
        tf = fminf(tf, fmaxf(t0, t1));
    }
    return !(tn > t_buf) && !(tf < t_min) && !(tn > tf);
}

/* would this ray's walk of the culling program reach element e: evaluate
 * the boxes that start at e (outermost first); a miss skips to the end of
 * that box's range */
static int cull_reach(oracle *o, const float *org, const float *inv,
                      float t_min, float t_buf, int e, int *skip_until)
{
    if (o->anc_off == NULL) return 1;
    if (e <= *skip_until) return 0;
    for (int k = o->anc_off[e]; k < o->anc_off[e + 1]; k++)
    {
        const qr_cull_step *st = &o->tab->steps[o->anc[k]];
        o->ops_box += 28;
        if (!box_meets(st, org, inv, t_min, t_buf))
        {
            *skip_until = o->box_last[o->anc[k]];
            return 0;
        }
    }
    return 1;
}

/* would a shadow ray's walk of the any-hit program reach element e: every
 * box enclosing it must meet the ray; each box is evaluated once per ray */
static int cull_reach_any(oracle *o, const float *org, const float *inv,
                          float t_min, float t_buf, int e)
{
    for (int k = o->any_off[e]; k < o->any_off[e + 1]; k++)
    {
        const int p = o->any_anc[k];
        if (o->any_res[p] < 0)
        {
            o->ops_box += 28;
            o->any_res[p] = (signed char)box_meets(&o->tab->steps_any[p], org, inv,
                                                   t_min, t_buf);
        }
        if (!o->any_res[p]) return 0;
    }
    return 1;
}

/* culling-box audit: is an accepted hit point inside its surface's box */
static void audit_hit(oracle *o, int si, const float *hit, const float *loc,
                      const float *org, const float *lorg)
{
    if (o->boxes == NULL) return;
    const float *b = o->boxes + 7 * (size_t)si;
    if (b[0] != b[0]) return;   /* no box */
    const int local = b[6] != 0.0f;
    const float *p = local ? loc : hit;      /* world or tracing frame */
    const float *q = local ? lorg : org;     /* the ray origin the test uses */
    float l1 = 0.0f;
    for (int a = 0; a < 3; a++) l1 += fabsf(b[a] - q[a]);
    const float m = QR_CULL_FUZZ * l1;
    o->box_checked++;
    int out = 0;
    for (int a = 0; a < 3; a++)
    {
        const float ex = fabsf(p[a] - b[a]) - b[3 + a];
        if (ex > 0.0f && l1 > 0.0f && ex / l1 > o->box_excess) o->box_excess = ex / l1;
        if (!(ex <= m)) out = 1;
    }
    o->box_outside += out;
}

#define FLAG_SIDE 1
#define FLAG_PASS 2
#define FLAG_SHAD 4

static inline float fneg_if(float v, int neg) { return neg ? -v : v; }
static inline float rsq(float x) { return 1.0f / sqrtf(x); }

static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline float fxor(float v, uint32_t m) { return bitsf(fbits(v) ^ m); }
static inline float fand(float v, uint32_t m) { return bitsf(fbits(v) & m); }

/* x86 float->int conversion: out-of-range and NaN give 0x80000000 */
static inline int32_t cvt_floor(float x)
{
    float f = floorf(x);
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) return (int32_t)0x80000000u;
    return (int32_t)f;
}
static inline int32_t cvt_rne(float x)
{
    float f = rintf(x);
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) return (int32_t)0x80000000u;
    return (int32_t)f;
}

/* ------------------------------------------------------------------------ */
/* list walk state: trnode transform cache (OO_cyc, 2430-2684)              */
/* ------------------------------------------------------------------------ */

typedef struct walk_state
{
    float dff[6];       /* DFF_X/Y/Z, DFF_I/J/K */
    float ray6[6];      /* RAY_X/Y/Z, RAY_I/J/K */
    float nrm6[6];      /* NRM fields used as clipper temporaries */
    float new6[6];      /* NEW_X/Y/Z, NEW_I/J/K: local hit temporaries */
    float hit[3];
    int   cache_last;   /* LOCAL(OBJ): list index of trnode's last element */
} walk_state;

/* diff or ray times the inverse matrix rows, diagonal first
 * (2536-2568, 2602-2643); diag_only for the scale fast path */
static void transform_in(oracle *o, const qr_surface *s, const float *v, float *out)
{
    float x = s->tc[0][0] * v[0];
    float y = s->tc[1][1] * v[1];
    float z = s->tc[2][2] * v[2];
    o->ops += 3;
    if (s->trm != 1)
    {
        x = x + s->tc[0][1] * v[1];
        x = x + s->tc[0][2] * v[2];
        y = y + s->tc[1][0] * v[0];
        y = y + s->tc[1][2] * v[2];
        z = z + s->tc[2][0] * v[0];
        z = z + s->tc[2][1] * v[1];
        o->ops += 12;
    }
    out[0] = x; out[1] = y; out[2] = z;
}

/* ------------------------------------------------------------------------ */
/* clipping subroutine CC_clp (2686-3236)                                   */
/* ------------------------------------------------------------------------ */

static int clip_eval(oracle *o, walk_state *w, const qr_surface *c, int rel)
{
    const int sh = c->shift * 3;
    float f;
    if (c->clp_kind == 1)
    {
        int k = c->amap[2] + sh;
        f = fneg_if(w->nrm6[k], c->asgn[2]);
        return rel < 0 ? (f >= 0.0f) : (f <= 0.0f);
    }
    const float *p = w->nrm6 + sh;
    if (c->clp_kind == 2)
    {
        float d[3];
        for (int a = 0; a < 3; a++)
        {
            float df1 = (c->scj[a] + c->scj[a]) * p[a];
            float df2 = (p[a] * p[a]) * c->sci[a];
            d[a] = df2 - df1;
        }
        f = d[0] - c->sci[3];
        f = f + d[1];
        f = f + d[2];
        o->ops += 18;
    }
    else
    {
        float d0 = (p[0] * p[0]) * c->sci[0];
        float d1 = (p[1] * p[1]) * c->sci[1];
        float d2 = (p[2] * p[2]) * c->sci[2];
        f = d0 - c->sci[3];
        f = f + d1;
        f = f + d2;
        o->ops += 9;
    }
    return rel < 0 ? (f >= 0.0f) : (f <= 0.0f);
}

/* returns 1 if the candidate root t on surface index si survives depth test,
 * axis clipping and custom clippers; fills w->hit and w->new6 */
static int clip_root(oracle *o, walk_state *w, const ray_ctx *rc, float t_buf,
                     int tie, int si, float t, int side, int dmask, float amask_sign)
{
    const qr_scene_tables *tab = o->tab;
    const qr_surface *s = &tab->surfaces[si];
    const int sh = s->shift * 3;
    int ok = 1;

    /* depth test (CC_clp 2690-2694) is strict, but the nearest-hit walks of
     * camera rays and of the rays leaving reflecting or transparent surfaces
     * run over lists that rt_SceneThread::ssort builds for them
     * (core/engine/engine.cpp:2180-2420) by prepending in the order of the
     * global list: their order is the global list's reversed, so of two
     * roots at the same distance the later one of the global list wins.
     * Walked in the global list's order, a root equal to the kept one (tie)
     * is taken. */
    if (!(t_buf > t) && !(tie && t_buf == t)) ok = 0;
    if (!(rc->t_min < t)) ok = 0;
    if (!ok) return 0;

    for (int a = 0; a < 3; a++)
    {
        w->hit[a] = rc->ray[a] * t + rc->org[a];
    }
    o->ops += 6;

    if (s->trm != 0)
    {
        for (int a = 0; a < 3; a++)
        {
            w->new6[3 + a] = w->ray6[3 + a] * t + w->dff[3 + a];
        }
    }
    else
    {
        for (int a = 0; a < 3; a++)
        {
            w->new6[a] = w->hit[a] - s->pos[a];
        }
    }
    o->ops += 6;

    /* conic apex snap for near-zero discriminant (2795-2945) */
    if (s->apex != 0 && dmask)
    {
        int mi = s->amap[0] + sh, mj = s->amap[1] + sh, mk = s->amap[2] + sh;
        float li = w->new6[mi];
        float r = li * li;
        if (s->apex != 2)
        {
            float lj = w->new6[mj];
            r = r + lj * lj;
        }
        float lk = w->new6[mk];
        r = r + lk * lk;
        o->ops += 5;
        if (r < s->t_eps)
        {
            o->n_apex++;
            float ci = bitsf((fbits(w->dff[mi]) & 0x80000000u) ^ fbits(1.0f));
            float cj = 0.0f;
            float sk = s->sci[s->amap[0]];
            float lr = 1.0f;
            if (s->apex != 2)
            {
                cj = bitsf((fbits(w->dff[mj]) & 0x80000000u) ^ fbits(1.0f));
                sk = sk + s->sci[s->amap[1]];
                lr = lr + 1.0f;
            }
            sk = sk / s->sci[s->amap[2]];
            sk = fxor(sk, 0x80000000u);
            float lc2 = sk;
            float ck = sqrtf(sk);
            lc2 = lc2 + lr;
            float inv = rsq(lc2);
            inv = inv * s->t_eps;
            ci = ci * inv;
            cj = cj * inv;
            ck = ck * inv;
            o->ops += 12;

            uint32_t am = fbits(amask_sign) & 0x80000000u;
            uint32_t ts = side ? 0x80000000u : 0u;
            ck = fxor(ck, fbits(w->dff[mk]) & 0x80000000u);
            ck = fxor(ck, (ts & am) ^ am);
            uint32_t tsg = (ts | am) ^ am;
            ci = fxor(ci, tsg);
            cj = fxor(cj, tsg);

            w->new6[mi] = ci;
            if (s->apex != 2) w->new6[mj] = cj;
            w->new6[mk] = ck;
        }
    }

    const float *l = w->new6 + sh;

    /* axis clipping (2963-3014) */
    for (int a = 0; a < 3; a++)
    {
        if ((s->cmin_on >> a) & 1)
        {
            if (!(s->cmin[a] <= l[a])) ok = 0;
        }
        if ((s->cmax_on >> a) & 1)
        {
            if (!(s->cmax[a] >= l[a])) ok = 0;
        }
    }

    /* custom clippers with accumulator segments (3020-3240) */
    if (s->clip_cnt > 0)
    {
        int acc_saved = 0;
        int cache = -1;
        for (int k = 0; k < s->clip_cnt; k++)
        {
            const qr_clip *e = &tab->clips[s->clip_beg + k];
            if (e->kind == QR_CLIP_ENTER)
            {
                acc_saved = ok;
                ok = 1;
                continue;
            }
            if (e->kind == QR_CLIP_LEAVE)
            {
                ok = (!ok) && acc_saved;
                continue;
            }
            const qr_surface *c = &tab->surfaces[e->surf];
            if (e->kind == QR_CLIP_TRNODE)
            {
                if (e->surf == s->trnode)
                {
                    for (int a = 0; a < 3; a++)
                    {
                        w->nrm6[a] = w->new6[3 + a] + s->pos[a];
                    }
                    o->ops += 3;
                }
                else
                {
                    float d[3];
                    for (int a = 0; a < 3; a++)
                    {
                        d[a] = w->hit[a] - c->pos[a];
                    }
                    o->ops += 3;
                    transform_in(o, c, d, w->nrm6);
                }
                cache = e->rel;
                continue;
            }
            if (cache >= 0)
            {
                for (int a = 0; a < 3; a++)
                {
                    w->nrm6[3 + a] = w->nrm6[a] - c->pos[a];
                }
                o->ops += 3;
                if (k == cache)
                {
                    cache = -1;
                }
            }
            else
            {
                for (int a = 0; a < 3; a++)
                {
                    w->nrm6[a] = w->hit[a] - c->pos[a];
                }
                o->ops += 3;
                if (c->trm != 0)
                {
                    transform_in(o, c, w->nrm6, w->nrm6 + 3);
                }
            }
            ok = ok && clip_eval(o, w, c, e->rel);
        }
    }
    return ok;
}

/* ------------------------------------------------------------------------ */
/* solvers (PL 5526-5712, QD 5714-6338, TP 6341-6497)                       */
/* ------------------------------------------------------------------------ */

/* material property bits of the side that was hit */
static inline uint32_t side_props(const qr_surface *s, int side)
{
    return s->props[side];
}

/* shadow check at a surviving root (CHECK_SHAD, 576-616): 1 if it shadows */
static int shadows(const qr_surface *s, int side)
{
    uint32_t p = side_props(s, side);
    if (p & QR_PROP_LIGHT) return 0;
    if ((p & QR_PROP_TRANSP) && !(p & QR_PROP_REFRACT)) return 0;
    return 1;
}

/* walk the surface list; nearest hit (shadow == 0) or any shadowing hit */
/* whether QuadRay gives a surface a nearest-hit list of its own for the
 * rays leaving it (reflecting or not opaque on either side,
 * core/engine/engine.cpp:2180-2201), built in the reverse of the global
 * list's order; other surfaces' rays walk the global list itself */
static int own_list(const qr_surface *s)
{
    const uint32_t p0 = s->props[0], p1 = s->props[1];
    return ((p0 | p1) & QR_PROP_REFLECT) != 0 || (p0 & p1 & QR_PROP_OPAQUE) == 0;
}

static int walk(oracle *o, const ray_ctx *rc, hit_rec *best)
{
    const qr_scene_tables *tab = o->tab;
    walk_state w;
    memset(&w, 0, sizeof(w));
    w.cache_last = -1;
    for (int a = 0; a < 3; a++) w.ray6[a] = rc->ray[a];

    float t_buf = rc->t_max;
    best->srf = -1;
    /* camera rays walk the camera's list (rt_Scene::render, engine.cpp:3663),
     * also built in reverse */
    const int rev = !rc->shadow &&
                    (rc->origin < 0 || own_list(&tab->surfaces[rc->origin]));

    /* op accounting of per-ray culling: settle each element's ops */
    float inv[3] = {0.0f, 0.0f, 0.0f};
    if (o->anc_off != NULL)
    {
        for (int a = 0; a < 3; a++) inv[a] = slab_rcp(rc->ray[a]);
        o->ops_box += 3;
    }
    uint64_t mark = o->ops;
    int reach = 1, skip_until = -1;
    const int any = rc->shadow && o->any_off != NULL;
    if (any) memset(o->any_res, -1, (size_t)tab->n_step_any);
#define SETTLE() do { if (!reach) o->ops_skip += o->ops - mark; } while (0)

    for (int e = 0; e < tab->n_elem; e++)
    {
        SETTLE();
        mark = o->ops;
        reach = any ? cull_reach_any(o, rc->org, inv, rc->t_min, t_buf, e)
                    : cull_reach(o, rc->org, inv, rc->t_min, t_buf, e, &skip_until);
        const int si = tab->elems[e].surf;
        const qr_surface *s = &tab->surfaces[si];
        const int sh = s->shift * 3;
        const int same = si == rc->origin;

        /* OO_loc: secondary rays reuse the origin's local hit as diff */
        if (same)
        {
            for (int a = 0; a < 3; a++) w.dff[sh + a] = rc->oloc[a];
        }

        if (s->tag >= 0 && w.cache_last >= 0)
        {
            /* cached transform from the trnode (2466-2510) */
            if (!same)
            {
                for (int a = 0; a < 3; a++) w.dff[3 + a] = w.dff[a] - s->pos[a];
                o->ops += 3;
            }
            if (e == w.cache_last) w.cache_last = -1;
        }
        else
        {
            int do_ray = 0;
            if (same)
            {
                do_ray = 1;
            }
            else
            {
                float d[3];
                for (int a = 0; a < 3; a++) d[a] = rc->org[a] - s->pos[a];
                o->ops += 3;
                for (int a = 0; a < 3; a++) w.dff[a] = d[a];
                if (s->trm != 0)
                {
                    float t3[3];
                    transform_in(o, s, d, t3);
                    if (s->tag < 0)
                    {
                        for (int a = 0; a < 3; a++) w.dff[a] = t3[a];
                        w.cache_last = tab->elems[e].last;
                    }
                    else
                    {
                        for (int a = 0; a < 3; a++) w.dff[3 + a] = t3[a];
                    }
                    do_ray = 1;
                }
            }
            if (do_ray)
            {
                transform_in(o, s, rc->ray, w.ray6 + 3);
            }
        }

        if (s->solver == QR_SOLVE_NONE) continue;

        /* op accounting: a transformed surface's own box, tested in its
         * tracing frame after the set-up above; work behind a miss is skipped */
        if (reach && o->elem_step != NULL && tab->steps[o->elem_step[e]].skip == -2)
        {
            const qr_cull_step *st = &tab->steps[o->elem_step[e]];
            float li[3];
            for (int a = 0; a < 3; a++) li[a] = slab_rcp(w.ray6[3 + a]);
            o->ops_box += 31;
            if (!box_meets(st, w.dff + 3, li, rc->t_min, t_buf))
            {
                mark = o->ops;
                reach = 0;
            }
        }

        if (s->solver == QR_SOLVE_PLANE)
        {
            if (same) continue;
            int k = s->amap[2] + sh;
            float dff_k = fneg_if(w.dff[k], s->asgn[2]);
            float ray_k = fneg_if(w.ray6[k], s->asgn[2]);
            dff_k = -dff_k;
            if (!(ray_k != 0.0f)) continue;
            float t = dff_k / ray_k;
            o->ops += 1;
            int side = ray_k < 0.0f ? 0 : 1;
            if (!clip_root(o, &w, rc, t_buf, rev && best->srf >= 0, si, t,
                           side, 0, 0.0f)) continue;
            audit_hit(o, si, w.hit, w.new6 + sh, rc->org, w.dff + 3);

            if (rc->shadow)
            {
                if (shadows(s, side)) { SETTLE(); return 1; }
                continue;
            }
            t_buf = t;
            best->srf = si;
            best->side = side;
            best->t = t;
            memcpy(best->hit, w.hit, sizeof(w.hit));
            memcpy(best->loc, w.new6 + sh, sizeof(best->loc));
            best->tex_uv[0] = fneg_if(w.new6[s->amap[0] + sh], s->asgn[0]);
            best->tex_uv[1] = fneg_if(w.new6[s->amap[1] + sh], s->asgn[1]);
            continue;
        }

        /* quadric coefficients */
        float a_val, b_val, c_val, d_val;
        if (s->solver == QR_SOLVE_QUADRIC)
        {
            const float *r = w.ray6 + sh;
            const float *d = w.dff + sh;
            float ax[3], bx[3], cx[3];
            for (int q = 0; q < 3; q++)
            {
                float sri = s->sci[q] * r[q];
                float sdi = s->sci[q] * d[q];
                sdi = sdi - s->scj[q];
                ax[q] = r[q] * sri;
                bx[q] = r[q] * sdi;
                sdi = sdi - s->scj[q];
                cx[q] = d[q] * sdi;
            }
            a_val = ax[0] + ax[1];
            b_val = bx[0] + bx[1];
            c_val = cx[0] + cx[1];
            a_val = a_val + ax[2];
            b_val = b_val + bx[2];
            c_val = c_val + cx[2];
            c_val = c_val - s->sci[3];
            d_val = b_val * b_val - c_val * a_val;
            o->ops += 31;
        }
        else
        {
            int mi = s->amap[0], mk = s->amap[2];
            float ri = w.ray6[mi + sh], di = w.dff[mi + sh];
            float rk = w.ray6[mk + sh], dk = w.dff[mk + sh];
            float si_ = s->sci[mi], sk_ = s->sci[mk];
            float dx = di * rk - dk * ri;
            dx = dx * dx;
            dx = dx * si_;
            dx = dx * sk_;
            d_val = fabsf(dx);
            float bi = si_ * di, bk = sk_ * dk;
            bi = bi * ri;
            bk = bk * rk;
            b_val = bi + bk;
            float ci = di * di, ck = dk * dk;
            ci = ci * si_;
            ck = ck * sk_;
            c_val = ci + ck;
            float ai = ri * ri, ak = rk * rk;
            ai = ai * si_;
            ak = ak * sk_;
            a_val = ai + ak;
            o->ops += 20;
            o->n_twoplane++;
        }

        /* QD_rts (5785-5960) */
        if (rc->primary)
        {
            float bb = b_val * b_val, ac = fabsf(a_val * c_val);
            float nm = bb > ac ? bb : ac;
            if (fabsf(d_val) < 1e-5f * nm) o->near_zero = 1;
        }
        if (!(0.0f <= d_val)) continue;
        float nb = -b_val;
        int dmask = d_val < s->d_eps;

        float sd = sqrtf(d_val);
        sd = fxor(sd, fbits(nb) & 0x80000000u);
        float q = nb + sd;
        int m_pos = 0.0f <= sd;
        int m_neg = 0.0f > sd;
        float t1n = m_neg ? q : (m_pos ? c_val : 0.0f);
        float t1d = m_neg ? a_val : (m_pos ? q : 0.0f);
        float t2n = m_neg ? c_val : (m_pos ? q : 0.0f);
        float t2d = m_neg ? q : (m_pos ? a_val : 0.0f);
        o->ops += 2;

        float t1 = 0.0f, t2 = 0.0f;
        int t1m = 0, t2m = 0;
        uint32_t amask = fbits(a_val) & 0x80000000u;
        if (dmask)
        {
            o->n_dmask++;
            /* nan roots to zero, equate then split (5908-5960) */
            if (t1n == 0.0f) t1d = 1.0f;
            if (t2n == 0.0f) t2d = 1.0f;
            t1 = t1n / t1d;
            t2 = t2n / t2d;
            t1m = t1d != 0.0f;
            t2m = t2d != 0.0f;
            float tdf = t1 - t2;
            tdf = fxor(tdf, amask);
            int fm = 0.0f <= tdf;
            if (!fm) tdf = 0.0f;
            float te = fm ? s->t_eps : 0.0f;
            te = te * t1;
            te = fabsf(te);
            tdf = tdf * -0.5f;
            tdf = tdf - te;
            tdf = fxor(tdf, amask);
            if (!(t1m && t2m)) tdf = 0.0f;
            t1 = t1 + tdf;
            t2 = t2 - tdf;
            o->ops += 8;
        }

        /* a-mixed order: inner root first when a < 0 (5982-5994) */
        int order[2];
        if (0.0f > a_val) { order[0] = 1; order[1] = 0; }
        else              { order[0] = 0; order[1] = 1; }

        for (int pass = 0; pass < 2; pass++)
        {
            int side = order[pass];
            if (same)
            {
                /* CHECK_SIDE (558-567) */
                int fl = rc->flags & (FLAG_SIDE | FLAG_PASS);
                if (fl == 1 - side || fl == 2 + side) continue;
            }
            float t;
            int tm;
            if (side == 0)
            {
                if (!dmask) { t1 = t1n / t1d; t1m = t1d != 0.0f; o->ops += 1; }
                t = t1; tm = t1m;
            }
            else
            {
                if (!dmask) { t2 = t2n / t2d; t2m = t2d != 0.0f; o->ops += 1; }
                t = t2; tm = t2m;
            }
            if (!tm) continue;
            if (!clip_root(o, &w, rc, t_buf, rev && best->srf >= 0, si, t,
                           side, dmask, bitsf(amask)))
                continue;
            audit_hit(o, si, w.hit, w.new6 + sh, rc->org, w.dff + 3);

            if (rc->shadow)
            {
                if (shadows(s, side)) { SETTLE(); return 1; }
            }
            else
            {
                t_buf = t;
                best->srf = si;
                best->side = side;
                best->t = t;
                memcpy(best->hit, w.hit, sizeof(w.hit));
                memcpy(best->loc, w.new6 + sh, sizeof(best->loc));
                best->tex_uv[0] = 0.0f;
                best->tex_uv[1] = 0.0f;
            }
            /* overdraw skip: the first root survived, the second cannot
             * be nearer (6092-6099) */
            break;
        }
    }
    SETTLE();
#undef SETTLE
    return best->srf >= 0;
}

/* ------------------------------------------------------------------------ */
/* shading (MT 3240-3416, LT 3931-4436, TR 4440-4936, RF 4944-5357)         */
/* ------------------------------------------------------------------------ */

static void trace(oracle *o, const ray_ctx *rc, float col[3]);

static void normal_at(oracle *o, const hit_rec *h, float nrm[3])
{
    const qr_scene_tables *tab = o->tab;
    const qr_surface *s = &tab->surfaces[h->srf];
    float n6[6] = {0, 0, 0, 0, 0, 0};
    const int sh = s->shift * 3;
    uint32_t ts = h->side ? 0x80000000u : 0u;

    if (s->nrm_kind == 1)
    {
        n6[s->amap[0] + sh] = 0.0f;
        n6[s->amap[1] + sh] = 0.0f;
        n6[s->amap[2] + sh] = fneg_if(fxor(1.0f, ts), s->asgn[2]);
    }
    else
    {
        float l[3];
        for (int a = 0; a < 3; a++)
        {
            l[a] = h->loc[a] * s->sci[a];
            if (s->nrm_kind == 2) l[a] = l[a] - s->scj[a];
        }
        float r = l[0] * l[0] + l[1] * l[1];
        r = r + l[2] * l[2];
        float inv = rsq(r);
        inv = fxor(inv, ts);
        for (int a = 0; a < 3; a++) n6[sh + a] = l[a] * inv;
        o->ops += 16;
    }

    if (s->trm == 0)
    {
        nrm[0] = n6[0]; nrm[1] = n6[1]; nrm[2] = n6[2];
        return;
    }

    /* transposed inverse of the trnode, renormalised when scaled (3273-3348) */
    const qr_surface *t = &tab->surfaces[s->trnode];
    const float *ni = n6 + 3;
    float x = t->tc[0][0] * ni[0];
    float y = t->tc[1][1] * ni[1];
    float z = t->tc[2][2] * ni[2];
    o->ops += 3;
    int renorm = 1;
    if (t->trm != 1)
    {
        x = x + t->tc[1][0] * ni[1];
        x = x + t->tc[2][0] * ni[2];
        y = y + t->tc[0][1] * ni[0];
        y = y + t->tc[2][1] * ni[2];
        z = z + t->tc[0][2] * ni[0];
        z = z + t->tc[1][2] * ni[1];
        o->ops += 12;
        if (t->trm == 2) renorm = 0;
    }
    if (renorm)
    {
        float r = x * x + y * y;
        r = r + z * z;
        float inv = rsq(r);
        x = x * inv; y = y * inv; z = z * inv;
        o->ops += 10;
    }
    nrm[0] = x; nrm[1] = y; nrm[2] = z;
}

static void texel_color(oracle *o, const hit_rec *h, const qr_material *m,
                        uint32_t props, float tex[3])
{
    uint32_t idx = 0;
    if (props & QR_PROP_TEXTURE)
    {
        float tx = h->tex_uv[m->tmap[0]];
        float ty = h->tex_uv[m->tmap[1]];
        tx = tx - m->xoffs;
        ty = ty - m->yoffs;
        tx = tx * m->xscal;
        ty = ty * m->yscal;
        uint32_t ix = (uint32_t)cvt_floor(tx) & m->xmask;
        uint32_t iy = (uint32_t)cvt_floor(ty) & m->ymask;
        iy = iy << m->yshft;
        idx = ix + iy;
        o->ops += 4;
    }
    uint32_t tv = o->tab->texels[m->tex_off + idx];
    tex[0] = (float)(int32_t)((tv >> 16) & 0xFF) / 255.0f;
    tex[1] = (float)(int32_t)((tv >> 8) & 0xFF) / 255.0f;
    tex[2] = (float)(int32_t)((tv >> 0) & 0xFF) / 255.0f;
    o->ops += 3;
    if (props & QR_PROP_GAMMA)
    {
        tex[0] = tex[0] * tex[0];
        tex[1] = tex[1] * tex[1];
        tex[2] = tex[2] * tex[2];
        o->ops += 3;
    }
}

/* specular power in fixed point 28.4 (4225-4283) */
static float spec_pow(oracle *o, float s, uint32_t pw)
{
    uint32_t fr = pw & 0xF;
    float x2 = s, x4 = s, r = 1.0f;
    while (fr != 0)
    {
        x4 = sqrtf(x4);
        uint32_t bit = 0x8 & fr;
        fr = (fr << 1) & 0xF;
        if (bit) r = r * x4;
        o->ops += 2;
    }
    uint32_t ip = pw >> 4;
    if (ip != 0)
    {
        float fp = r;
        r = 1.0f;
        while (ip != 0)
        {
            uint32_t bit = ip & 1;
            ip >>= 1;
            if (bit) r = r * x2;
            x2 = x2 * x2;
            o->ops += 2;
        }
        r = r * fp;
        o->ops += 1;
    }
    return r;
}

/* secondary ray leaving hit h along dir (TR/RF/PT ENTER blocks) */
static void child_ray(const ray_ctx *rc, const hit_rec *h, const float dir[3],
                      int flags, ray_ctx *c)
{
    memset(c, 0, sizeof(*c));
    memcpy(c->org, h->hit, sizeof(c->org));
    memcpy(c->ray, dir, sizeof(c->ray));
    c->t_min = 0.0f;
    c->t_max = FLT_MAX;
    c->flags = flags;
    c->origin = h->srf;
    memcpy(c->oloc, h->loc, sizeof(c->oloc));
    c->depth = rc->depth - 1;
}

/* refraction and reflection set-up of a hit: which children it spawns,
 * their directions and the transmitted / reflected fractions with Fresnel
 * (TR 4442-4936, RF 4944-5357) */
typedef struct trrf
{
    int   tr, rf;
    int   fresnel_tr;   /* Fresnel computed in the transparency block */
    float c_trn, c_rfl;
    float tdir[3], rdir[3];
} trrf;

static void trrf_setup(oracle *o, const ray_ctx *rc, const hit_rec *h,
                       const float nrm[3], const qr_material *m, uint32_t props,
                       trrf *out)
{
    (void)h;
    /* transparency and refraction (4442-4936) */
    float c_trn = m->c_trn;
    float c_rfl = m->c_rfl;
    out->tr = 0;
    out->rf = 0;
    out->fresnel_tr = !(props & QR_PROP_OPAQUE) && (props & QR_PROP_FRESNEL);
    if (!(props & QR_PROP_OPAQUE))
    {
        int m_trn = 1;
        float nray[3];
        if ((props & QR_PROP_REFRACT) || (props & QR_PROP_FRESNEL))
        {
            float r2 = rc->ray[0] * rc->ray[0];
            r2 = r2 + rc->ray[1] * rc->ray[1];
            r2 = r2 + rc->ray[2] * rc->ray[2];
            float inv = rsq(r2);
            float rn[3] = {rc->ray[0] * inv, rc->ray[1] * inv, rc->ray[2] * inv};
            float dot = rn[0] * nrm[0];
            dot = dot + rn[1] * nrm[1];
            dot = dot + rn[2] * nrm[2];
            float x0 = dot * m->c_rfr;
            float k = x0 * x0;
            k = k + 1.0f;
            k = k - m->rfr_2;
            o->ops += 18;
            int tir = 0;
            if (props & QR_PROP_FRESNEL)
            {
                if (!(0.0f <= k)) { tir = 1; m_trn = 0; }
            }
            if (tir)
            {
                c_trn = 0.0f;
                c_rfl = m->c_rfl + m->c_trn;
                o->ops += 1;
            }
            else
            {
                float sk = sqrtf(k);
                x0 = x0 + sk;
                o->ops += 2;
                if (props & QR_PROP_REFRACT)
                {
                    for (int a = 0; a < 3; a++)
                    {
                        float v = nrm[a] * x0;
                        nray[a] = rn[a] * m->c_rfr;
                        nray[a] = nray[a] - v;
                    }
                    o->ops += 9;
                }
                else
                {
                    memcpy(nray, rc->ray, sizeof(nray));
                }
                if (props & QR_PROP_FRESNEL)
                {
                    float n = m->c_rfr;
                    float a2 = dot * n;
                    a2 = a2 - sk;
                    float skn = sk * n;
                    float a1 = dot + skn;
                    float a3 = dot - skn;
                    float f0 = x0 / a2;
                    float f1 = a1 / a3;
                    f0 = f0 * f0;
                    f1 = f1 * f1;
                    f0 = f0 + f1;
                    f0 = f0 * -0.5f;
                    f0 = fabsf(f0);
                    f0 = f0 * m->c_trn;
                    c_trn = m->c_trn - f0;
                    c_rfl = m->c_rfl + f0;
                    o->ops += 14;
                }
            }
        }
        else
        {
            memcpy(nray, rc->ray, sizeof(nray));
        }

        if (m_trn)
        {
            out->tr = 1;
            memcpy(out->tdir, nray, sizeof(out->tdir));
        }
    }

    /* reflection (4944-5357) */
    int do_rfl = (props & QR_PROP_REFLECT) ||
                 (!(props & QR_PROP_OPAQUE) && (props & QR_PROP_FRESNEL));
    if (do_rfl)
    {
        float r2 = rc->ray[0] * rc->ray[0];
        r2 = r2 + rc->ray[1] * rc->ray[1];
        r2 = r2 + rc->ray[2] * rc->ray[2];
        float inv = rsq(r2);
        float rn[3] = {rc->ray[0] * inv, rc->ray[1] * inv, rc->ray[2] * inv};
        float dot = rn[0] * nrm[0];
        dot = dot + rn[1] * nrm[1];
        dot = dot + rn[2] * nrm[2];
        float nray[3];
        for (int a = 0; a < 3; a++)
        {
            float v = nrm[a] * dot;
            nray[a] = rn[a] - v;
            nray[a] = nray[a] - v;
        }
        o->ops += 20;

        if ((props & QR_PROP_FRESNEL) && (props & QR_PROP_OPAQUE))
        {
            float f0;
            if (props & QR_PROP_METAL)
            {
                float x6 = m->c_rcp;
                float x4 = dot * x6;
                x4 = x4 + x4;
                float x0 = dot * dot;
                x6 = x6 * x6;
                x6 = x6 + m->ext_2;
                float x1 = x0 * x6;
                x0 = x0 + x6;
                x1 = x1 + 1.0f;
                float x2 = x0, x3 = x1;
                x0 = x0 + x4;
                x1 = x1 + x4;
                x2 = x2 - x4;
                x3 = x3 - x4;
                x0 = x0 / x2;
                x1 = x1 / x3;
                x0 = x0 + x1;
                x0 = x0 * -0.5f;
                f0 = fabsf(x0);
                o->ops += 16;
            }
            else
            {
                float n = m->c_rfr;
                float x0 = dot * n;
                float k = x0 * x0;
                k = k + 1.0f;
                k = k - m->rfr_2;
                float sk = sqrtf(k);
                x0 = x0 + sk;
                float a2 = dot * n;
                a2 = a2 - sk;
                float skn = sk * n;
                float a1 = dot + skn;
                float a3 = dot - skn;
                float q0 = x0 / a2;
                float q1 = a1 / a3;
                q0 = q0 * q0;
                q1 = q1 * q1;
                q0 = q0 + q1;
                q0 = q0 * -0.5f;
                f0 = fabsf(q0);
                o->ops += 19;
            }
            f0 = f0 - 1.0f;
            f0 = f0 * m->c_rfl;
            c_rfl = m->c_rfl + f0;
            o->ops += 3;
        }

        out->rf = 1;
        memcpy(out->rdir, nray, sizeof(out->rdir));
    }
    out->c_trn = c_trn;
    out->c_rfl = c_rfl;
}

static void shade(oracle *o, const ray_ctx *rc, const hit_rec *h, float col[3])
{
    const qr_scene_tables *tab = o->tab;
    const qr_frame_desc *fd = o->fd;
    const qr_surface *s = &tab->surfaces[h->srf];
    const uint32_t props = s->props[h->side];
    const qr_material *m = &tab->materials[s->mat[h->side]];

    float nrm[3];
    normal_at(o, h, nrm);

    float tex[3];
    texel_color(o, h, m, props, tex);

    /* ambient (3933-3996) */
    float c[3];
    if (props & QR_PROP_LIGHT)
    {
        c[0] = tex[0]; c[1] = tex[1]; c[2] = tex[2];
    }
    else
    {
        c[0] = tex[0] * fd->amb[0];
        c[1] = tex[1] * fd->amb[1];
        c[2] = tex[2] * fd->amb[2];
        float f = 1.0f - m->c_trn;
        f = f - m->c_rfl;
        if (!(0.0f <= f)) f = 0.0f;
        c[0] = c[0] * f; c[1] = c[1] * f; c[2] = c[2] * f;
        o->ops += 8;
    }
    col[0] = c[0]; col[1] = c[1]; col[2] = c[2];

    /* lights (4002-4391) */
    if (!(props & QR_PROP_LIGHT))
    {
        for (int li = 0; li < tab->n_lgt; li++)
        {
            const qr_light *l = &tab->lights[li];
            float L[3];
            L[0] = l->pos[0] - h->hit[0];
            L[1] = l->pos[1] - h->hit[1];
            L[2] = l->pos[2] - h->hit[2];
            float dot = L[0] * nrm[0] + L[1] * nrm[1];
            dot = dot + L[2] * nrm[2];
            o->ops += 8;
            if (!(0.0f < dot)) continue;

            /* hard shadow: any-hit walk to the light (4040-4095) */
            ray_ctx sr;
            memset(&sr, 0, sizeof(sr));
            memcpy(sr.org, h->hit, sizeof(sr.org));
            memcpy(sr.ray, L, sizeof(sr.ray));
            sr.t_min = 0.0f;
            sr.t_max = l->t_max;
            sr.flags = h->side | FLAG_SHAD;
            sr.origin = h->srf;
            memcpy(sr.oloc, h->loc, sizeof(sr.oloc));
            sr.depth = rc->depth - 1;
            sr.shadow = 1;
            hit_rec sh;
            o->rays++;
            if (walk(o, &sr, &sh)) continue;

            float r2 = L[0] * L[0] + L[1] * L[1];
            r2 = r2 + L[2] * L[2];
            o->ops += 5;

            float dff, sdot;
            if (props & QR_PROP_DIFFUSE)
            {
                float ir = rsq(r2);
                float r1 = ir * r2;
                float at = r2 * l->a_qdr;
                float lr = r1 * l->a_lnr;
                at = at + l->a_cnt;
                at = at + lr;
                float ia = rsq(at);
                sdot = dot;
                dff = dot * ia;
                dff = dff * ir;
                dff = dff * m->l_dff;
                o->ops += 12;
            }
            else
            {
                sdot = dot;
                dff = 0.0f;
            }

            int plain_spc = 0;
            float spc = 0.0f;
            if (props & QR_PROP_SPECULAR)
            {
                float R[3], sn;
                for (int a = 0; a < 3; a++)
                {
                    float v = sdot * nrm[a];
                    R[a] = L[a] - v;
                    R[a] = R[a] - v;
                }
                float rr = rc->ray[2] * rc->ray[2];
                rr = rr + rc->ray[0] * rc->ray[0];
                rr = rr + rc->ray[1] * rc->ray[1];
                sn = R[0] * rc->ray[0] + R[1] * rc->ray[1];
                sn = sn + R[2] * rc->ray[2];
                o->ops += 19;
                if (0.0f < sn)
                {
                    sn = sn * rsq(rr);
                    sn = sn * rsq(r2);
                    o->ops += 6;
                    spc = spec_pow(o, sn, m->l_pow);
                    spc = spc * m->l_spc;
                    o->ops += 1;
                    if (props & QR_PROP_METAL)
                    {
                        /* metal: highlight tinted by the surface color */
                        dff = dff + spc;
                        o->ops += 1;
                    }
                    else
                    {
                        /* plain: highlight in the light's color (4323-4371) */
                        plain_spc = 1;
                    }
                }
            }

            float lc[3];
            if (plain_spc)
            {
                for (int a = 0; a < 3; a++)
                {
                    float v = tex[a] * dff;
                    v = v * l->col[a];
                    lc[a] = v + l->col[a] * spc;
                }
                o->ops += 12;
            }
            else
            {
                for (int a = 0; a < 3; a++)
                {
                    float v = tex[a] * l->col[a];
                    lc[a] = v * dff;
                }
                o->ops += 6;
            }
            col[0] = lc[0] + col[0];
            col[1] = lc[1] + col[1];
            col[2] = lc[2] + col[2];
            o->ops += 3;
        }
    }

    /* transparency, refraction, reflection (4442-5357) */
    trrf t;
    trrf_setup(o, rc, h, nrm, m, props, &t);
    if (t.tr && rc->depth != 0)
    {
        ray_ctx tr;
        child_ray(rc, h, t.tdir, h->side | FLAG_PASS, &tr);
        float tc[3];
        trace(o, &tr, tc);
        for (int a = 0; a < 3; a++)
        {
            float v = tc[a] * t.c_trn;
            col[a] = v + col[a];
        }
        o->ops += 6;
    }
    if (t.rf && rc->depth != 0)
    {
        ray_ctx rr;
        child_ray(rc, h, t.rdir, h->side, &rr);
        float rcol[3];
        trace(o, &rr, rcol);
        for (int a = 0; a < 3; a++)
        {
            float v = rcol[a] * t.c_rfl;
            col[a] = v + col[a];
        }
        o->ops += 6;
    }
}

static void trace(oracle *o, const ray_ctx *rc, float col[3])
{
    hit_rec h;
    o->rays++;
    col[0] = col[1] = col[2] = 0.0f;
    if (!walk(o, rc, &h))
    {
        if (rc->primary) o->hit_id = -1;
        return;
    }
    if (rc->primary)
    {
        const qr_surface *s = &o->tab->surfaces[h.srf];
        o->hit_id = (s->id << 1) | h.side;
    }
    ray_ctx sub = *rc;
    sub.primary = 0;
    shade(o, &sub, &h, col);
}

/* anti-aliasing resolve of a pixel's clamped samples, halve then pairwise
 * add per pass (7047-7133), and ARGB8 packing (1922-1945) */
static uint32_t resolve_pixel(const qr_frame_desc *fd, const float acc[4][3])
{
    const int ns = fd->fsaa == 2 ? 4 : fd->fsaa == 1 ? 2 : 1;
    float c[3];
    for (int a = 0; a < 3; a++)
    {
        if (ns == 1)
        {
            c[a] = acc[0][a];
        }
        else if (ns == 2)
        {
            c[a] = acc[0][a] * 0.5f + acc[1][a] * 0.5f;
        }
        else
        {
            float p0 = acc[0][a] * 0.5f + acc[1][a] * 0.5f;
            float p1 = acc[2][a] * 0.5f + acc[3][a] * 0.5f;
            c[a] = p0 * 0.5f + p1 * 0.5f;
        }
    }
    uint32_t px = 0;
    const int shifts[3] = {16, 8, 0};
    for (int a = 0; a < 3; a++)
    {
        float v = c[a];
        if (fd->gamma & QR_PROP_GAMMA) v = sqrtf(v);
        v = v * 255.0f;
        px |= ((uint32_t)cvt_rne(v) & 0xFF) << shifts[a];
    }
    return px;
}

/* ------------------------------------------------------------------------ */
/* path tracer (PT domain of render0)                                       */
/* ------------------------------------------------------------------------ */

/* next number of the 32-bit LCG as a float in [0, 1): the top 24 bits over
 * 2^24 (GET_RANDOM_F, 1949-1980, RT_PRNG LCG24 tracer.h:58; constants
 * engine.cpp:885-892) */
static float pt_random(uint32_t *st)
{
    *st = *st * 214013u + 2531011u;
    return (float)((*st >> 8) & 0xFFFFFFu) / 16777216.0f;
}

/* power series of sin and cos (sinps_rr / cosps_rr, 1982-2015; coefficients
 * engine.cpp:912-918), multiply and add kept separate */
static float pt_sin(float x)
{
    const float t = x * x;
    float d = x, v = x * t;
    d = fmaf(v, -0.1666666666666666666666666666666666666666666f, d);
    v = v * t; d = fmaf(v, +0.0083333333333333333333333333333333333333333f, d);
    v = v * t; d = fmaf(v, -0.0001984126984126984126984126984126984126984f, d);
    v = v * t; d = fmaf(v, +0.0000027557319223985890652557319223985890652f, d);
    return d;
}

static float pt_cos(float x)
{
    const float t = x * x;
    float d = 1.0f, v = d * t;
    d = fmaf(v, -0.5f, d);
    v = v * t; d = fmaf(v, +0.0416666666666666666666666666666666666666666f, d);
    v = v * t; d = fmaf(v, -0.0013888888888888888888888888888888888888888f, d);
    v = v * t; d = fmaf(v, +0.0000248015873015873015873015873015873015873f, d);
    return d;
}

/* tent-filtered pixel jitter (2266-2342): two draws, scaled by 1/2, and by
 * 1/2 again with anti-aliasing */
static float pt_tent(uint32_t *st, int fsaa)
{
    float r = pt_random(st);
    r = r + r;
    float d = r < 1.0f ? sqrtf(r) - 1.0f : 1.0f - sqrtf(2.0f - r);
    d = d * 0.5f;
    if (fsaa != 0) d = d * 0.5f;
    return d;
}

typedef struct pt_job
{
    ray_ctx  rc;
    float    mul[3];        /* path throughput (MUL_R/G/B) */
    uint32_t st;            /* the path's random state */
} pt_job;

#define PT_STACK 32
#define PT_ROULETTE_DEPTH 5     /* RT_STACK_DEPTH - 5, 3488 */
#define PT_SPLIT_DEPTH    8     /* RT_STACK_DEPTH - 2, 4690 */

/* one path-traced camera sample: every hit adds its emission times the
 * path throughput to acc (PT_emt, 3425-3470); a diffuse hit continues along a
 * cosine-distributed direction with throughput times l_dff and the surface
 * colour, with Russian roulette by the colour's maximum at remaining depth
 * <= 5 (3479-3800); transparent Fresnel surfaces pick refraction or
 * reflection at random at remaining depth <= 8 with smallpt's weights
 * (4686-4752); the refraction / reflection children carry throughput times
 * c_trn / c_rfl (4762-4800, 5175-5215).  The diffuse child continues the
 * random state after its two draws, the refraction and reflection children
 * after every draw of the hit. */
static void pt_sample(oracle *o, const ray_ctx *primary, uint32_t st,
                      float pts_o, float acc[3])
{
    const qr_scene_tables *tab = o->tab;
    pt_job stk[PT_STACK];
    int sp = 0;
    stk[sp].rc = *primary;
    stk[sp].mul[0] = stk[sp].mul[1] = stk[sp].mul[2] = 1.0f;
    stk[sp].st = st;
    sp++;
    while (sp > 0)
    {
        pt_job j = stk[--sp];
        hit_rec h;
        o->rays++;
        if (!walk(o, &j.rc, &h)) continue;
        const qr_surface *s = &tab->surfaces[h.srf];
        const uint32_t props = s->props[h.side];
        const qr_material *m = &tab->materials[s->mat[h.side]];
        float nrm[3], tex[3];
        normal_at(o, &h, nrm);
        texel_color(o, &h, m, props, tex);

        for (int a = 0; a < 3; a++)
        {
            float c = m->col[a] * pts_o;
            c = c * j.mul[a];
            acc[a] = acc[a] + c;
        }

        if (props & QR_PROP_DIFFUSE)
        {
            int alive = 1;
            if (j.rc.depth <= PT_ROULETTE_DEPTH)
            {
                float p = fmaxf(fmaxf(tex[0], tex[1]), tex[2]);
                float r = pt_random(&j.st);
                if (r < p)
                {
                    float ip = 1.0f / p;
                    tex[0] = tex[0] * ip; tex[1] = tex[1] * ip; tex[2] = tex[2] * ip;
                }
                else
                {
                    alive = 0;
                }
            }
            if (alive)
            {
                const float *ry = j.rc.ray;
                float u[3], v[3];
                u[0] = nrm[1] * ry[2] - nrm[2] * ry[1];
                u[1] = nrm[2] * ry[0] - nrm[0] * ry[2];
                u[2] = nrm[0] * ry[1] - nrm[1] * ry[0];
                float l2 = u[0] * u[0] + u[1] * u[1];
                l2 = l2 + u[2] * u[2];
                float il = rsq(l2);
                u[0] = u[0] * il; u[1] = u[1] * il; u[2] = u[2] * il;
                v[0] = nrm[1] * u[2] - nrm[2] * u[1];
                v[1] = nrm[2] * u[0] - nrm[0] * u[2];
                v[2] = nrm[0] * u[1] - nrm[1] * u[0];
                float r1 = pt_random(&j.st);
                float st_ = sqrtf(r1);
                float ct_ = sqrtf(1.0f - r1);
                float r2 = pt_random(&j.st);
                float phi = (r2 + r2) * 3.14159265358979323846f;
                phi = phi - 3.14159265358979323846f;
                float cu = pt_cos(phi) * st_, sv = pt_sin(phi) * st_;
                float dir[3];
                for (int a = 0; a < 3; a++)
                {
                    dir[a] = nrm[a] * ct_;
                    dir[a] = dir[a] + u[a] * cu;
                    dir[a] = dir[a] + v[a] * sv;
                }
                if (j.rc.depth != 0 && sp < PT_STACK)
                {
                    pt_job *c = &stk[sp++];
                    child_ray(&j.rc, &h, dir, h.side, &c->rc);
                    for (int a = 0; a < 3; a++)
                    {
                        float w = j.mul[a] * m->l_dff;
                        c->mul[a] = w * tex[a];
                    }
                    c->st = j.st;
                }
            }
        }

        trrf t;
        trrf_setup(o, &j.rc, &h, nrm, m, props, &t);
        if (t.fresnel_tr && j.rc.depth <= PT_SPLIT_DEPTH)
        {
            float r = pt_random(&j.st);
            float q = t.c_rfl / (t.c_trn + t.c_rfl);
            q = q * 0.5f;
            float pr = 0.25f + q;
            int pick_tr = r >= pr;
            t.c_trn = pick_tr ? t.c_trn / (1.0f - pr) : 0.0f;
            t.c_rfl = pick_tr ? 0.0f : t.c_rfl / pr;
            t.tr = t.tr && pick_tr;
            t.rf = t.rf && !pick_tr;
        }
        if (t.tr && j.rc.depth != 0 && sp < PT_STACK)
        {
            pt_job *c = &stk[sp++];
            child_ray(&j.rc, &h, t.tdir, h.side | FLAG_PASS, &c->rc);
            for (int a = 0; a < 3; a++) c->mul[a] = j.mul[a] * t.c_trn;
            c->st = j.st;
        }
        if (t.rf && j.rc.depth != 0 && sp < PT_STACK)
        {
            pt_job *c = &stk[sp++];
            child_ray(&j.rc, &h, t.rdir, h.side, &c->rc);
            for (int a = 0; a < 3; a++) c->mul[a] = j.mul[a] * t.c_rfl;
            c->st = j.st;
        }
    }
}

int qr_oracle_render_pt(const qr_scene_tables *tab, const qr_frame_desc *fd,
                        uint32_t *frame, float *plane_r, float *plane_g,
                        float *plane_b, uint32_t *seeds, int y0, int y1,
                        qr_oracle_aux *aux)
{
    oracle o;
    memset(&o, 0, sizeof(o));
    o.tab = tab;
    o.fd = fd;
    if (cull_setup(&o) != 0 || cull_setup_any(&o) != 0)
    {
        cull_free(&o);
        return -1;
    }
    const int ns = fd->fsaa == 2 ? 4 : fd->fsaa == 1 ? 2 : 1;
    /* running mean over passes (FF_pts, 2113-2140): pts_c counts this pass */
    const float pts_o = 1.0f / fd->pts_c;
    const float pts_u = 1.0f - pts_o;

    for (int y = y0; y < y1; y++)
    {
        for (int x = 0; x < fd->x_res; x++)
        {
            float acc[4][3];
            for (int k = 0; k < ns; k++)
            {
                const size_t slot = ((size_t)y * fd->x_res + x) * ns + k;
                uint32_t st = seeds[slot];
                int sk = k;
                if (fd->fsaa == 1 && (x & 1)) sk = k + 2;
                float hs = (float)x;
                float vs = (float)y;
                if (fd->fsaa != 0)
                {
                    hs = hs + fd->aa_h[sk];
                    vs = vs + fd->aa_v[sk];
                }
                hs = hs + pt_tent(&st, fd->fsaa);
                vs = vs + pt_tent(&st, fd->fsaa);
                ray_ctx rc;
                memset(&rc, 0, sizeof(rc));
                for (int a = 0; a < 3; a++)
                {
                    float hx = fd->hor[a] * hs;
                    float vx = fd->ver[a] * vs;
                    hx = hx + vx;
                    rc.ray[a] = hx + fd->dir[a];
                    rc.org[a] = fd->org[a];
                }
                rc.t_min = fd->t_min;
                rc.t_max = FLT_MAX;
                rc.origin = -1;
                rc.depth = fd->depth;
                float pl[3] = {plane_r[slot] * pts_u, plane_g[slot] * pts_u,
                               plane_b[slot] * pts_u};
                pt_sample(&o, &rc, st, pts_o, pl);
                plane_r[slot] = pl[0];
                plane_g[slot] = pl[1];
                plane_b[slot] = pl[2];
                /* the slot's seed steps once per pass (6963-6980) */
                seeds[slot] = seeds[slot] * 214013u + 2531011u;
                for (int a = 0; a < 3; a++) acc[k][a] = pl[a] < 1.0f ? pl[a] : 1.0f;
            }
            frame[(ptrdiff_t)y * fd->x_row + x] = resolve_pixel(fd, acc);
        }
    }
    if (aux != NULL)
    {
        aux->ops += o.ops;
        aux->rays += o.rays;
        aux->ops_skip += o.ops_skip;
        aux->ops_box += o.ops_box;
    }
    cull_free(&o);
    return 0;
}

/* seed plane of a path-traced frame: the low 32 bits of a 48-bit LCG chain
 * started at 1, one value per sample slot (reset_pseed, engine.cpp:3719-3740) */
void qr_oracle_pt_seeds(uint32_t *seeds, size_t n)
{
    uint64_t s = 1;
    for (size_t k = 0; k < n; k++)
    {
        s = (s * 25214903917ull + 11ull) & 0x0000FFFFFFFFFFFFull;
        seeds[k] = (uint32_t)s;
    }
}

int qr_oracle_render(const qr_scene_tables *tab, const qr_frame_desc *fd,
                     uint32_t *frame, int y0, int y1, qr_oracle_aux *aux)
{
    oracle o;
    memset(&o, 0, sizeof(o));
    o.tab = tab;
    o.fd = fd;
    o.boxes = aux != NULL ? aux->boxes : NULL;
    if (cull_setup(&o) != 0 || cull_setup_any(&o) != 0)
    {
        cull_free(&o);
        return -1;
    }

    const int ns = fd->fsaa == 2 ? 4 : fd->fsaa == 1 ? 2 : 1;

    for (int y = y0; y < y1; y++)
    {
        for (int x = 0; x < fd->x_res; x++)
        {
            float acc[4][3];
            for (int k = 0; k < ns; k++)
            {
                int sk = k;
                if (fd->fsaa == 1 && (x & 1)) sk = k + 2;
                float hs = (float)x;
                float vs = (float)y;
                if (fd->fsaa != 0)
                {
                    hs = hs + fd->aa_h[sk];
                    vs = vs + fd->aa_v[sk];
                }
                hs = hs + 0.0f;
                vs = vs + 0.0f;
                ray_ctx rc;
                memset(&rc, 0, sizeof(rc));
                for (int a = 0; a < 3; a++)
                {
                    float hx = fd->hor[a] * hs;
                    float vx = fd->ver[a] * vs;
                    hx = hx + vx;
                    rc.ray[a] = hx + fd->dir[a];
                    rc.org[a] = fd->org[a];
                }
                o.ops += 15;
                rc.t_min = fd->t_min;
                rc.t_max = FLT_MAX;
                rc.flags = 0;
                rc.origin = -1;
                rc.depth = fd->depth;
                rc.primary = 1;
                o.near_zero = 0;
                o.hit_id = -1;
                float col[3];
                trace(&o, &rc, col);
                for (int a = 0; a < 3; a++)
                {
                    acc[k][a] = col[a] < 1.0f ? col[a] : 1.0f;
                }
                if (aux != NULL)
                {
                    size_t idx = ((size_t)y * fd->x_res + x) * ns + k;
                    if (aux->ids) aux->ids[idx] = o.hit_id;
                    if (aux->flags) aux->flags[idx] = (uint8_t)o.near_zero;
                }
            }

            frame[(ptrdiff_t)y * fd->x_row + x] = resolve_pixel(fd, acc);
        }
    }
    if (aux != NULL)
    {
        aux->ops += o.ops;
        aux->rays += o.rays;
        aux->box_checked += o.box_checked;
        aux->box_outside += o.box_outside;
        if (o.box_excess > aux->box_excess) aux->box_excess = o.box_excess;
        aux->ops_skip += o.ops_skip;
        aux->ops_box += o.ops_box;
        aux->n_dmask += o.n_dmask;
        aux->n_twoplane += o.n_twoplane;
        aux->n_apex += o.n_apex;
    }
    cull_free(&o);
    return 0;
}
