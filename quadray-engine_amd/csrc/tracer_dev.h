/*
 * tracer_dev.h -- device code shared by the Whitted kernels (gfx950).
 *
 * QuadRay's per-ray path restated for one work-item per ray: the surface
 * list walk with transform caching, quadric / plane / two-plane roots,
 * axis and custom clippers, normals, texturing, lighting with hard shadows,
 * and the refraction / reflection set-up with Fresnel.  Arithmetic is fp32
 * in QuadRay's operation order: compiled with -ffp-contract=off and correctly
 * rounded division and square root, 1/sqrt(x) as an IEEE sqrt followed by an
 * IEEE divide, as in QuadRay's full-precision builds
 * (core/config/rtconf.h:166-211).
 *
 * Scene tables are read through the constant address space, so a read at a
 * wave-uniform index (the list walk, clip lists, lights) is a scalar load
 * into SGPRs and never occupies VGPRs or LDS.
 */
#ifndef QR_TRACER_DEV_H
#define QR_TRACER_DEV_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qr_soa.h"
#include "tracer_kernels.h"

namespace {

constexpr uint32_t SIGN = 0x80000000u;
constexpr int FLAG_PASS = 2;
constexpr int FLAG_SHAD = 4;
constexpr int MAX_STACK = 16;

struct V3 { float x, y, z; };

__device__ __forceinline__ float bitsf(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float fxor(float v, uint32_t m) { return bitsf(fbits(v) ^ m); }
__device__ __forceinline__ float negif(float v, int n) { return n ? -v : v; }
__device__ __forceinline__ float rsq(float x) { return 1.0f / sqrtf(x); }

/* Component access at a run-time index.  pick() selects values, not
 * references: a branch between two fields is merged by the optimiser into
 * one access at a computed address, which gives the walk registers an
 * address and moves them to scratch memory. */
__device__ __forceinline__ float sel(const V3 &v, int i)
{
    return i == 0 ? v.x : (i == 1 ? v.y : v.z);
}
__device__ __forceinline__ V3 pick(bool c, V3 a, V3 b)
{
    return V3{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z};
}
__device__ __forceinline__ float sel6(const V3 &a, const V3 &b, int i)
{
    return i < 3 ? sel(a, i) : sel(b, i - 3);
}
__device__ __forceinline__ void put(V3 &v, int i, float f)
{
    if (i == 0) v.x = f; else if (i == 1) v.y = f; else v.z = f;
}
__device__ __forceinline__ void put6(V3 &a, V3 &b, int i, float f)
{
    if (i < 3) put(a, i, f); else put(b, i - 3, f);
}

/* x86 float->int conversion semantics: out of range and NaN give INT_MIN */
__device__ __forceinline__ int32_t cvt_floor(float x)
{
    float f = floorf(x);
    return (f >= -2147483648.0f && f < 2147483648.0f) ? (int32_t)f : (int32_t)SIGN;
}
__device__ __forceinline__ int32_t cvt_rne(float x)
{
    float f = rintf(x);
    return (f >= -2147483648.0f && f < 2147483648.0f) ? (int32_t)f : (int32_t)SIGN;
}

/* Scene tables are read through the constant address space: a read at a
 * wave-uniform index (list walk, clip lists, lights) becomes a scalar load
 * into SGPRs, a read at a per-lane index a vector load. */
#define QC __attribute__((address_space(4)))

struct Scene
{
    const QC qr_surface  *srf;
    const QC qr_material *mat;
    const QC qr_light    *lgt;
    const QC qr_clip     *clp;
    const QC qr_elem     *elem;
    const QC uint32_t    *texel;
    const QC qr_dev_step *step;         /* nearest-hit program (ends in an end step) */
    const QC qr_dev_step *step_any;     /* any-hit program (ends in an end step) */
    int n_elem, n_lgt;
    float ref_x, ref_y, ref_z;          /* reference point of the world boxes */
};

/* kernel view of the device tables */
inline Scene scene_of(const qr_device_tables *t)
{
    Scene sc;
    sc.srf = (const QC qr_surface *)t->surfaces;
    sc.mat = (const QC qr_material *)t->materials;
    sc.lgt = (const QC qr_light *)t->lights;
    sc.clp = (const QC qr_clip *)t->clips;
    sc.elem = (const QC qr_elem *)t->elems;
    sc.texel = (const QC uint32_t *)t->texels;
    sc.step = (const QC qr_dev_step *)t->steps;
    sc.step_any = (const QC qr_dev_step *)t->steps_any;
    sc.n_elem = t->n_elem;
    sc.n_lgt = t->n_lgt;
    sc.ref_x = t->cull_ref[0];
    sc.ref_y = t->cull_ref[1];
    sc.ref_z = t->cull_ref[2];
    return sc;
}

struct Ray
{
    V3    org, dir;
    float t_min, t_max;
    int   flags;        /* side | pass (| shadow) of the originating hit */
    int   origin;       /* originating surface record or -1 */
    V3    oloc;         /* origin's local hit */
    int   depth;
};

struct Hit
{
    int   srf, side;
    float t;
    V3    hit, loc;
    float u, v;
};

/* inverse-matrix product, diagonal first (tracer.cpp:2536-2568) */
__device__ __forceinline__ V3 xform(const QC qr_surface &s, const V3 &d)
{
    V3 o;
    o.x = s.tc[0][0] * d.x;
    o.y = s.tc[1][1] * d.y;
    o.z = s.tc[2][2] * d.z;
    if (s.trm != 1)
    {
        o.x = o.x + s.tc[0][1] * d.y;
        o.x = o.x + s.tc[0][2] * d.z;
        o.y = o.y + s.tc[1][0] * d.x;
        o.y = o.y + s.tc[1][2] * d.z;
        o.z = o.z + s.tc[2][0] * d.x;
        o.z = o.z + s.tc[2][1] * d.y;
    }
    return o;
}

template <class P>
__device__ __forceinline__ V3 sub3(const V3 &a, P p)
{
    return V3{a.x - p[0], a.y - p[1], a.z - p[2]};
}

/* list-walk registers mirroring QuadRay's context fields */
struct Walk
{
    V3  dff_x, dff_i;   /* DFF_X/Y/Z, DFF_I/J/K */
    V3  ray_i;          /* RAY_I/J/K */
    int cache_last;     /* trnode: last element of the group, -1 off */
    int cache_first;    /* the group's transform-node element */
};

/* hit point and local hit of a root that passed clipping */
struct ClipOut
{
    V3 hit;             /* HIT */
    V3 loc;             /* NEW: the local hit in the surface's tracing frame */
};

/* clipper evaluation (PL_clp 5696, QD_clp 6292, TP_clp 6466) */
__device__ __forceinline__ bool clip_keep(const QC qr_surface &c, const V3 &nx,
                                          const V3 &ni, int rel)
{
    float f;
    if (c.clp_kind == 1)
    {
        int k = c.amap[2] + 3 * c.shift;
        f = negif(sel6(nx, ni, k), c.asgn[2]);
    }
    else
    {
        const V3 p = pick(c.shift, ni, nx);
        float d0 = (p.x * p.x) * c.sci[0];
        float d1 = (p.y * p.y) * c.sci[1];
        float d2 = (p.z * p.z) * c.sci[2];
        if (c.clp_kind == 2)
        {
            d0 = d0 - (c.scj[0] + c.scj[0]) * p.x;
            d1 = d1 - (c.scj[1] + c.scj[1]) * p.y;
            d2 = d2 - (c.scj[2] + c.scj[2]) * p.z;
        }
        f = d0 - c.sci[3];
        f = f + d1;
        f = f + d2;
    }
    return rel < 0 ? (f >= 0.0f) : (f <= 0.0f);
}

/* depth test, local hit, apex snap, axis and custom clipping of one root
 * (CC_clp, tracer.cpp:2686-3236).
 *
 * QuadRay keeps two local-hit registers, NEW_X (world-aligned, from HIT) and
 * NEW_I (transformed, from RAY_I and DFF_I): a surface with a transform
 * (trm) refreshes NEW_I, one without refreshes NEW_X, and clipping reads
 * NEW_I when the surface is traced in its transform node's frame (shift),
 * NEW_X otherwise.  The engine only emits surfaces whose trm and shift
 * agree (a transformed surface always has a transform node, and a surface
 * under one always inherits its rotation), so the register read is always
 * the one refreshed for that root: one local hit, "loc". */
__device__ bool clip_root(const Scene &sc, const Walk &w, const Ray &r, float t_buf,
                          bool tie, const QC qr_surface &s, float t, int side,
                          bool dmask, uint32_t amask, ClipOut *out = nullptr)
{
    /* the depth test is strict, but a nearest-hit walk takes a root equal to
     * the kept one (tie): QuadRay walks camera rays and the secondary rays of
     * reflecting and transparent surfaces over lists built in the reverse of
     * the global list's order (core/engine/engine.cpp:2180-2420), so of two
     * roots at the same distance the later one of the global list wins */
    if (!(t_buf > t || (tie && t_buf == t)) || !(r.t_min < t)) return false;

    V3 hit;
    hit.x = r.dir.x * t + r.org.x;
    hit.y = r.dir.y * t + r.org.y;
    hit.z = r.dir.z * t + r.org.z;
    const V3 ni = {w.ray_i.x * t + w.dff_i.x, w.ray_i.y * t + w.dff_i.y,
                   w.ray_i.z * t + w.dff_i.z};
    V3 loc = pick(s.shift != 0, ni, sub3(hit, s.pos));
    const int sh = 3 * s.shift;

    if (s.apex != 0 && dmask)
    {
        const int ai = s.amap[0], aj = s.amap[1], ak = s.amap[2];
        const int mi = ai + sh, mj = aj + sh, mk = ak + sh;
        float li = sel(loc, ai);
        float rr = li * li;
        if (s.apex != 2)
        {
            float lj = sel(loc, aj);
            rr = rr + lj * lj;
        }
        float lk = sel(loc, ak);
        rr = rr + lk * lk;
        if (rr < s.t_eps)
        {
            float ci = bitsf((fbits(sel6(w.dff_x, w.dff_i, mi)) & SIGN) ^ fbits(1.0f));
            float cj = 0.0f;
            float sk = s.sci[s.amap[0]];
            float lr = 1.0f;
            if (s.apex != 2)
            {
                cj = bitsf((fbits(sel6(w.dff_x, w.dff_i, mj)) & SIGN) ^ fbits(1.0f));
                sk = sk + s.sci[s.amap[1]];
                lr = lr + 1.0f;
            }
            sk = sk / s.sci[s.amap[2]];
            sk = fxor(sk, SIGN);
            float ck = sqrtf(sk);
            float lc2 = sk + lr;
            float inv = rsq(lc2);
            inv = inv * s.t_eps;
            ci = ci * inv;
            cj = cj * inv;
            ck = ck * inv;
            uint32_t ts = side ? SIGN : 0u;
            ck = fxor(ck, fbits(sel6(w.dff_x, w.dff_i, mk)) & SIGN);
            ck = fxor(ck, (ts & amask) ^ amask);
            uint32_t tsg = (ts | amask) ^ amask;
            ci = fxor(ci, tsg);
            cj = fxor(cj, tsg);
            put(loc, ai, ci);
            if (s.apex != 2) put(loc, aj, cj);
            put(loc, ak, ck);
        }
    }

    const V3 l = loc;
    bool ok = true;
    if ((s.cmin_on & 1) && !(s.cmin[0] <= l.x)) ok = false;
    if ((s.cmin_on & 2) && !(s.cmin[1] <= l.y)) ok = false;
    if ((s.cmin_on & 4) && !(s.cmin[2] <= l.z)) ok = false;
    if ((s.cmax_on & 1) && !(s.cmax[0] >= l.x)) ok = false;
    if ((s.cmax_on & 2) && !(s.cmax[1] >= l.y)) ok = false;
    if ((s.cmax_on & 4) && !(s.cmax[2] >= l.z)) ok = false;

    if (s.clip_cnt > 0)
    {
        bool saved = false;
        int cache = -1;
        V3 nx = {0, 0, 0}, ni = {0, 0, 0};
        for (int k = 0; k < s.clip_cnt; k++)
        {
            const QC qr_clip &e = sc.clp[s.clip_beg + k];
            if (e.kind == QR_CLIP_ENTER) { saved = ok; ok = true; continue; }
            if (e.kind == QR_CLIP_LEAVE) { ok = !ok && saved; continue; }
            const QC qr_surface &c = sc.srf[e.surf];
            if (e.kind == QR_CLIP_TRNODE)
            {
                if (e.surf == s.trnode)
                {
                    nx.x = loc.x + s.pos[0];
                    nx.y = loc.y + s.pos[1];
                    nx.z = loc.z + s.pos[2];
                }
                else
                {
                    nx = xform(c, sub3(hit, c.pos));
                }
                cache = e.rel;
                continue;
            }
            if (cache >= 0)
            {
                ni = sub3(nx, c.pos);
                if (k == cache) cache = -1;
            }
            else
            {
                nx = sub3(hit, c.pos);
                if (c.trm != 0) ni = xform(c, nx);
            }
            ok = ok && clip_keep(c, nx, ni, e.rel);
        }
    }
    if (out != nullptr)
    {
        out->hit = hit;
        out->loc = loc;
    }
    return ok;
}

__device__ __forceinline__ bool shadows(uint32_t p)
{
    if (p & QR_PROP_LIGHT) return false;
    if ((p & QR_PROP_TRANSP) && !(p & QR_PROP_REFRACT)) return false;
    return true;
}

/* walk the surface list: nearest hit, or any shadowing hit (SHADOW) */
/* transform set-up of list element e (surface record s; same: it is the
 * ray's origin surface) for ray r: the
 * DFF_X / DFF_I / RAY_I registers as QuadRay's list walk leaves them
 * (OO_loc, tracer.cpp:2466-2684); "last" opens a transform-node cache */
__device__ __forceinline__ void elem_setup(Walk &w, const Ray &r,
                                           const QC qr_surface &s, bool same,
                                           int e, int last)
{
    if (same)
    {
        if (s.shift) w.dff_i = r.oloc; else w.dff_x = r.oloc;
    }

    if (s.tag >= 0 && w.cache_last >= 0)
    {
        if (!same) w.dff_i = sub3(w.dff_x, s.pos);
        if (e == w.cache_last) w.cache_last = -1;
    }
    else
    {
        bool do_ray = same;
        if (!same)
        {
            V3 d = sub3(r.org, s.pos);
            w.dff_x = d;
            if (s.trm != 0)
            {
                V3 t3 = xform(s, d);
                if (s.tag < 0)
                {
                    w.dff_x = t3;
                    w.cache_last = last;
                    w.cache_first = e;
                }
                else
                {
                    w.dff_i = t3;
                }
                do_ray = true;
            }
        }
        if (do_ray) w.ray_i = xform(s, r.dir);
    }
}

/* roots of a quadric or two-plane surface in QuadRay's operation order */
struct Roots
{
    float    t1, t2;
    bool     t1m, t2m;      /* root defined (non-zero denominator) */
    bool     dmask;         /* near-zero discriminant */
    uint32_t amask;         /* sign of a */
    int      first;         /* side tested first */
};

/* discriminant and roots (QD_rt1 5722-6206, TP_rt1 6337-6413); false when
 * the discriminant is negative */
__device__ __forceinline__ bool solve(const QC qr_surface &s, const V3 &rr,
                                      const V3 &dd, Roots &q)
{
    float a_val, b_val, c_val, d_val;
    if (s.solver == QR_SOLVE_QUADRIC)
    {
        float ax, bx, cx, ay, by, cy, az, bz, cz;
        {
            float sri = s.sci[0] * rr.x;
            float sdi = s.sci[0] * dd.x;
            sdi = sdi - s.scj[0];
            ax = rr.x * sri;
            bx = rr.x * sdi;
            sdi = sdi - s.scj[0];
            cx = dd.x * sdi;
        }
        {
            float sri = s.sci[1] * rr.y;
            float sdi = s.sci[1] * dd.y;
            sdi = sdi - s.scj[1];
            ay = rr.y * sri;
            by = rr.y * sdi;
            sdi = sdi - s.scj[1];
            cy = dd.y * sdi;
        }
        {
            float sri = s.sci[2] * rr.z;
            float sdi = s.sci[2] * dd.z;
            sdi = sdi - s.scj[2];
            az = rr.z * sri;
            bz = rr.z * sdi;
            sdi = sdi - s.scj[2];
            cz = dd.z * sdi;
        }
        a_val = ax + ay;
        b_val = bx + by;
        c_val = cx + cy;
        a_val = a_val + az;
        b_val = b_val + bz;
        c_val = c_val + cz;
        c_val = c_val - s.sci[3];
        d_val = b_val * b_val - c_val * a_val;
    }
    else
    {
        const int mi = s.amap[0], mk = s.amap[2];
        float ri = sel(rr, mi), di = sel(dd, mi);
        float rk = sel(rr, mk), dk = sel(dd, mk);
        float si_ = s.sci[mi], sk_ = s.sci[mk];
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
    }

    if (!(0.0f <= d_val)) return false;
    float nb = -b_val;
    q.dmask = d_val < s.d_eps;
    float sd = sqrtf(d_val);
    sd = fxor(sd, fbits(nb) & SIGN);
    float qq = nb + sd;
    const bool m_pos = 0.0f <= sd;
    const bool m_neg = 0.0f > sd;
    float t1n = m_neg ? qq : (m_pos ? c_val : 0.0f);
    float t1d = m_neg ? a_val : (m_pos ? qq : 0.0f);
    float t2n = m_neg ? c_val : (m_pos ? qq : 0.0f);
    float t2d = m_neg ? qq : (m_pos ? a_val : 0.0f);
    q.amask = fbits(a_val) & SIGN;

    if (q.dmask)
    {
        if (t1n == 0.0f) t1d = 1.0f;
        if (t2n == 0.0f) t2d = 1.0f;
        float t1 = t1n / t1d;
        float t2 = t2n / t2d;
        q.t1m = t1d != 0.0f;
        q.t2m = t2d != 0.0f;
        float tdf = t1 - t2;
        tdf = fxor(tdf, q.amask);
        bool fm = 0.0f <= tdf;
        if (!fm) tdf = 0.0f;
        float te = fm ? s.t_eps : 0.0f;
        te = te * t1;
        te = fabsf(te);
        tdf = tdf * -0.5f;
        tdf = tdf - te;
        tdf = fxor(tdf, q.amask);
        if (!(q.t1m && q.t2m)) tdf = 0.0f;
        q.t1 = t1 + tdf;
        q.t2 = t2 - tdf;
    }
    else
    {
        q.t1 = t1n / t1d;
        q.t2 = t2n / t2d;
        q.t1m = t1d != 0.0f;
        q.t2m = t2d != 0.0f;
    }
    q.first = (0.0f > a_val) ? 1 : 0;
    return true;
}

/* reciprocal of a direction component for the slab tests: approximate, and
 * clamped to +-1e30 so that a zero component gives finite slab distances
 * (never 0 * inf) */
__device__ __forceinline__ float slab_rcp(float d)
{
    return fminf(fmaxf(__builtin_amdgcn_rcpf(d), -1e30f), 1e30f);
}

/* does any active lane of the wave hold b (the lane mask stays in scalar
 * registers) */
__device__ __forceinline__ bool wave_any(bool b)
{
    return __builtin_amdgcn_ballot_w64(b) != 0;
}

/* a ray's side of the box tests: its clamped reciprocal direction and, per
 * axis, the offsets that place a box's bounds on the ray, with the ray's
 * share of the allowance (QR_CULL_FUZZ times the L1 distance of its origin
 * o from the boxes' reference point) moving each bound outwards */
/* the step at byte offset "off" of a culling program (wave-uniform: one
 * 32-byte scalar load) */
__device__ __forceinline__ qr_dev_step step_at(const QC qr_dev_step *prog, uint32_t off)
{
    qr_dev_step v;
    __builtin_memcpy(&v, (const QC qr_dev_step *)((const QC char *)prog + off), sizeof(v));
    return v;
}

struct BoxRay
{
    float ix, iy, iz;
    float lx, ly, lz;       /* -(o + m) * i, for the lower bounds */
    float hx, hy, hz;       /* -(o - m) * i, for the upper bounds */
};

__device__ __forceinline__ BoxRay box_ray(const V3 &o, const V3 &dir, float rx, float ry,
                                          float rz)
{
    BoxRay b;
    b.ix = slab_rcp(dir.x);
    b.iy = slab_rcp(dir.y);
    b.iz = slab_rcp(dir.z);
    const float m = QR_CULL_FUZZ * (fabsf(o.x - rx) + fabsf(o.y - ry) + fabsf(o.z - rz));
    b.lx = -((o.x + m) * b.ix);
    b.ly = -((o.y + m) * b.iy);
    b.lz = -((o.z + m) * b.iz);
    b.hx = -((o.x - m) * b.ix);
    b.hy = -((o.y - m) * b.iy);
    b.hz = -((o.z - m) * b.iz);
    return b;
}

/* can the segment (t_min, t_buf) of the ray meet the box?  Slab test: the
 * ray's entry into the box must come no later than its exit and overlap
 * the segment.  The slab distances are formed with fused multiply-adds: the
 * test only decides what may be skipped, and the allowance is far above
 * their rounding. */
__device__ __forceinline__ bool box_pass(const qr_dev_step &st, const BoxRay &b,
                                         float t_min, float t_buf)
{
    const float x0 = __builtin_fmaf(st.lo[0], b.ix, b.lx);
    const float x1 = __builtin_fmaf(st.hi[0], b.ix, b.hx);
    const float y0 = __builtin_fmaf(st.lo[1], b.iy, b.ly);
    const float y1 = __builtin_fmaf(st.hi[1], b.iy, b.hy);
    const float z0 = __builtin_fmaf(st.lo[2], b.iz, b.lz);
    const float z1 = __builtin_fmaf(st.hi[2], b.iz, b.hz);
    const float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fminf(z0, z1));
    const float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1));
    return fmaxf(tn, t_min) <= fminf(tf, t_buf);
}

/* material properties of one side of a surface at a per-lane side: both
 * read as wave-uniform scalars, then selected (an indexed read would be a
 * vector memory load) */
__device__ __forceinline__ uint32_t side_props(const QC qr_surface &s, int side)
{
    const uint32_t p0 = __builtin_amdgcn_readfirstlane(s.props[0]);
    const uint32_t p1 = __builtin_amdgcn_readfirstlane(s.props[1]);
    return side ? p1 : p0;
}

/* one root of a quadric (side 0: t1, side 1: t2) through the origin-side
 * rule, the depth test and clipping; true when it is accepted */
__device__ __forceinline__ bool try_root(const Scene &sc, Walk &w, const Ray &r,
                                         float t_buf, bool tie, const QC qr_surface &s,
                                         const Roots &q, int side, bool same)
{
    const int fl = r.flags & 3;
    if (same && (fl == 1 - side || fl == 2 + side)) return false;
    const float t = side == 0 ? q.t1 : q.t2;
    const bool tm = side == 0 ? q.t1m : q.t2m;
    if (!tm) return false;
    return clip_root(sc, w, r, t_buf, tie, s, t, side, q.dmask, q.amask);
}

/* whether QuadRay gives a surface a nearest-hit list of its own for the
 * rays leaving it (reflecting or not opaque on either side,
 * core/engine/engine.cpp:2180-2201), built in the reverse of the global
 * list's order; other surfaces' rays walk the global list itself */
__device__ __forceinline__ bool own_list(const QC qr_surface &s)
{
    const uint32_t p0 = s.props[0], p1 = s.props[1];
    return ((p0 | p1) & QR_PROP_REFLECT) != 0 || (p0 & p1 & QR_PROP_OPAQUE) == 0;
}

/* the roots of list element e (surface s, set up in w) for ray r that the
 * walk accepts: true with the accepted side and distance.  A tie goes by
 * list order ("rev": the later element wins), or with SPATIAL by element
 * index against best_e. */
template <bool SHADOW, bool SPATIAL>
__device__ __forceinline__ bool elem_roots(const Scene &sc, Walk &w, const Ray &r,
                                           const QC qr_surface &s, bool same, int e,
                                           const Hit &best, int best_e, bool rev, float t_buf,
                                           int &side, float &t)
{
    const int sh = 3 * s.shift;
    bool acc = false;
    if (s.solver == QR_SOLVE_PLANE)
    {
        const int k = s.amap[2] + sh;
        float dff_k = negif(sel6(w.dff_x, w.dff_i, k), s.asgn[2]);
        float ray_k = negif(k < 3 ? sel(r.dir, k) : sel(w.ray_i, k - 3), s.asgn[2]);
        dff_k = -dff_k;
        if (!same && ray_k != 0.0f)
        {
            t = dff_k / ray_k;
            side = ray_k < 0.0f ? 0 : 1;
            const bool tie = !SHADOW && best.srf >= 0 &&
                             (SPATIAL ? (rev ? e > best_e : e < best_e) : rev);
            acc = clip_root(sc, w, r, t_buf, tie, s, t, side, false, 0u);
        }
    }
    else
    {
        Roots q;
        if (solve(s, pick(s.shift, w.ray_i, r.dir), pick(s.shift, w.dff_i, w.dff_x), q))
        {
            const bool tie = !SHADOW && best.srf >= 0 &&
                             (SPATIAL ? (rev ? e > best_e : e < best_e) : rev);
            side = q.first;
            acc = try_root(sc, w, r, t_buf, tie, s, q, side, same);
            if (!acc)
            {
                side = 1 - side;
                acc = try_root(sc, w, r, t_buf, tie, s, q, side, same);
            }
            t = side == 0 ? q.t1 : q.t2;
        }
    }
    return acc;
}

/* most work-items in a workgroup of a kernel that walks the surface list */
#define QR_WALK_MAX_WG 256

/* walk the surface list through the culling program: nearest accepted root
 * (h.srf, h.side, h.t), or any shadowing root (SHADOW).  Accepted roots follow
 * QuadRay's overwrite rules: the second root of a surface is skipped once the
 * first is accepted, and of roots at the same distance the first in list
 * order wins, or with "rev" the last: camera rays, and the rays leaving a
 * surface that has its own list (own_list: every secondary ray of the
 * Whitted tracer), walk lists that QuadRay builds in the reverse of this
 * one's order.
 *
 * The loop has one exit, taken by the whole wave: a lane whose shadow test
 * is decided stops keeping boxes open, and the wave leaves once every lane
 * is decided (a lane leaving on its own would split the loop's exit and cost
 * every iteration a per-lane state dispatch). */
template <bool SHADOW, bool SPATIAL = false>
__device__ bool walk(const Scene &sc, const Ray &r, Hit &best, bool rev = true)
{
    Walk w;
    w.cache_last = -1;
    w.cache_first = 0;
    w.dff_x = V3{0, 0, 0};
    w.dff_i = V3{0, 0, 0};
    w.ray_i = V3{0, 0, 0};
    float t_buf = r.t_max;
    best.srf = -1;
    bool shadowed = false;
    /* lanes that still keep boxes open (any-hit walks: not yet shadowed),
     * as a wave mask that the box tests' ballots are masked with */
    uint64_t live = ~0ull;

    /* the ray's side of the world boxes (the reciprocal direction is
     * approximate: the boxes carry an allowance far above its error) */
    const BoxRay br = box_ray(r.org, r.dir, sc.ref_x, sc.ref_y, sc.ref_z);

    /* nearest-hit walks over the spatial (any-hit) program visit units out
     * of list order: exact ties then go by element index (best_e) */
    const QC qr_dev_step *prog = SHADOW || SPATIAL ? sc.step_any : sc.step;
    int best_e = -1;
    /* the ray's side of the open transform-node group's member boxes, in the
     * node's frame: kept in LDS, a slot per work-item, not in registers the
     * whole loop would hold.  A member's step is reached only through its
     * node's step: the upload gives a member step that kind only in a
     * program where that holds. */
    __shared__ float s_gb[9][QR_WALK_MAX_WG];
    const unsigned tid = threadIdx.x;
    /* the byte offset of the ray's origin surface record (none: matches no
     * record) */
    const uint32_t org_off = r.origin >= 0 ? (uint32_t)r.origin * (uint32_t)sizeof(qr_surface)
                                           : 0xffffffffu;
    /* two loops: the inner one runs the box steps up to the next element
     * step and touches no walk register, so the element work is the only
     * path back to the outer loop's head (the walk registers then need no
     * copies where paths merge).  The step is addressed by its byte offset
     * and a box step's skip is one, and the program's end step stops the
     * walk, so a box step costs its test, one select and one load. */
    for (uint32_t off = 0;; off += QR_DEV_STEP_BYTES)
    {
        /* a step is one 32-byte scalar load (box, element and skip
         * together), so a box step waits for one load, not three in a row */
        qr_dev_step st = step_at(prog, off);
        while (st.elem < 0)
        {
            const uint64_t near = __builtin_amdgcn_ballot_w64(box_pass(st, br, r.t_min, t_buf));
            off = (near & live) != 0 ? off + QR_DEV_STEP_BYTES : (uint32_t)st.skip;
            st = step_at(prog, off);
        }
        if (st.elem == QR_DEV_STEP_END) break;
        const int e = st.elem;
        const int kind = qr_dev_step_local(&st);
        /* a member of the open group: its box in the group's frame, before
         * anything of the surface is loaded (a member's skipped set-up
         * changes nothing for the elements after it) */
        /* the loop body has one path back to its head (no continue): the
         * element is tested only while some live lane keeps it */
        uint64_t keep = live;
        if (kind == QR_LOCAL_GROUP)
        {
            BoxRay gb;
            gb.ix = s_gb[0][tid];
            gb.iy = s_gb[1][tid];
            gb.iz = s_gb[2][tid];
            gb.lx = s_gb[3][tid];
            gb.ly = s_gb[4][tid];
            gb.lz = s_gb[5][tid];
            gb.hx = s_gb[6][tid];
            gb.hy = s_gb[7][tid];
            gb.hz = s_gb[8][tid];
            keep &= __builtin_amdgcn_ballot_w64(box_pass(st, gb, r.t_min, t_buf));
        }
        if (keep != 0)
        {
            /* the cache holds only inside its group, whose last element may
             * have been skipped (the any-hit program also visits groups out
             * of list order) */
            if (w.cache_last >= 0 && (e > w.cache_last || e < w.cache_first))
            {
                w.cache_last = -1;
            }
            /* the record's byte offset (the index only for a kept hit) */
            const uint32_t soff = (uint32_t)st.skip & ~3u;
            const QC qr_surface &s = *(const QC qr_surface *)((const QC char *)sc.srf + soff);
            const bool same = soff == org_off;
            /* a transform node's group end (only a node's step has one;
             * other surfaces ignore it) */
            const int last = __float_as_int(st.lo[0]);
            elem_setup(w, r, s, same, e, last);
            if (s.tag < 0 && s.trm != 0)
            {
                /* a transform node (never a ray's origin) opened its group:
                 * the ray in the node's frame, about the node's origin */
                const BoxRay gb = box_ray(w.dff_x, w.ray_i, 0.0f, 0.0f, 0.0f);
                s_gb[0][tid] = gb.ix;
                s_gb[1][tid] = gb.iy;
                s_gb[2][tid] = gb.iz;
                s_gb[3][tid] = gb.lx;
                s_gb[4][tid] = gb.ly;
                s_gb[5][tid] = gb.lz;
                s_gb[6][tid] = gb.hx;
                s_gb[7][tid] = gb.hy;
                s_gb[8][tid] = gb.hz;
            }
            if (s.solver != QR_SOLVE_NONE)
            {
                if (kind == QR_LOCAL_OWN)
                {
                    /* the surface's own box, in its tracing frame (the
                     * registers the roots are computed from), whose
                     * reference point is the surface's origin */
                    const BoxRay bl = box_ray(w.dff_i, w.ray_i, 0.0f, 0.0f, 0.0f);
                    keep &= __builtin_amdgcn_ballot_w64(box_pass(st, bl, r.t_min, t_buf));
                }
                if (keep != 0)
                {
                    int side = 0;
                    float t = 0.0f;
                    const bool acc = elem_roots<SHADOW, SPATIAL>(sc, w, r, s, same, e, best,
                                                                 best_e, rev, t_buf, side, t);
                    if (SHADOW)
                    {
                        shadowed = shadowed || (acc && shadows(side_props(s, side)));
                        live = __builtin_amdgcn_ballot_w64(!shadowed);
                        if (live == 0) break;
                    }
                    else if (acc)
                    {
                        t_buf = t;
                        best_e = e;
                        best.srf = (int)(soff / (uint32_t)sizeof(qr_surface));
                        best.side = side;
                        best.t = t;
                    }
                }
            }
        }
    }
    return SHADOW ? shadowed : best.srf >= 0;
}

/* hit point, local hit and plane texture coordinates of an accepted root,
 * recomputed with the operations the walk applied to it: the element set-up
 * (after its transform node's, for a grouped surface), the roots and the
 * clipping stage of that surface */
__device__ void hit_info(const Scene &sc, const Ray &r, Hit &h)
{
    Walk w;
    w.cache_last = -1;
    w.cache_first = 0;
    w.dff_x = V3{0, 0, 0};
    w.dff_i = V3{0, 0, 0};
    w.ray_i = V3{0, 0, 0};
    const QC qr_surface &s = sc.srf[h.srf];
    if (s.tag >= 0 && s.trnode >= 0 && s.trnode != h.srf)
    {
        const QC qr_surface &tn = sc.srf[s.trnode];
        if (tn.tag < 0 && tn.trm != 0)
        {
            elem_setup(w, r, tn, s.trnode == r.origin, -2, 0x7fffffff);
        }
    }
    elem_setup(w, r, s, h.srf == r.origin, -1, -1);
    bool dmask = false;
    uint32_t amask = 0u;
    if (s.solver != QR_SOLVE_PLANE)
    {
        Roots q;
        solve(s, pick(s.shift, w.ray_i, r.dir), pick(s.shift, w.dff_i, w.dff_x), q);
        dmask = q.dmask;
        amask = q.amask;
    }
    ClipOut co;
    co.hit = V3{0, 0, 0};
    co.loc = V3{0, 0, 0};
    clip_root(sc, w, r, 3.402823466e+38f, false, s, h.t, h.side, dmask, amask, &co);
    h.hit = co.hit;
    h.loc = co.loc;
    if (s.solver == QR_SOLVE_PLANE)
    {
        h.u = negif(sel(co.loc, s.amap[0]), s.asgn[0]);
        h.v = negif(sel(co.loc, s.amap[1]), s.asgn[1]);
    }
    else
    {
        h.u = 0.0f;
        h.v = 0.0f;
    }
}


/* surface normal in world space (PL 5669-5685, QD 6237-6281, TP 6415-6455,
 * transform 3273-3348) */
__device__ V3 normal_at(const Scene &sc, const QC qr_surface &s, const Hit &h)
{
    V3 n;
    const uint32_t ts = h.side ? SIGN : 0u;
    if (s.nrm_kind == 1)
    {
        V3 z = {0.0f, 0.0f, 0.0f};
        put(z, s.amap[2], negif(fxor(1.0f, ts), s.asgn[2]));
        n = z;
    }
    else
    {
        V3 l;
        l.x = h.loc.x * s.sci[0];
        l.y = h.loc.y * s.sci[1];
        l.z = h.loc.z * s.sci[2];
        if (s.nrm_kind == 2)
        {
            l.x = l.x - s.scj[0];
            l.y = l.y - s.scj[1];
            l.z = l.z - s.scj[2];
        }
        float rr = l.x * l.x + l.y * l.y;
        rr = rr + l.z * l.z;
        float inv = fxor(rsq(rr), ts);
        n = V3{l.x * inv, l.y * inv, l.z * inv};
    }
    if (s.trm == 0) return n;

    const QC qr_surface &t = sc.srf[s.trnode];
    float x = t.tc[0][0] * n.x;
    float y = t.tc[1][1] * n.y;
    float z = t.tc[2][2] * n.z;
    bool renorm = true;
    if (t.trm != 1)
    {
        x = x + t.tc[1][0] * n.y;
        x = x + t.tc[2][0] * n.z;
        y = y + t.tc[0][1] * n.x;
        y = y + t.tc[2][1] * n.z;
        z = z + t.tc[0][2] * n.x;
        z = z + t.tc[1][2] * n.y;
        renorm = t.trm != 2;
    }
    if (renorm)
    {
        float rr = x * x + y * y;
        rr = rr + z * z;
        float inv = rsq(rr);
        x = x * inv; y = y * inv; z = z * inv;
    }
    return V3{x, y, z};
}

/* vector to a light and its dot product with the normal (LT, 4015-4035) */
__device__ __forceinline__ float light_vec(const QC qr_light &l, const V3 &hit,
                                           const V3 &nrm, V3 &L)
{
    L = V3{l.pos[0] - hit.x, l.pos[1] - hit.y, l.pos[2] - hit.z};
    float dot = L.x * nrm.x + L.y * nrm.y;
    dot = dot + L.z * nrm.z;
    return dot;
}

/* fixed point 28.4 specular power (tracer.cpp:4225-4283) */
__device__ float spec_pow(float s, uint32_t pw)
{
    uint32_t fr = pw & 0xF;
    float x2 = s, x4 = s, r = 1.0f;
    while (fr != 0)
    {
        x4 = sqrtf(x4);
        uint32_t bit = 0x8 & fr;
        fr = (fr << 1) & 0xF;
        if (bit) r = r * x4;
    }
    uint32_t ip = pw >> 4;
    if (ip != 0)
    {
        float fp = r;
        r = 1.0f;
        while (ip != 0)
        {
            if (ip & 1) r = r * x2;
            ip >>= 1;
            x2 = x2 * x2;
        }
        r = r * fp;
    }
    return r;
}

/* surface colour at a hit: texel lookup and gamma to linear (MT_tex
 * 3382-3416, PAINT_SIMD 985) */
__device__ __forceinline__ V3 texel_color(const Scene &sc, const QC qr_material &m,
                                          uint32_t props, const Hit &h)
{
    uint32_t idx = 0;
    if (props & QR_PROP_TEXTURE)
    {
        float tx = m.tmap[0] ? h.v : h.u;
        float ty = m.tmap[1] ? h.v : h.u;
        tx = tx - m.xoffs;
        ty = ty - m.yoffs;
        tx = tx * m.xscal;
        ty = ty * m.yscal;
        uint32_t ix = (uint32_t)cvt_floor(tx) & m.xmask;
        uint32_t iy = (uint32_t)cvt_floor(ty) & m.ymask;
        idx = ix + (iy << m.yshft);
    }
    const uint32_t tv = sc.texel[m.tex_off + idx];
    V3 tex;
    tex.x = (float)(int32_t)((tv >> 16) & 0xFF) / 255.0f;
    tex.y = (float)(int32_t)((tv >> 8) & 0xFF) / 255.0f;
    tex.z = (float)(int32_t)(tv & 0xFF) / 255.0f;
    if (props & QR_PROP_GAMMA)
    {
        tex.x = tex.x * tex.x;
        tex.y = tex.y * tex.y;
        tex.z = tex.z * tex.z;
    }

    return tex;
}

/* refraction and reflection set-up of a hit: which children it spawns,
 * their directions, and the transmitted / reflected fractions with Fresnel
 * and total inner reflection (TR 4442-4936, RF 4944-5357) */
struct TrRf
{
    bool  tr, rf;
    bool  fresnel_tr;       /* Fresnel computed in the transparency block */
    float c_trn, c_rfl;
    V3    tdir, rdir;
};

__device__ __forceinline__ TrRf trrf_setup(const QC qr_material &m, uint32_t props,
                                           const V3 &rdir_in, const V3 &nrm)
{
    TrRf t;
    t.tr = false;
    t.rf = false;
    t.fresnel_tr = !(props & QR_PROP_OPAQUE) && (props & QR_PROP_FRESNEL);
    t.tdir = V3{0.0f, 0.0f, 0.0f};
    t.rdir = V3{0.0f, 0.0f, 0.0f};
    const V3 &rd = rdir_in;
    /* refraction with Fresnel (4442-4936) */
    float c_trn = m.c_trn;
    float c_rfl = m.c_rfl;
    if (!(props & QR_PROP_OPAQUE))
    {
        bool m_trn = true;
        V3 nray = rd;
        if ((props & QR_PROP_REFRACT) || (props & QR_PROP_FRESNEL))
        {
            float r2 = rd.x * rd.x;
            r2 = r2 + rd.y * rd.y;
            r2 = r2 + rd.z * rd.z;
            float inv = rsq(r2);
            V3 rn = {rd.x * inv, rd.y * inv, rd.z * inv};
            float dot = rn.x * nrm.x;
            dot = dot + rn.y * nrm.y;
            dot = dot + rn.z * nrm.z;
            float x0 = dot * m.c_rfr;
            float k = x0 * x0;
            k = k + 1.0f;
            k = k - m.rfr_2;
            if ((props & QR_PROP_FRESNEL) && !(0.0f <= k))
            {
                m_trn = false;
                c_trn = 0.0f;
                c_rfl = m.c_rfl + m.c_trn;
            }
            else
            {
                float sk = sqrtf(k);
                x0 = x0 + sk;
                if (props & QR_PROP_REFRACT)
                {
                    float v;
                    v = nrm.x * x0; nray.x = rn.x * m.c_rfr; nray.x = nray.x - v;
                    v = nrm.y * x0; nray.y = rn.y * m.c_rfr; nray.y = nray.y - v;
                    v = nrm.z * x0; nray.z = rn.z * m.c_rfr; nray.z = nray.z - v;
                }
                if (props & QR_PROP_FRESNEL)
                {
                    float n = m.c_rfr;
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
                    f0 = f0 * m.c_trn;
                    c_trn = m.c_trn - f0;
                    c_rfl = m.c_rfl + f0;
                }
            }
        }
        if (m_trn)
        {
            t.tr = true;
            t.tdir = nray;
        }
    }

    /* reflection with Fresnel (4944-5357) */
    if ((props & QR_PROP_REFLECT) ||
        (!(props & QR_PROP_OPAQUE) && (props & QR_PROP_FRESNEL)))
    {
        float r2 = rd.x * rd.x;
        r2 = r2 + rd.y * rd.y;
        r2 = r2 + rd.z * rd.z;
        float inv = rsq(r2);
        V3 rn = {rd.x * inv, rd.y * inv, rd.z * inv};
        float dot = rn.x * nrm.x;
        dot = dot + rn.y * nrm.y;
        dot = dot + rn.z * nrm.z;
        float v;
        V3 nray;
        v = nrm.x * dot; nray.x = rn.x - v; nray.x = nray.x - v;
        v = nrm.y * dot; nray.y = rn.y - v; nray.y = nray.y - v;
        v = nrm.z * dot; nray.z = rn.z - v; nray.z = nray.z - v;

        if ((props & QR_PROP_FRESNEL) && (props & QR_PROP_OPAQUE))
        {
            float f0;
            if (props & QR_PROP_METAL)
            {
                float x6 = m.c_rcp;
                float x4 = dot * x6;
                x4 = x4 + x4;
                float x0 = dot * dot;
                x6 = x6 * x6;
                x6 = x6 + m.ext_2;
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
            }
            else
            {
                float n = m.c_rfr;
                float x0 = dot * n;
                float k = x0 * x0;
                k = k + 1.0f;
                k = k - m.rfr_2;
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
            }
            f0 = f0 - 1.0f;
            f0 = f0 * m.c_rfl;
            c_rfl = m.c_rfl + f0;
        }
        t.rf = true;
        t.rdir = nray;
    }
    t.c_trn = c_trn;
    t.c_rfl = c_rfl;
    return t;
}

/* ARGB8 packing of a resolved colour with optional gamma (1922-1945) */
__device__ __forceinline__ uint32_t pack_argb(const qr_frame_desc &fd, const V3 &c)
{
    const float v[3] = {c.x, c.y, c.z};
    uint32_t px = 0;
    #pragma unroll
    for (int ch = 0; ch < 3; ch++)
    {
        float f = v[ch];
        if (fd.gamma & QR_PROP_GAMMA) f = sqrtf(f);
        f = f * 255.0f;
        px |= ((uint32_t)cvt_rne(f) & 0xFFu) << (16 - 8 * ch);
    }
    return px;
}

/* anti-aliasing resolve of a pixel's clamped samples, halve then pairwise
 * add per pass (7047-7133), and ARGB8 packing */
__device__ __forceinline__ uint32_t resolve_pack(const qr_frame_desc &fd, int ns,
                                                 const V3 &a0, const V3 &a1,
                                                 const V3 &a2, const V3 &a3)
{
    V3 c;
    if (ns == 1)
    {
        c = a0;
    }
    else if (ns == 2)
    {
        c.x = a0.x * 0.5f + a1.x * 0.5f;
        c.y = a0.y * 0.5f + a1.y * 0.5f;
        c.z = a0.z * 0.5f + a1.z * 0.5f;
    }
    else
    {
        float p0, p1;
        p0 = a0.x * 0.5f + a1.x * 0.5f; p1 = a2.x * 0.5f + a3.x * 0.5f;
        c.x = p0 * 0.5f + p1 * 0.5f;
        p0 = a0.y * 0.5f + a1.y * 0.5f; p1 = a2.y * 0.5f + a3.y * 0.5f;
        c.y = p0 * 0.5f + p1 * 0.5f;
        p0 = a0.z * 0.5f + a1.z * 0.5f; p1 = a2.z * 0.5f + a3.z * 0.5f;
        c.z = p0 * 0.5f + p1 * 0.5f;
    }
    return pack_argb(fd, c);
}

/* the same resolve of a pixel's samples held by ns neighbouring lanes
 * (sample k in lane k of an aligned group): halve, then add the partner
 * lane's value per pass; every lane of the wave calls it with the same ns,
 * and each lane of the group gets the pixel */
__device__ __forceinline__ uint32_t resolve_lanes(const qr_frame_desc &fd, int ns, V3 c)
{
    for (int d = 1; d < ns; d <<= 1)
    {
        const V3 h = V3{c.x * 0.5f, c.y * 0.5f, c.z * 0.5f};
        c.x = h.x + __shfl_xor(h.x, d);
        c.y = h.y + __shfl_xor(h.y, d);
        c.z = h.z + __shfl_xor(h.z, d);
    }
    return pack_argb(fd, c);
}

/* a pending node of the Whitted tree */
struct Node
{
    V3    col;          /* colour accumulated so far */
    float c_trn, c_rfl;
    V3    org;          /* hit point: origin of the children */
    V3    tdir, rdir;   /* refraction and reflection directions */
    V3    loc;          /* local hit for the children's self-tests */
    int   srf, side;
    int   depth;
    int   todo;         /* bit 0: refraction pending, bit 1: reflection */
};

/* local shading of a hit (MT, LT): its colour; occluded(light, shadow ray)
 * answers the hard-shadow test of one light */
template <class Occluded>
__device__ V3 shade_color(const Scene &sc, const qr_frame_desc &fd, const Ray &r,
                          const Hit &h, const V3 &nrm, Occluded occluded)
{
    const QC qr_surface &s = sc.srf[h.srf];
    const uint32_t props = s.props[h.side];
    const QC qr_material &m = sc.mat[s.mat[h.side]];

    const V3 tex = texel_color(sc, m, props, h);

    V3 col;
    if (props & QR_PROP_LIGHT)
    {
        col = tex;
    }
    else
    {
        col.x = tex.x * fd.amb[0];
        col.y = tex.y * fd.amb[1];
        col.z = tex.z * fd.amb[2];
        float f = 1.0f - m.c_trn;
        f = f - m.c_rfl;
        if (!(0.0f <= f)) f = 0.0f;
        col.x = col.x * f;
        col.y = col.y * f;
        col.z = col.z * f;

        for (int li = 0; li < sc.n_lgt; li++)
        {
            const QC qr_light &l = sc.lgt[li];
            V3 L;
            const float dot = light_vec(l, h.hit, nrm, L);
            if (!(0.0f < dot)) continue;

            Ray sr;
            sr.org = h.hit;
            sr.dir = L;
            sr.t_min = 0.0f;
            sr.t_max = l.t_max;
            sr.flags = h.side | FLAG_SHAD;
            sr.origin = h.srf;
            sr.oloc = h.loc;
            sr.depth = r.depth - 1;
            if (occluded(li, sr)) continue;

            float r2 = L.x * L.x + L.y * L.y;
            r2 = r2 + L.z * L.z;

            float dff = 0.0f;
            if (props & QR_PROP_DIFFUSE)
            {
                float ir = rsq(r2);
                float r1 = ir * r2;
                float at = r2 * l.a_qdr;
                float lr = r1 * l.a_lnr;
                at = at + l.a_cnt;
                at = at + lr;
                float ia = rsq(at);
                dff = dot * ia;
                dff = dff * ir;
                dff = dff * m.l_dff;
            }

            bool plain_spc = false;
            float spc = 0.0f;
            if (props & QR_PROP_SPECULAR)
            {
                V3 R;
                float v = dot * nrm.x;
                R.x = L.x - v; R.x = R.x - v;
                v = dot * nrm.y;
                R.y = L.y - v; R.y = R.y - v;
                v = dot * nrm.z;
                R.z = L.z - v; R.z = R.z - v;
                float rr2 = r.dir.z * r.dir.z;
                rr2 = rr2 + r.dir.x * r.dir.x;
                rr2 = rr2 + r.dir.y * r.dir.y;
                float sn = R.x * r.dir.x + R.y * r.dir.y;
                sn = sn + R.z * r.dir.z;
                if (0.0f < sn)
                {
                    sn = sn * rsq(rr2);
                    sn = sn * rsq(r2);
                    spc = spec_pow(sn, m.l_pow);
                    spc = spc * m.l_spc;
                    if (props & QR_PROP_METAL) dff = dff + spc;
                    else plain_spc = true;
                }
            }

            V3 lc;
            if (plain_spc)
            {
                float v;
                v = tex.x * dff; v = v * l.col[0]; lc.x = v + l.col[0] * spc;
                v = tex.y * dff; v = v * l.col[1]; lc.y = v + l.col[1] * spc;
                v = tex.z * dff; v = v * l.col[2]; lc.z = v + l.col[2] * spc;
            }
            else
            {
                lc.x = (tex.x * l.col[0]) * dff;
                lc.y = (tex.y * l.col[1]) * dff;
                lc.z = (tex.z * l.col[2]) * dff;
            }
            col.x = lc.x + col.x;
            col.y = lc.y + col.y;
            col.z = lc.z + col.z;
        }
    }
    return col;
}

/* set-up of a hit's children (TR, RF), which does not depend on its
 * lighting: origin, directions, coefficients and which are pending */
__device__ __forceinline__ void node_children(const Scene &sc, const Ray &r, const Hit &h,
                                              const V3 &nrm, Node &nd)
{
    const QC qr_surface &s = sc.srf[h.srf];
    const uint32_t props = s.props[h.side];
    const QC qr_material &m = sc.mat[s.mat[h.side]];
    nd.org = h.hit;
    nd.loc = h.loc;
    nd.srf = h.srf;
    nd.side = h.side;
    nd.depth = r.depth;
    nd.todo = 0;

    const TrRf t = trrf_setup(m, props, r.dir, nrm);
    nd.todo |= (t.tr ? 1 : 0) | (t.rf ? 2 : 0);
    nd.tdir = t.tdir;
    nd.rdir = t.rdir;
    nd.c_trn = t.c_trn;
    nd.c_rfl = t.c_rfl;
    if (r.depth == 0) nd.todo = 0;
}

/* local shading of a hit and set-up of its children (MT, LT, TR, RF) */
template <class Occluded>
__device__ void shade(const Scene &sc, const qr_frame_desc &fd, const Ray &r,
                      const Hit &h, const V3 &nrm, Node &nd, Occluded occluded)
{
    nd.col = shade_color(sc, fd, r, h, nrm, occluded);
    node_children(sc, r, h, nrm, nd);
}

} /* namespace */

#endif /* QR_TRACER_DEV_H */
