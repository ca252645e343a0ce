/*
 * tracer_pt.hip -- QuadRay's path tracer (PT domain of render0) for MI355X,
 * traced as a wavefront: one queue of rays per bounce, compacted between
 * bounces.
 *
 * A launch traces a run of passes over a chunk of the frame (its 8x8-pixel
 * tiles tile0 .. tile0 + ntiles - 1): every (pass, sample slot) pair is one
 * camera sample.  Bounce L is one kernel over the rays of queue L; each ray
 * walks the surface list, adds its hit's emission times its throughput to
 * its sample, and appends its children (reflection, refraction, diffuse) to
 * queue L + 1 with one ballot / popcount prefix per wave and one global
 * atomic per workgroup, so every wave of the next bounce is full and rays
 * of one workgroup stay together.  QuadRay reaches the same coherence by
 * sorting its rays into per-surface SIMD buffers (core/tracer/tracer.h
 * 467-547, tracer.cpp 783-964, 6686-7611).
 *
 * A path is a tree (a diffuse hit and a transparent one may both continue),
 * and QuadRay adds the contributions of a sample in depth-first order: the
 * hit, then its reflection subtree, its refraction subtree, its diffuse
 * subtree.  Each ray carries its position in that order (two bits per
 * bounce: reflection 1, refraction 2, diffuse 3), each non-zero
 * contribution is kept with it in a per-sample list, and after the last
 * bounce one work-item per slot folds its passes in order: the running mean
 * times (1 - 1/n), plus the pass's contributions in depth-first order, in
 * QuadRay's fp32 operations; then the slot's seed steps once per pass
 * (QuadRay keeps the same planes and seed plane per framebuffer,
 * core/engine/engine.cpp 2910-2935, 3719-3756).  Zero contributions are
 * skipped: adding +0 to the non-negative sum changes no bit.
 *
 * A path's random numbers come from its slot's 32-bit LCG in the order
 * QuadRay's buffered tracer draws them along the path: two for the tent
 * jitter of the camera ray, then per hit one for Russian roulette, two for
 * the cosine-distributed diffuse direction and one for the refraction /
 * reflection pick; a diffuse child continues the state after its two draws,
 * refraction and reflection children after all draws of the hit.  Pass p of
 * a launch starts from the slot's seed stepped p times.
 *
 * Queues and lists are preallocated; an append beyond capacity sets a flag
 * and keeps counting, the fold and the resolve then write nothing, and the
 * run is traced again once the buffers have grown to the counted need.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstring>
#include <new>

#include "qr_soa.h"
#include "tracer_dev.h"
#include "tracer_kernels.h"

namespace {

constexpr int PT_TILE = 8;
constexpr int PT_ROULETTE_DEPTH = 5;    /* RT_STACK_DEPTH - 5 (tracer.cpp:3488) */
constexpr int PT_SPLIT_DEPTH = 8;       /* RT_STACK_DEPTH - 2 (tracer.cpp:4690) */
constexpr int PT_MAX_DEPTH = 14;
constexpr int WG = 256;
constexpr int WAVES = WG / 64;
/* samples (slots x passes) per run of passes: a launch of more is split
 * into runs, which bounds the queues' memory (about 100 bytes a sample) */
constexpr uint64_t PT_RUN_SAMPLES = 1ull << 25;

/* counters of a run: rays queued per bounce, contributions, overflow */
constexpr int CNT_CONTRIB = 16;
constexpr int CNT_OVERFLOW = 17;
constexpr int CNT_N = 18;

/* the children of a hit in depth-first order, as digits of the order key */
constexpr uint32_t KEY_RF = 1, KEY_TR = 2, KEY_DF = 3;

struct PtQueue
{
    float4   *a, *b, *c, *d;    /* org|origin, dir|flags, oloc|state, mul|key */
    uint32_t *sid;              /* sample: pass * n0 + chunk slot */
};

struct PtArgs
{
    Scene          sc;
    qr_frame_desc  fd;
    float         *pr, *pg, *pb;        /* colour planes, one value per slot */
    uint32_t      *seeds;
    uint32_t      *frame;
    uint32_t       tile0, ntiles, tiles_x;
    int            ns;
    uint32_t       n0;                  /* slots of the chunk */
    uint32_t       passes;              /* passes of this run */
    float          pts_c0;              /* count of the run's first pass */
    PtQueue        q[2];                /* bounce L reads q[L & 1], appends to
                                           q[(L + 1) & 1] */
    uint32_t       cap_q;
    float4        *ct;                  /* contributions: rgb | order key */
    uint32_t      *ct_next;             /* next contribution of the sample */
    uint32_t       cap_c;
    uint32_t      *heads;               /* first contribution per sample */
    uint32_t      *cnt;                 /* CNT_N counters */
};

/* next number of the slot's LCG as a float in [0, 1): top 24 bits over 2^24
 * (GET_RANDOM_F tracer.cpp:1949-1980, LCG24 tracer.h:58) */
__device__ __forceinline__ float pt_random(uint32_t &st)
{
    st = st * 214013u + 2531011u;
    return (float)((st >> 8) & 0xFFFFFFu) / 16777216.0f;
}

__device__ __forceinline__ uint32_t lcg_steps(uint32_t st, uint32_t n)
{
    for (uint32_t k = 0; k < n; k++) st = st * 214013u + 2531011u;
    return st;
}

/* power series of sin and cos with QuadRay's fused multiply-adds
 * (sinps_rr / cosps_rr tracer.cpp:1982-2015, coefficients
 * engine.cpp:912-918) */
__device__ __forceinline__ float pt_sin(float x)
{
    const float t = x * x;
    float d = x, v = x * t;
    d = fmaf(v, -0.1666666666666666666666666666666666666666666f, d);
    v = v * t; d = fmaf(v, +0.0083333333333333333333333333333333333333333f, d);
    v = v * t; d = fmaf(v, -0.0001984126984126984126984126984126984126984f, d);
    v = v * t; d = fmaf(v, +0.0000027557319223985890652557319223985890652f, d);
    return d;
}

__device__ __forceinline__ float pt_cos(float x)
{
    const float t = x * x;
    float d = 1.0f, v = d * t;
    d = fmaf(v, -0.5f, d);
    v = v * t; d = fmaf(v, +0.0416666666666666666666666666666666666666666f, d);
    v = v * t; d = fmaf(v, -0.0013888888888888888888888888888888888888888f, d);
    v = v * t; d = fmaf(v, +0.0000248015873015873015873015873015873015873f, d);
    return d;
}

/* tent-filtered pixel offset (tracer.cpp:2266-2342) */
__device__ __forceinline__ float pt_tent(uint32_t &st, int fsaa)
{
    float r = pt_random(st);
    r = r + r;
    float d = r < 1.0f ? sqrtf(r) - 1.0f : 1.0f - sqrtf(2.0f - r);
    d = d * 0.5f;
    if (fsaa != 0) d = d * 0.5f;
    return d;
}

/* chunk slot j -> pixel, sub-sample and plane slot: tiles of 8x8 pixels in
 * row-major order, pixels row-major in a tile, samples innermost */
__device__ __forceinline__ bool chunk_pixel(const PtArgs &a, uint32_t j, int &x, int &y,
                                            int &k, size_t &slot)
{
    const uint32_t per_tile = PT_TILE * PT_TILE * (uint32_t)a.ns;
    const uint32_t lt = j / per_tile;
    const uint32_t rem = j - lt * per_tile;
    const uint32_t p = rem / (uint32_t)a.ns;
    k = (int)(rem - p * (uint32_t)a.ns);
    const uint32_t tile = a.tile0 + lt;
    x = (int)((tile % a.tiles_x) * PT_TILE + p % PT_TILE);
    y = (int)((tile / a.tiles_x) * PT_TILE + p / PT_TILE);
    slot = ((size_t)y * a.fd.x_res + x) * a.ns + k;
    return x < a.fd.x_res && y < a.fd.y_res;
}

/* a lane's first slot for "cnt" (0..3) entries of one queue: the wave's
 * prefix from two ballots of the count's bits, the workgroup's from the
 * waves' totals in LDS, one global atomic per workgroup.  Every thread of
 * the workgroup calls it. */
__device__ __forceinline__ uint32_t wg_append(uint32_t cnt, uint32_t *counter,
                                              uint32_t *sh_tot, uint32_t *sh_base)
{
    const int wv = threadIdx.x >> 6;
    const int lane = __lane_id();
    const uint64_t b0 = __ballot(cnt & 1u), b1 = __ballot((cnt >> 1) & 1u);
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t pre = (uint32_t)__popcll(b0 & below) + 2u * (uint32_t)__popcll(b1 & below);
    __syncthreads();                    /* the previous call's totals are read */
    if (lane == 0) sh_tot[wv] = (uint32_t)__popcll(b0) + 2u * (uint32_t)__popcll(b1);
    __syncthreads();
    if (threadIdx.x == 0)
    {
        uint32_t run = 0;
        for (int v = 0; v < WAVES; v++)
        {
            const uint32_t t = sh_tot[v];
            sh_tot[v] = run;
            run += t;
        }
        *sh_base = run != 0 ? atomicAdd(counter, run) : 0u;
    }
    __syncthreads();
    return *sh_base + sh_tot[wv] + pre;
}

__device__ __forceinline__ void put_ray(const PtArgs &a, int out, uint32_t idx, const Hit &h,
                                        const V3 &dir, int flags, uint32_t st, const V3 &mul,
                                        uint32_t key, uint32_t sid)
{
    const PtQueue &q = a.q[out];
    q.a[idx] = make_float4(h.hit.x, h.hit.y, h.hit.z, __int_as_float(h.srf));
    q.b[idx] = make_float4(dir.x, dir.y, dir.z, __int_as_float(flags));
    q.c[idx] = make_float4(h.loc.x, h.loc.y, h.loc.z, __uint_as_float(st));
    q.d[idx] = make_float4(mul.x, mul.y, mul.z, __uint_as_float(key));
    q.sid[idx] = sid;
}

/* bounce L: the rays of queue L (for L = 0 the camera samples of every pass
 * of the run), each hit's emission into its sample's list and its children
 * into queue L + 1 (3420-3920 emission, roulette and diffuse bounce;
 * 4686-4752 the refraction / reflection pick; 4762-4800, 5175-5215 their
 * throughputs) */
/* six waves per SIMD (80 VGPRs, 7 spilled): with workgroups handed out as
 * others finish, more resident waves hide more of the walks' latency
 * (measured per frame: 40.1 ms at five, 37.6 at six, 37.5 at seven with 34
 * spilled, 40.8 at eight) */
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(6, 6)))
k_pt_bounce(PtArgs a, int L)
{
    __shared__ uint32_t sh_tot[WAVES];
    __shared__ uint32_t sh_base;
    /* an overflowed queue holds only its first cap_q rays */
    const uint32_t n = L == 0 ? a.n0 * a.passes : min(a.cnt[L], a.cap_q);
    const int in = L & 1, out = (L + 1) & 1;
    const int depth = a.fd.depth - L;
    const uint32_t shift = 2u * (uint32_t)(PT_MAX_DEPTH - L);
    for (uint32_t base = blockIdx.x * WG; base < n; base += gridDim.x * WG)
    {
        const uint32_t i = base + threadIdx.x;
        bool live = i < n;
        Ray r;
        r.org = V3{0.0f, 0.0f, 0.0f};
        r.dir = V3{0.0f, 0.0f, 1.0f};
        r.flags = 0;
        r.origin = -1;
        r.oloc = V3{0.0f, 0.0f, 0.0f};
        r.t_min = 0.0f;
        V3 mul = V3{1.0f, 1.0f, 1.0f};
        uint32_t st = 0, key = 0, sid = 0;
        if (live && L == 0)
        {
            sid = i;
            const uint32_t pass = i / a.n0, j = i - pass * a.n0;
            int x, y, k;
            size_t slot;
            live = chunk_pixel(a, j, x, y, k, slot);
            if (live)
            {
                st = lcg_steps(a.seeds[slot], pass);
                const qr_frame_desc &fd = a.fd;
                int sk = k;
                if (fd.fsaa == 1 && (x & 1)) sk = k + 2;
                float hs = (float)x;
                float vs = (float)y;
                if (fd.fsaa != 0)
                {
                    hs = hs + fd.aa_h[sk];
                    vs = vs + fd.aa_v[sk];
                }
                hs = hs + pt_tent(st, fd.fsaa);
                vs = vs + pt_tent(st, fd.fsaa);
                float hx, vx;
                hx = fd.hor[0] * hs; vx = fd.ver[0] * vs; hx = hx + vx; r.dir.x = hx + fd.dir[0];
                hx = fd.hor[1] * hs; vx = fd.ver[1] * vs; hx = hx + vx; r.dir.y = hx + fd.dir[1];
                hx = fd.hor[2] * hs; vx = fd.ver[2] * vs; hx = hx + vx; r.dir.z = hx + fd.dir[2];
                r.org = V3{fd.org[0], fd.org[1], fd.org[2]};
                r.t_min = fd.t_min;
            }
        }
        else if (live)
        {
            const PtQueue &q = a.q[in];
            const float4 qa = q.a[i], qb = q.b[i], qc = q.c[i], qd = q.d[i];
            r.org = V3{qa.x, qa.y, qa.z};
            r.origin = __float_as_int(qa.w);
            r.dir = V3{qb.x, qb.y, qb.z};
            r.flags = __float_as_int(qb.w);
            r.oloc = V3{qc.x, qc.y, qc.z};
            st = __float_as_uint(qc.w);
            mul = V3{qd.x, qd.y, qd.z};
            key = __float_as_uint(qd.w);
            sid = q.sid[i];
        }
        r.t_max = 3.402823466e+38f;
        r.depth = depth;

        Hit h;
        h.srf = -1;
        h.side = 0;
        h.hit = V3{0.0f, 0.0f, 0.0f};
        h.loc = V3{0.0f, 0.0f, 0.0f};
        bool hit = false;
        if (live)
        {
            const bool rev = r.origin < 0 || own_list(a.sc.srf[r.origin]);
            hit = walk<false>(a.sc, r, h, rev);
        }
        bool emit = false;
        V3 col = V3{0.0f, 0.0f, 0.0f};
        uint32_t kids = 0;              /* bit 0 reflection, 1 refraction, 2 diffuse */
        V3 d_dir = V3{0, 0, 0}, t_dir = V3{0, 0, 0}, r_dir = V3{0, 0, 0};
        V3 d_mul = V3{0, 0, 0}, t_mul = V3{0, 0, 0}, r_mul = V3{0, 0, 0};
        uint32_t d_st = 0;
        if (hit)
        {
            hit_info(a.sc, r, h);
            const QC qr_surface &s = a.sc.srf[h.srf];
            const uint32_t props = s.props[h.side];
            const QC qr_material &m = a.sc.mat[s.mat[h.side]];
            const V3 nrm = normal_at(a.sc, s, h);
            V3 tex = texel_color(a.sc, m, props, h);

            const uint32_t pass = sid / a.n0;
            const float pts_o = 1.0f / (a.pts_c0 + (float)pass);
            float c;
            c = m.col[0] * pts_o; col.x = c * mul.x;
            c = m.col[1] * pts_o; col.y = c * mul.y;
            c = m.col[2] * pts_o; col.z = c * mul.z;
            emit = col.x != 0.0f || col.y != 0.0f || col.z != 0.0f;

            if (props & QR_PROP_DIFFUSE)
            {
                bool alive = true;
                if (depth <= PT_ROULETTE_DEPTH)
                {
                    const float p = fmaxf(fmaxf(tex.x, tex.y), tex.z);
                    const float rr = pt_random(st);
                    if (rr < p)
                    {
                        const float ip = 1.0f / p;
                        tex.x = tex.x * ip; tex.y = tex.y * ip; tex.z = tex.z * ip;
                    }
                    else
                    {
                        alive = false;
                    }
                }
                if (alive)
                {
                    const V3 &ry = r.dir;
                    V3 u, v;
                    u.x = nrm.y * ry.z - nrm.z * ry.y;
                    u.y = nrm.z * ry.x - nrm.x * ry.z;
                    u.z = nrm.x * ry.y - nrm.y * ry.x;
                    float l2 = u.x * u.x + u.y * u.y;
                    l2 = l2 + u.z * u.z;
                    const float il = rsq(l2);
                    u.x = u.x * il; u.y = u.y * il; u.z = u.z * il;
                    v.x = nrm.y * u.z - nrm.z * u.y;
                    v.y = nrm.z * u.x - nrm.x * u.z;
                    v.z = nrm.x * u.y - nrm.y * u.x;
                    const float r1 = pt_random(st);
                    const float sn = sqrtf(r1);
                    const float cs = sqrtf(1.0f - r1);
                    const float r2 = pt_random(st);
                    float phi = (r2 + r2) * 3.14159265358979323846f;
                    phi = phi - 3.14159265358979323846f;
                    const float cu = pt_cos(phi) * sn, sv = pt_sin(phi) * sn;
                    d_dir.x = nrm.x * cs; d_dir.x = d_dir.x + u.x * cu; d_dir.x = d_dir.x + v.x * sv;
                    d_dir.y = nrm.y * cs; d_dir.y = d_dir.y + u.y * cu; d_dir.y = d_dir.y + v.y * sv;
                    d_dir.z = nrm.z * cs; d_dir.z = d_dir.z + u.z * cu; d_dir.z = d_dir.z + v.z * sv;
                    if (depth != 0)
                    {
                        float w;
                        w = mul.x * m.l_dff; d_mul.x = w * tex.x;
                        w = mul.y * m.l_dff; d_mul.y = w * tex.y;
                        w = mul.z * m.l_dff; d_mul.z = w * tex.z;
                        d_st = st;
                        kids |= 4u;
                    }
                }
            }

            TrRf t = trrf_setup(m, props, r.dir, nrm);
            if (t.fresnel_tr && depth <= PT_SPLIT_DEPTH)
            {
                const float rr = pt_random(st);
                float qv = t.c_rfl / (t.c_trn + t.c_rfl);
                qv = qv * 0.5f;
                const float pr = 0.25f + qv;
                const bool pick_tr = rr >= pr;
                t.c_trn = pick_tr ? t.c_trn / (1.0f - pr) : 0.0f;
                t.c_rfl = pick_tr ? 0.0f : t.c_rfl / pr;
                t.tr = t.tr && pick_tr;
                t.rf = t.rf && !pick_tr;
            }
            if (t.tr && depth != 0)
            {
                t_dir = t.tdir;
                t_mul = V3{mul.x * t.c_trn, mul.y * t.c_trn, mul.z * t.c_trn};
                kids |= 2u;
            }
            if (t.rf && depth != 0)
            {
                r_dir = t.rdir;
                r_mul = V3{mul.x * t.c_rfl, mul.y * t.c_rfl, mul.z * t.c_rfl};
                kids |= 1u;
            }
        }

        /* the sample's contribution, linked into its list */
        const uint32_t ci = wg_append(emit ? 1u : 0u, &a.cnt[CNT_CONTRIB], sh_tot, &sh_base);
        if (emit)
        {
            if (ci < a.cap_c)
            {
                a.ct[ci] = make_float4(col.x, col.y, col.z, __uint_as_float(key << shift));
                a.ct_next[ci] = atomicExch(&a.heads[sid], ci);
            }
            else
            {
                a.cnt[CNT_OVERFLOW] = 1u;
            }
        }

        /* the children, compacted into the next bounce's queue; every slot
         * below the capacity is written, so the next bounce reads only rays
         * of this run even after an overflow */
        const uint32_t nk = (uint32_t)__popc(kids);
        uint32_t at = wg_append(nk, &a.cnt[L + 1], sh_tot, &sh_base);
        if (nk != 0 && (uint64_t)at + nk > a.cap_q) a.cnt[CNT_OVERFLOW] = 1u;
        if ((kids & 1u) && at < a.cap_q)
            put_ray(a, out, at, h, r_dir, h.side, st, r_mul, key * 4u + KEY_RF, sid);
        at += kids & 1u;
        if ((kids & 2u) && at < a.cap_q)
            put_ray(a, out, at, h, t_dir, h.side | FLAG_PASS, st, t_mul, key * 4u + KEY_TR, sid);
        at += (kids >> 1) & 1u;
        if ((kids & 4u) && at < a.cap_q)
            put_ray(a, out, at, h, d_dir, h.side, d_st, d_mul, key * 4u + KEY_DF, sid);
    }
}

/* start of a run: counters and the samples' list heads cleared (a fill
 * kernel rather than a memset, so the launch can be captured) */
__global__ void __launch_bounds__(WG) k_pt_reset(PtArgs a)
{
    const uint32_t n = a.n0 * a.passes;
    for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < n; i += gridDim.x * WG)
    {
        a.heads[i] = 0xFFFFFFFFu;
    }
    if (blockIdx.x == 0 && threadIdx.x < CNT_N) a.cnt[threadIdx.x] = 0u;
}

/* per slot, the run's passes in order: the running mean times (1 - 1/n),
 * plus the pass's contributions in depth-first order (FF_pts
 * tracer.cpp:2113-2140, PT_emt 3420-3460); then the seed stepped once per
 * pass */
__global__ void __launch_bounds__(WG) k_pt_fold(PtArgs a)
{
    const uint32_t j = blockIdx.x * WG + threadIdx.x;
    if (j >= a.n0 || a.cnt[CNT_OVERFLOW] != 0u) return;
    int x, y, k;
    size_t slot;
    if (!chunk_pixel(a, j, x, y, k, slot)) return;
    float pr = a.pr[slot], pg = a.pg[slot], pb = a.pb[slot];
    for (uint32_t p = 0; p < a.passes; p++)
    {
        const float pts_o = 1.0f / (a.pts_c0 + (float)p);
        const float pts_u = 1.0f - pts_o;
        pr = pr * pts_u;
        pg = pg * pts_u;
        pb = pb * pts_u;
        const uint32_t head = a.heads[p * a.n0 + j];
        int64_t last = -1;              /* keys are non-negative, distinct */
        for (;;)
        {
            /* the next contribution in order: the smallest key above the
             * last one taken (a sample's list holds a few entries) */
            uint32_t best = 0xFFFFFFFFu;
            int64_t bkey = INT64_MAX;
            for (uint32_t c = head; c != 0xFFFFFFFFu; c = a.ct_next[c])
            {
                const int64_t kc = (int64_t)__float_as_uint(a.ct[c].w);
                if (kc > last && kc < bkey)
                {
                    bkey = kc;
                    best = c;
                }
            }
            if (best == 0xFFFFFFFFu) break;
            const float4 v = a.ct[best];
            pr = pr + v.x;
            pg = pg + v.y;
            pb = pb + v.z;
            last = bkey;
        }
    }
    a.pr[slot] = pr;
    a.pg[slot] = pg;
    a.pb[slot] = pb;
    a.seeds[slot] = lcg_steps(a.seeds[slot], a.passes);
}

/* the chunk's pixels from the planes: clamp, anti-aliasing resolve, ARGB8;
 * four tiles of 8x8 pixels per workgroup */
__global__ void __launch_bounds__(WG) k_pt_resolve(PtArgs a)
{
    const uint32_t lt = blockIdx.x * 4 + threadIdx.x / 64;
    if (lt >= a.ntiles || a.cnt[CNT_OVERFLOW] != 0u) return;
    const uint32_t tile = a.tile0 + lt, p = threadIdx.x % 64;
    const int x = (int)((tile % a.tiles_x) * PT_TILE + p % PT_TILE);
    const int y = (int)((tile / a.tiles_x) * PT_TILE + p / PT_TILE);
    if (x >= a.fd.x_res || y >= a.fd.y_res) return;
    const size_t s0 = ((size_t)y * a.fd.x_res + x) * a.ns;
    V3 c[4];
    #pragma unroll
    for (int k = 0; k < 4; k++)
    {
        const size_t sl = s0 + (k < a.ns ? k : 0);
        c[k].x = a.pr[sl] < 1.0f ? a.pr[sl] : 1.0f;
        c[k].y = a.pg[sl] < 1.0f ? a.pg[sl] : 1.0f;
        c[k].z = a.pb[sl] < 1.0f ? a.pb[sl] : 1.0f;
    }
    a.frame[(ptrdiff_t)y * a.fd.x_row + x] = resolve_pack(a.fd, a.ns, c[0], c[1], c[2], c[3]);
}

struct Buf
{
    void  *ptr = nullptr;
    size_t size = 0;
    hipError_t reserve(size_t n)
    {
        if (n <= size) return hipSuccess;
        if (ptr != nullptr) (void)hipFree(ptr);
        ptr = nullptr;
        size = 0;
        hipError_t e = hipMalloc(&ptr, n);
        if (e == hipSuccess) size = n;
        return e;
    }
    void release()
    {
        if (ptr != nullptr) (void)hipFree(ptr);
        ptr = nullptr;
        size = 0;
    }
};

uint32_t round_up(uint64_t n)
{
    uint64_t c = 1024;
    while (c < n) c += c >> 1;
    return (uint32_t)std::min<uint64_t>(c, 0xFFFFFF00ull);
}

} /* namespace */

struct qr_pt_state
{
    Buf qa[2], qb[2], qc[2], qd[2], qs[2];
    Buf ct, ct_next, heads, cnt;
    uint32_t cap_q = 0, cap_c = 0;
    uint64_t heads_cap = 0;
    uint32_t *cnt_host = nullptr;       /* pinned copy after a run */
    uint32_t want_q = 0, want_c = 0;    /* counted need of an overflowed run */
    int cus = 256;
    PtArgs last;                        /* the last run, traced again by
                                           qr_pt_retry after an overflow */
    hipStream_t last_stream = nullptr;
};

namespace {

hipError_t pt_reserve(qr_pt_state *s, uint64_t samples)
{
    hipError_t e;
    const uint32_t want_q = std::max(round_up(samples), s->want_q);
    if (s->cap_q < want_q)
    {
        const size_t n = want_q;
        for (int k = 0; k < 2; k++)
        {
            if ((e = s->qa[k].reserve(16 * n)) != hipSuccess) return e;
            if ((e = s->qb[k].reserve(16 * n)) != hipSuccess) return e;
            if ((e = s->qc[k].reserve(16 * n)) != hipSuccess) return e;
            if ((e = s->qd[k].reserve(16 * n)) != hipSuccess) return e;
            if ((e = s->qs[k].reserve(4 * n)) != hipSuccess) return e;
        }
        s->cap_q = want_q;
    }
    const uint32_t want_c = std::max(round_up(samples / 4 + 1024), s->want_c);
    if (s->cap_c < want_c)
    {
        if ((e = s->ct.reserve(16 * (size_t)want_c)) != hipSuccess) return e;
        if ((e = s->ct_next.reserve(4 * (size_t)want_c)) != hipSuccess) return e;
        s->cap_c = want_c;
    }
    if (s->heads_cap < samples)
    {
        if ((e = s->heads.reserve(4 * (size_t)samples)) != hipSuccess) return e;
        s->heads_cap = samples;
    }
    return hipSuccess;
}

/* enqueue one run (reset, bounces, fold) with the state's buffers */
hipError_t pt_run(qr_pt_state *s, PtArgs &a, hipStream_t stream)
{
    hipError_t e;
    if ((e = pt_reserve(s, (uint64_t)a.n0 * a.passes)) != hipSuccess) return e;
    for (int k = 0; k < 2; k++)
    {
        a.q[k].a = (float4 *)s->qa[k].ptr;
        a.q[k].b = (float4 *)s->qb[k].ptr;
        a.q[k].c = (float4 *)s->qc[k].ptr;
        a.q[k].d = (float4 *)s->qd[k].ptr;
        a.q[k].sid = (uint32_t *)s->qs[k].ptr;
    }
    a.cap_q = s->cap_q;
    a.ct = (float4 *)s->ct.ptr;
    a.ct_next = (uint32_t *)s->ct_next.ptr;
    a.cap_c = s->cap_c;
    a.heads = (uint32_t *)s->heads.ptr;
    a.cnt = (uint32_t *)s->cnt.ptr;
    /* 256 workgroups per compute unit: a bounce's rays (their count is on
     * the device) take one or two blocks of 256 per workgroup, and the
     * hardware hands out workgroups as others finish, which balances the
     * incoherent bounces' uneven waves (measured on the 1280x720 64-spp
     * box: 8 workgroups per CU 48.2 ms per frame, 40 -> 42.0, 128 -> 40.7,
     * 256 -> 40.2, 512 -> 40.4) */
    const uint32_t grid = (uint32_t)s->cus * 256u;
    const uint64_t n = (uint64_t)a.n0 * a.passes;
    const uint32_t g0 = (uint32_t)std::min<uint64_t>((n + WG - 1) / WG, grid);
    hipLaunchKernelGGL(k_pt_reset, dim3(g0), dim3(WG), 0, stream, a);
    /* a bounce's ray count is on the device; its bound here (at most three
     * children a ray, at most the queue's capacity) keeps a small run from
     * launching workgroups that find nothing to do */
    uint64_t bound = n;
    for (int L = 0; L <= a.fd.depth; L++)
    {
        const uint32_t gl = (uint32_t)std::min<uint64_t>((bound + WG - 1) / WG, grid);
        hipLaunchKernelGGL(k_pt_bounce, dim3(std::max(gl, 1u)), dim3(WG), 0, stream, a, L);
        bound = std::min<uint64_t>(bound * 3, a.cap_q);
    }
    hipLaunchKernelGGL(k_pt_fold, dim3((a.n0 + WG - 1) / WG), dim3(WG), 0, stream, a);
    return hipGetLastError();
}

hipError_t pt_counters(qr_pt_state *s, hipStream_t stream)
{
    return hipMemcpyAsync(s->cnt_host, s->cnt.ptr, sizeof(uint32_t) * CNT_N,
                          hipMemcpyDeviceToHost, stream);
}

/* after an overflow: the counted need for the next try */
void pt_grow(qr_pt_state *s)
{
    uint32_t need = 0;
    for (int L = 1; L < CNT_CONTRIB; L++) need = std::max(need, s->cnt_host[L]);
    s->want_q = std::max(s->want_q, round_up(need));
    s->want_c = std::max(s->want_c, round_up(s->cnt_host[CNT_CONTRIB]));
    s->cnt_host[CNT_OVERFLOW] = 0u;
}

} /* namespace */

hipError_t qr_pt_create(qr_pt_state **out)
{
    *out = nullptr;
    qr_pt_state *s = new (std::nothrow) qr_pt_state();
    if (s == nullptr) return hipErrorOutOfMemory;
    hipError_t e = hipHostMalloc((void **)&s->cnt_host, sizeof(uint32_t) * CNT_N,
                                 hipHostMallocDefault);
    if (e == hipSuccess) e = s->cnt.reserve(sizeof(uint32_t) * CNT_N);
    if (e != hipSuccess)
    {
        qr_pt_destroy(s);
        return e;
    }
    memset(s->cnt_host, 0, sizeof(uint32_t) * CNT_N);
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        cus > 0)
    {
        s->cus = cus;
    }
    *out = s;
    return hipSuccess;
}

void qr_pt_destroy(qr_pt_state *s)
{
    if (s == nullptr) return;
    for (int k = 0; k < 2; k++)
    {
        s->qa[k].release(); s->qb[k].release(); s->qc[k].release(); s->qd[k].release();
        s->qs[k].release();
    }
    s->ct.release();
    s->ct_next.release();
    s->heads.release();
    s->cnt.release();
    if (s->cnt_host != nullptr) (void)hipHostFree(s->cnt_host);
    delete s;
}

uint32_t qr_pt_tiles(int x_res, int y_res)
{
    return (uint32_t)((x_res + PT_TILE - 1) / PT_TILE) * (uint32_t)((y_res + PT_TILE - 1) / PT_TILE);
}

hipError_t qr_pt_launch(qr_pt_state *s, const qr_device_tables *t, const qr_frame_desc *fd,
                        float *planes, uint32_t *seeds, uint32_t *frame, uint32_t tile0,
                        uint32_t ntiles, int passes, hipStream_t stream)
{
    if (fd->depth < 0 || fd->depth > PT_MAX_DEPTH || !(fd->pts_c >= 1.0f) || passes < 1 ||
        !(fd->pts_c + (float)passes < 16777216.0f))
    {
        return hipErrorInvalidValue;
    }
    const uint32_t all = qr_pt_tiles(fd->x_res, fd->y_res);
    if (ntiles == 0)
    {
        tile0 = 0;
        ntiles = all;
    }
    if ((uint64_t)tile0 + ntiles > all) return hipErrorInvalidValue;
    PtArgs a;
    a.sc = scene_of(t);
    a.fd = *fd;
    a.fd.x_row = fd->x_res;
    a.ns = fd->fsaa == 2 ? 4 : fd->fsaa == 1 ? 2 : 1;
    const size_t slots = (size_t)fd->x_res * fd->y_res * a.ns;
    a.pr = planes;
    a.pg = planes + slots;
    a.pb = planes + 2 * slots;
    a.seeds = seeds;
    a.frame = frame;
    a.tiles_x = (uint32_t)(fd->x_res + PT_TILE - 1) / PT_TILE;
    a.tile0 = tile0;
    a.ntiles = ntiles;
    a.n0 = ntiles * PT_TILE * PT_TILE * (uint32_t)a.ns;
    /* runs of passes of at most PT_RUN_SAMPLES samples (at least one pass);
     * counts of passes are whole floats below 2^24, so the sums are exact */
    const uint32_t run = (uint32_t)std::max<uint64_t>(1, PT_RUN_SAMPLES / a.n0);
    hipError_t e;
    float pts_c = fd->pts_c;
    uint32_t done = 0;
    for (;;)
    {
        a.passes = std::min<uint32_t>(run, (uint32_t)passes - done);
        a.pts_c0 = pts_c;
        if ((e = pt_run(s, a, stream)) != hipSuccess) return e;
        if ((e = pt_counters(s, stream)) != hipSuccess) return e;
        done += a.passes;
        pts_c = pts_c + (float)a.passes;
        if (done >= (uint32_t)passes) break;
        /* the next run reads the planes and seeds this one wrote: an
         * overflowed run wrote none and is traced again first */
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
        for (int attempt = 0; s->cnt_host[CNT_OVERFLOW] != 0u; attempt++)
        {
            if (attempt == 16) return hipErrorOutOfMemory;
            pt_grow(s);
            if ((e = pt_run(s, a, stream)) != hipSuccess) return e;
            if ((e = pt_counters(s, stream)) != hipSuccess) return e;
            if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
        }
    }
    hipLaunchKernelGGL(k_pt_resolve, dim3((ntiles + 3) / 4), dim3(WG), 0, stream, a);
    s->last = a;
    s->last_stream = stream;
    return hipGetLastError();
}

int qr_pt_check(qr_pt_state *s)
{
    if (s->cnt_host[CNT_OVERFLOW] == 0u) return 0;
    pt_grow(s);
    return 1;
}

hipError_t qr_pt_retry(qr_pt_state *s, hipStream_t stream)
{
    PtArgs a = s->last;
    hipError_t e;
    if ((e = pt_run(s, a, stream)) != hipSuccess) return e;
    if ((e = pt_counters(s, stream)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_pt_resolve, dim3((a.ntiles + 3) / 4), dim3(WG), 0, stream, a);
    s->last = a;
    return hipGetLastError();
}
