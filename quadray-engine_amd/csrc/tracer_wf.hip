/*
 * tracer_wf.hip -- level-by-level Whitted pipeline for MI355X (gfx950).
 *
 * The Whitted tree of every camera sample is traced one tree level at a
 * time.  Per level L, over a compacted queue of that level's rays:
 *
 *   walk_prep   nearest-hit walk of each ray; for a hit, the hit point,
 *               local hit and normal, and one shadow job per light facing
 *               the hit, appended to that light's queue
 *   shadow      any-hit walk of each shadow job; an unoccluded light sets
 *               its bit in the hit's visibility mask
 *   final       shading with the visibility mask (QuadRay's MT/LT/TR/RF
 *               arithmetic), children appended to level L+1's queue
 *
 * then the tree is folded bottom-up, child colour times its coefficient
 * plus the node's colour, refraction before reflection (QuadRay's TR_mix /
 * RF_mix order, core/tracer/tracer.cpp 4862-4936, 5277-5357), and level 0 is
 * resolved to pixels.  Appends are counted per wave (ballot + prefix count)
 * and per workgroup in LDS, with one global atomic per workgroup and queue
 * pair, so queues stay dense and lanes stay busy between bounces; the walk
 * kernels keep only a ray and its walk registers live.
 *
 * Queues are preallocated; an append beyond capacity only sets a flag and
 * keeps counting, the host grows the queue to the counted size and traces
 * the frame again, so capacity never changes a result.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstring>

#include "qr_soa.h"
#include "tracer_dev.h"
#include "tracer_kernels.h"

namespace {

constexpr int TILE = 8;             /* level-0 samples: 8x8-pixel tiles */
constexpr int MAX_LEVELS = 16;      /* depth 0..14 plus one */
constexpr int MAX_LIGHTS = 32;      /* visibility mask bits */

/* counter block: [0] overflow flag, [1] largest shadow queue of the frame,
 * [2 + L] queue length of level L, [CNT_SH + l] shadow jobs of light l */
constexpr int CNT_OVF = 0;
constexpr int CNT_SHMAX = 1;
constexpr int CNT_LVL = 2;
constexpr int CNT_SH = CNT_LVL + MAX_LEVELS;
constexpr int CNT_N = CNT_SH + MAX_LIGHTS;
static_assert(CNT_SH % 2 == 0, "shadow counters are updated in 64-bit pairs");

/* one tree level in HBM (structure of float4 arrays, one entry per ray) */
struct Level
{
    float4   *ra, *rb, *rc;     /* ray: org|origin, dir|flags, oloc (L >= 1) */
    float4   *na, *nb, *nc;     /* hit: point|srf<<1|side, loc|u, normal|v */
    uint32_t *vis;              /* visible-light mask */
    float4   *fa, *fb;          /* fold: colour|c_trn, c_rfl|tr|rf */
    uint32_t  cap;
};

struct Args
{
    Scene          sc;
    qr_frame_desc  fd;
    Level          cur, nxt;
    uint32_t      *cnt;
    uint32_t      *shq;         /* shadow jobs: light l at [l * sh_cap] */
    uint32_t       sh_cap;
    int            level;
    uint32_t       n0;          /* level-0 samples (padded to whole tiles) */
    uint32_t       tiles_x;
    int            ns;          /* samples per pixel */
    int32_t       *ids;
    uint32_t      *frame;
};

__device__ __forceinline__ void overflow(uint32_t *cnt)
{
    atomicOr(&cnt[CNT_OVF], 1u);
}

constexpr int WAVES = 256 / 64;     /* waves of a 256-thread workgroup */

/* Workgroup-wide append to up to MAX_LIGHTS queues whose counters are
 * consecutive 32-bit words starting 8-byte aligned.  Waves count their lanes
 * in LDS; then one 64-bit atomic per pair of queues and workgroup reserves
 * the workgroup's slots (two 32-bit totals never carry into each other: each
 * stays below 2^32).  Every thread of the workgroup calls it with bit q of
 * "mask" set for one slot in queue q, and reads its slot with slot(). */
struct BlockAppend
{
    uint32_t base[MAX_LIGHTS];          /* workgroup's first slot per queue */
    uint32_t woff[WAVES][MAX_LIGHTS];   /* wave's offset within the workgroup */
};

__device__ __forceinline__ void block_append(BlockAppend &b, uint32_t mask, int nq,
                                             uint32_t *counters)
{
    const int wv = threadIdx.x >> 6;
    __syncthreads();    /* the previous call's slots have been read */
    if ((int)threadIdx.x < nq) b.base[threadIdx.x] = 0u;
    __syncthreads();
    for (int q = 0; q < nq; q++)
    {
        const uint64_t m = __ballot((mask >> q) & 1u);
        if (__lane_id() == 0)
        {
            b.woff[wv][q] = m != 0 ? atomicAdd(&b.base[q], (uint32_t)__popcll(m)) : 0u;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < (nq + 1) / 2)
    {
        const int q = 2 * threadIdx.x;
        const uint64_t lo = b.base[q];
        const uint64_t hi = q + 1 < nq ? b.base[q + 1] : 0u;
        uint64_t old = 0;
        if ((lo | hi) != 0)
        {
            old = atomicAdd((unsigned long long *)&counters[q],
                            (unsigned long long)(lo | (hi << 32)));
        }
        b.base[q] = (uint32_t)old;
        if (q + 1 < nq) b.base[q + 1] = (uint32_t)(old >> 32);
    }
    __syncthreads();
}

/* Workgroup-wide append of up to two slots per lane (bit 0, then bit 1 of
 * mask) to one queue: one 32-bit atomic per workgroup.  Every thread of the
 * workgroup calls it; the lane's first slot is returned, its second follows
 * the wave's bit-0 slots (pair_slot). */
struct PairAppend
{
    uint32_t base;
    uint32_t woff[WAVES];
};

__device__ __forceinline__ void pair_append(PairAppend &b, uint32_t mask, uint32_t *counter)
{
    const int wv = threadIdx.x >> 6;
    __syncthreads();
    if (threadIdx.x == 0) b.base = 0u;
    __syncthreads();
    const uint64_t m0 = __ballot(mask & 1u), m1 = __ballot((mask >> 1) & 1u);
    if (__lane_id() == 0)
    {
        const uint32_t c = (uint32_t)(__popcll(m0) + __popcll(m1));
        b.woff[wv] = c != 0 ? atomicAdd(&b.base, c) : 0u;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        b.base = b.base != 0 ? atomicAdd(counter, b.base) : 0u;
    }
    __syncthreads();
}

/* slot of bit k (0 or 1) of a lane's mask; every lane of the wave calls it */
__device__ __forceinline__ uint32_t pair_slot(const PairAppend &b, uint32_t mask, int k)
{
    const uint64_t m0 = __ballot(mask & 1u), m1 = __ballot((mask >> 1) & 1u);
    const int lane = __lane_id();
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t w = b.base + b.woff[threadIdx.x >> 6];
    return k == 0 ? w + (uint32_t)__popcll(m0 & below)
                  : w + (uint32_t)__popcll(m0) + (uint32_t)__popcll(m1 & below);
}

/* slot in queue q of a lane with bit q of mask set; every lane of the wave
 * calls it */
__device__ __forceinline__ uint32_t block_slot(const BlockAppend &b, uint32_t mask, int q)
{
    const uint64_t m = __ballot((mask >> q) & 1u);
    const int lane = __lane_id();
    const uint64_t below = lane == 0 ? 0ull : (m & (~0ull >> (64 - lane)));
    return b.base[q] + b.woff[threadIdx.x >> 6][q] + (uint32_t)__popcll(below);
}

/* level-0 sample i -> pixel and sub-sample (8x8 tiles, samples innermost) */
__device__ __forceinline__ bool sample_pixel(const Args &a, uint32_t i, int &x,
                                             int &y, int &k)
{
    const uint32_t per_tile = TILE * TILE * a.ns;
    const uint32_t tile = i / per_tile;
    const uint32_t rem = i - tile * per_tile;
    const uint32_t p = rem / a.ns;
    k = (int)(rem - p * a.ns);
    x = (int)((tile % a.tiles_x) * TILE + p % TILE);
    y = (int)((tile / a.tiles_x) * TILE + p / TILE);
    return x < a.fd.x_res && y < a.fd.y_res;
}

/* camera ray of a sub-sample (render_slice, engine.cpp:3514-3697 and the
 * 2X / 4X offsets of 6942-7150) */
__device__ __forceinline__ Ray camera_ray(const qr_frame_desc &fd, int x, int y, int k)
{
    int sk = k;
    if (fd.fsaa == 1 && (x & 1)) sk = k + 2;
    float hs = (float)x;
    float vs = (float)y;
    if (fd.fsaa != 0)
    {
        hs = hs + fd.aa_h[sk];
        vs = vs + fd.aa_v[sk];
    }
    Ray r;
    float hx, vx;
    hx = fd.hor[0] * hs; vx = fd.ver[0] * vs; hx = hx + vx; r.dir.x = hx + fd.dir[0];
    hx = fd.hor[1] * hs; vx = fd.ver[1] * vs; hx = hx + vx; r.dir.y = hx + fd.dir[1];
    hx = fd.hor[2] * hs; vx = fd.ver[2] * vs; hx = hx + vx; r.dir.z = hx + fd.dir[2];
    r.org = V3{fd.org[0], fd.org[1], fd.org[2]};
    r.t_min = fd.t_min;
    r.t_max = 3.402823466e+38f;
    r.flags = 0;
    r.origin = -1;
    r.oloc = V3{0.0f, 0.0f, 0.0f};
    r.depth = fd.depth;
    return r;
}

/* ray i of the current level; false for padding samples outside the frame */
template <bool L0>
__device__ __forceinline__ bool level_ray(const Args &a, uint32_t i, Ray &r)
{
    if (L0)
    {
        int x, y, k;
        const bool in = sample_pixel(a, i, x, y, k);
        r = camera_ray(a.fd, x, y, k);
        return in;
    }
    const float4 ra = a.cur.ra[i], rb = a.cur.rb[i], rc = a.cur.rc[i];
    r.org = V3{ra.x, ra.y, ra.z};
    r.origin = __float_as_int(ra.w);
    r.dir = V3{rb.x, rb.y, rb.z};
    r.flags = __float_as_int(rb.w);
    r.oloc = V3{rc.x, rc.y, rc.z};
    r.t_min = 0.0f;
    r.t_max = 3.402823466e+38f;
    r.depth = a.fd.depth - a.level;
    return true;
}

__global__ void k_reset(uint32_t *cnt)
{
    for (int i = threadIdx.x; i < CNT_N; i += blockDim.x) cnt[i] = 0u;
}

/* nearest hit of every ray of the level, hit records, shadow jobs */
template <bool L0>
__global__ void __launch_bounds__(256) k_walk_prep(Args a)
{
    __shared__ BlockAppend ba;
    const uint32_t n = L0 ? a.n0 : min(a.cnt[CNT_LVL + a.level], a.cur.cap);
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride)
    {
        const uint32_t i = base + threadIdx.x;
        Ray r;
        bool valid = false;
        if (i < n) valid = level_ray<L0>(a, i, r);

        Hit h;
        h.srf = -1;
        bool hit = false;
        if (valid) hit = walk<false>(a.sc, r, h);

        V3 nrm = V3{0.0f, 0.0f, 0.0f};
        if (hit)
        {
            hit_info(a.sc, r, h);
            nrm = normal_at(a.sc, a.sc.srf[h.srf], h);
            a.cur.na[i] = make_float4(h.hit.x, h.hit.y, h.hit.z,
                                      __int_as_float((h.srf << 1) | h.side));
            a.cur.nb[i] = make_float4(h.loc.x, h.loc.y, h.loc.z, h.u);
            a.cur.nc[i] = make_float4(nrm.x, nrm.y, nrm.z, h.v);
            a.cur.vis[i] = 0u;
        }
        else if (i < n)
        {
            a.cur.na[i] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
        }
        if (L0 && a.ids != nullptr && valid)
        {
            int x, y, k;
            sample_pixel(a, i, x, y, k);
            a.ids[((size_t)y * a.fd.x_res + x) * a.ns + k] =
                hit ? (a.sc.srf[h.srf].id << 1) | h.side : -1;
        }

        /* one shadow job per light in front of a lit hit (shade's gate) */
        const bool lit = hit && !(a.sc.srf[h.srf].props[h.side] & QR_PROP_LIGHT);
        uint32_t need = 0u;
        if (lit)
        {
            for (int li = 0; li < a.sc.n_lgt; li++)
            {
                V3 L;
                if (0.0f < light_vec(a.sc.lgt[li], h.hit, nrm, L)) need |= 1u << li;
            }
        }
        block_append(ba, need, a.sc.n_lgt, &a.cnt[CNT_SH]);
        for (int li = 0; li < a.sc.n_lgt; li++)
        {
            const uint32_t slot = block_slot(ba, need, li);
            if ((need >> li) & 1u)
            {
                if (slot < a.sh_cap) a.shq[(size_t)li * a.sh_cap + slot] = i;
                else overflow(a.cnt);
            }
        }
    }
}

/* any-hit walk of every shadow job of the level, light by light */
__global__ void __launch_bounds__(256) k_shadow(Args a)
{
    const uint32_t stride = gridDim.x * blockDim.x;
    for (int li = 0; li < a.sc.n_lgt; li++)
    {
        const uint32_t n = min(a.cnt[CNT_SH + li], a.sh_cap);
        const QC qr_light &l = a.sc.lgt[li];
        for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride)
        {
            const uint32_t j = base + threadIdx.x;
            if (j >= n) continue;
            const uint32_t i = a.shq[(size_t)li * a.sh_cap + j];
            const float4 na = a.cur.na[i], nb = a.cur.nb[i], nc = a.cur.nc[i];
            const int code = __float_as_int(na.w);
            Ray sr;
            sr.org = V3{na.x, na.y, na.z};
            light_vec(l, sr.org, V3{nc.x, nc.y, nc.z}, sr.dir);
            sr.t_min = 0.0f;
            sr.t_max = l.t_max;
            sr.flags = (code & 1) | FLAG_SHAD;
            sr.origin = code >> 1;
            sr.oloc = V3{nb.x, nb.y, nb.z};
            sr.depth = 0;
            Hit sh;
            if (!walk<true>(a.sc, sr, sh)) atomicOr(&a.cur.vis[i], 1u << li);
        }
    }
}

/* shading with the visibility masks, fold records, the next level's rays */
template <bool L0>
__global__ void __launch_bounds__(256) k_final(Args a)
{
    __shared__ PairAppend pa;
    const uint32_t n = L0 ? a.n0 : min(a.cnt[CNT_LVL + a.level], a.cur.cap);
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride)
    {
        const uint32_t i = base + threadIdx.x;
        bool hit = false;
        Node nd;
        nd.todo = 0;
        if (i < n)
        {
            const float4 na = a.cur.na[i];
            const int code = __float_as_int(na.w);
            hit = code >= 0;
            if (hit)
            {
                Ray r;
                level_ray<L0>(a, i, r);
                const float4 nb = a.cur.nb[i], nc = a.cur.nc[i];
                Hit h;
                h.srf = code >> 1;
                h.side = code & 1;
                h.hit = V3{na.x, na.y, na.z};
                h.loc = V3{nb.x, nb.y, nb.z};
                h.u = nb.w;
                h.v = nc.w;
                const V3 nrm = V3{nc.x, nc.y, nc.z};
                const uint32_t vis = a.cur.vis[i];
                shade(a.sc, a.fd, r, h, nrm, nd, [&](int li, const Ray &) {
                    return ((vis >> li) & 1u) == 0u;
                });
            }
        }
        const bool tr = hit && (nd.todo & 1);
        const bool rf = hit && (nd.todo & 2);
        const uint32_t kids = (tr ? 1u : 0u) | (rf ? 2u : 0u);
        pair_append(pa, kids, &a.cnt[CNT_LVL + a.level + 1]);
        int s_tr = (int)pair_slot(pa, kids, 0);
        int s_rf = (int)pair_slot(pa, kids, 1);
        if (tr)
        {
            if ((uint32_t)s_tr < a.nxt.cap)
            {
                a.nxt.ra[s_tr] = make_float4(nd.org.x, nd.org.y, nd.org.z,
                                             __int_as_float(nd.srf));
                a.nxt.rb[s_tr] = make_float4(nd.tdir.x, nd.tdir.y, nd.tdir.z,
                                             __int_as_float(nd.side | FLAG_PASS));
                a.nxt.rc[s_tr] = make_float4(nd.loc.x, nd.loc.y, nd.loc.z, 0.0f);
            }
            else
            {
                overflow(a.cnt);
                s_tr = -1;
            }
        }
        if (rf)
        {
            if ((uint32_t)s_rf < a.nxt.cap)
            {
                a.nxt.ra[s_rf] = make_float4(nd.org.x, nd.org.y, nd.org.z,
                                             __int_as_float(nd.srf));
                a.nxt.rb[s_rf] = make_float4(nd.rdir.x, nd.rdir.y, nd.rdir.z,
                                             __int_as_float(nd.side));
                a.nxt.rc[s_rf] = make_float4(nd.loc.x, nd.loc.y, nd.loc.z, 0.0f);
            }
            else
            {
                overflow(a.cnt);
                s_rf = -1;
            }
        }
        if (i < n)
        {
            if (hit)
            {
                a.cur.fa[i] = make_float4(nd.col.x, nd.col.y, nd.col.z, nd.c_trn);
                a.cur.fb[i] = make_float4(nd.c_rfl, __int_as_float(tr ? s_tr : -1),
                                          __int_as_float(rf ? s_rf : -1), 0.0f);
            }
            else
            {
                a.cur.fa[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                a.cur.fb[i] = make_float4(0.0f, __int_as_float(-1),
                                          __int_as_float(-1), 0.0f);
            }
        }
    }
    /* the shadow queues are free again for the next level */
    if (blockIdx.x == 0 && threadIdx.x < (unsigned)a.sc.n_lgt)
    {
        atomicMax(&a.cnt[CNT_SHMAX], a.cnt[CNT_SH + threadIdx.x]);
        a.cnt[CNT_SH + threadIdx.x] = 0u;
    }
}

/* node colour with its children's results folded in */
__device__ __forceinline__ V3 fold(const Level &cur, const Level &nxt, uint32_t i)
{
    const float4 fa = cur.fa[i], fb = cur.fb[i];
    V3 col = V3{fa.x, fa.y, fa.z};
    const int tr = __float_as_int(fb.y), rf = __float_as_int(fb.z);
    if (tr >= 0)
    {
        const float4 c = nxt.fa[tr];
        col.x = c.x * fa.w + col.x;
        col.y = c.y * fa.w + col.y;
        col.z = c.z * fa.w + col.z;
    }
    if (rf >= 0)
    {
        const float4 c = nxt.fa[rf];
        col.x = c.x * fb.x + col.x;
        col.y = c.y * fb.x + col.y;
        col.z = c.z * fb.x + col.z;
    }
    return col;
}

__global__ void __launch_bounds__(256) k_fold(Args a)
{
    const uint32_t n = min(a.cnt[CNT_LVL + a.level], a.cur.cap);
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    {
        const V3 c = fold(a.cur, a.nxt, i);
        const float4 fa = a.cur.fa[i];
        a.cur.fa[i] = make_float4(c.x, c.y, c.z, fa.w);
    }
}

/* level 0: fold, clamp, anti-aliasing resolve and ARGB8 packing
 * (6942-7150, 1922-1945) */
__global__ void __launch_bounds__(256) k_resolve(Args a)
{
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= a.fd.x_res || y >= a.fd.y_res) return;
    const uint32_t tile = (uint32_t)(y / TILE) * a.tiles_x + (uint32_t)(x / TILE);
    const uint32_t p = (uint32_t)(y % TILE) * TILE + (uint32_t)(x % TILE);
    const uint32_t i0 = (tile * (TILE * TILE) + p) * a.ns;

    V3 acc[4];
    for (int k = 0; k < a.ns; k++)
    {
        V3 c = fold(a.cur, a.nxt, i0 + k);
        c.x = c.x < 1.0f ? c.x : 1.0f;
        c.y = c.y < 1.0f ? c.y : 1.0f;
        c.z = c.z < 1.0f ? c.z : 1.0f;
        if (k == 0) acc[0] = c;
        else if (k == 1) acc[1] = c;
        else if (k == 2) acc[2] = c;
        else acc[3] = c;
    }
    V3 c;
    if (a.ns == 1)
    {
        c = acc[0];
    }
    else if (a.ns == 2)
    {
        c.x = acc[0].x * 0.5f + acc[1].x * 0.5f;
        c.y = acc[0].y * 0.5f + acc[1].y * 0.5f;
        c.z = acc[0].z * 0.5f + acc[1].z * 0.5f;
    }
    else
    {
        float p0, p1;
        p0 = acc[0].x * 0.5f + acc[1].x * 0.5f; p1 = acc[2].x * 0.5f + acc[3].x * 0.5f;
        c.x = p0 * 0.5f + p1 * 0.5f;
        p0 = acc[0].y * 0.5f + acc[1].y * 0.5f; p1 = acc[2].y * 0.5f + acc[3].y * 0.5f;
        c.y = p0 * 0.5f + p1 * 0.5f;
        p0 = acc[0].z * 0.5f + acc[1].z * 0.5f; p1 = acc[2].z * 0.5f + acc[3].z * 0.5f;
        c.z = p0 * 0.5f + p1 * 0.5f;
    }
    const float v[3] = {c.x, c.y, c.z};
    uint32_t px = 0;
    #pragma unroll
    for (int ch = 0; ch < 3; ch++)
    {
        float f = v[ch];
        if (a.fd.gamma & QR_PROP_GAMMA) f = sqrtf(f);
        f = f * 255.0f;
        px |= ((uint32_t)cvt_rne(f) & 0xFFu) << (16 - 8 * ch);
    }
    a.frame[(ptrdiff_t)y * a.fd.x_row + x] = px;
}

/* grow-only device allocation */
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

struct LevelBufs
{
    Buf ra, rb, rc, na, nb, nc, vis, fa, fb;
    uint32_t cap = 0;
};

} /* namespace */

struct qr_wf_state
{
    LevelBufs lv[MAX_LEVELS];
    uint32_t  want[MAX_LEVELS] = {};   /* capacity wanted per level */
    Buf       shq;
    uint32_t  sh_cap = 0, sh_want = 0;
    Buf       cnt;
    uint32_t *cnt_host = nullptr;      /* pinned copy of the counters */
    int       levels_used = 0;
    int       n_lgt = 0;
    int       grid = 2048;
};

int qr_wf_max_lights(void)
{
    return MAX_LIGHTS;
}

hipError_t qr_wf_create(qr_wf_state **out)
{
    *out = nullptr;
    qr_wf_state *s = new (std::nothrow) qr_wf_state();
    if (s == nullptr) return hipErrorOutOfMemory;
    hipError_t e = s->cnt.reserve(sizeof(uint32_t) * CNT_N);
    if (e == hipSuccess)
    {
        e = hipHostMalloc((void **)&s->cnt_host, sizeof(uint32_t) * CNT_N,
                          hipHostMallocDefault);
    }
    if (e != hipSuccess)
    {
        s->cnt.release();
        delete s;
        return e;
    }
    memset(s->cnt_host, 0, sizeof(uint32_t) * CNT_N);
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) ==
            hipSuccess && cus > 0)
    {
        s->grid = cus * 8;
    }
    *out = s;
    return hipSuccess;
}

void qr_wf_destroy(qr_wf_state *s)
{
    if (s == nullptr) return;
    for (LevelBufs &l : s->lv)
    {
        l.ra.release(); l.rb.release(); l.rc.release();
        l.na.release(); l.nb.release(); l.nc.release();
        l.vis.release(); l.fa.release(); l.fb.release();
    }
    s->shq.release();
    s->cnt.release();
    if (s->cnt_host != nullptr) (void)hipHostFree(s->cnt_host);
    delete s;
}

static hipError_t level_reserve(LevelBufs &l, uint32_t cap, bool rays)
{
    hipError_t e;
    const size_t f4 = sizeof(float4) * (size_t)cap;
    if (rays)
    {
        if ((e = l.ra.reserve(f4)) != hipSuccess) return e;
        if ((e = l.rb.reserve(f4)) != hipSuccess) return e;
        if ((e = l.rc.reserve(f4)) != hipSuccess) return e;
    }
    if ((e = l.na.reserve(f4)) != hipSuccess) return e;
    if ((e = l.nb.reserve(f4)) != hipSuccess) return e;
    if ((e = l.nc.reserve(f4)) != hipSuccess) return e;
    if ((e = l.vis.reserve(sizeof(uint32_t) * (size_t)cap)) != hipSuccess) return e;
    if ((e = l.fa.reserve(f4)) != hipSuccess) return e;
    if ((e = l.fb.reserve(f4)) != hipSuccess) return e;
    l.cap = cap;
    return hipSuccess;
}

static Level level_view(LevelBufs &l)
{
    Level v;
    v.ra = (float4 *)l.ra.ptr;
    v.rb = (float4 *)l.rb.ptr;
    v.rc = (float4 *)l.rc.ptr;
    v.na = (float4 *)l.na.ptr;
    v.nb = (float4 *)l.nb.ptr;
    v.nc = (float4 *)l.nc.ptr;
    v.vis = (uint32_t *)l.vis.ptr;
    v.fa = (float4 *)l.fa.ptr;
    v.fb = (float4 *)l.fb.ptr;
    v.cap = l.cap;
    return v;
}

static uint32_t round_cap(uint64_t n)
{
    uint64_t c = 256;
    while (c < n) c <<= 1;
    return (uint32_t)std::min<uint64_t>(c, 0x7fffff00u);
}

hipError_t qr_wf_launch(qr_wf_state *s, const qr_device_tables *t,
                        const qr_frame_desc *fd, uint32_t *frame, int32_t *ids,
                        hipStream_t stream)
{
    if (fd->depth < 0 || fd->depth > MAX_LEVELS - 2 || t->n_lgt > MAX_LIGHTS)
    {
        return hipErrorInvalidValue;
    }
    const int ns = fd->fsaa == 2 ? 4 : fd->fsaa == 1 ? 2 : 1;
    const uint32_t tiles_x = (uint32_t)(fd->x_res + TILE - 1) / TILE;
    const uint32_t tiles_y = (uint32_t)(fd->y_res + TILE - 1) / TILE;
    const uint64_t n0 = (uint64_t)tiles_x * tiles_y * TILE * TILE * ns;
    if (n0 > 0x7fffff00ull) return hipErrorInvalidValue;
    const int levels = fd->depth + 1;

    /* capacities: level 0 exact, deeper levels as last measured (at least
     * half of level 0), shadow jobs per light as large as level 0 */
    hipError_t e;
    if ((e = level_reserve(s->lv[0], (uint32_t)n0, false)) != hipSuccess) return e;
    for (int L = 1; L < levels; L++)
    {
        uint32_t want = std::max(s->want[L], round_cap(n0 / 2));
        if (s->lv[L].cap < want)
        {
            if ((e = level_reserve(s->lv[L], want, true)) != hipSuccess) return e;
        }
    }
    {
        uint32_t want = std::max(s->sh_want, (uint32_t)n0);
        if (s->sh_cap < want || s->n_lgt < t->n_lgt)
        {
            const size_t per = std::max(1, t->n_lgt);
            if ((e = s->shq.reserve(sizeof(uint32_t) * per * want)) != hipSuccess) return e;
            s->sh_cap = want;
            s->n_lgt = t->n_lgt;
        }
    }
    s->levels_used = levels;

    Args a;
    memset(&a, 0, sizeof(a));
    a.sc = scene_of(t);
    a.fd = *fd;
    a.cnt = (uint32_t *)s->cnt.ptr;
    a.shq = (uint32_t *)s->shq.ptr;
    a.sh_cap = s->sh_cap;
    a.n0 = (uint32_t)n0;
    a.tiles_x = tiles_x;
    a.ns = ns;
    a.ids = ids;
    a.frame = frame;

    hipLaunchKernelGGL(k_reset, dim3(1), dim3(256), 0, stream, a.cnt);
    for (int L = 0; L < levels; L++)
    {
        a.level = L;
        a.cur = level_view(s->lv[L]);
        a.nxt = L + 1 < levels ? level_view(s->lv[L + 1]) : Level{};
        const uint32_t cap = L == 0 ? (uint32_t)n0 : a.cur.cap;
        const uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>(
            (cap + 255) / 256, (uint32_t)s->grid));
        const uint32_t g0 = (uint32_t)((n0 + 255) / 256);
        if (L == 0)
        {
            hipLaunchKernelGGL(k_walk_prep<true>, dim3(g0), dim3(256), 0, stream, a);
        }
        else
        {
            hipLaunchKernelGGL(k_walk_prep<false>, dim3(g), dim3(256), 0, stream, a);
        }
        hipLaunchKernelGGL(k_shadow, dim3(s->grid), dim3(256), 0, stream, a);
        if (L == 0)
        {
            hipLaunchKernelGGL(k_final<true>, dim3(g0), dim3(256), 0, stream, a);
        }
        else
        {
            hipLaunchKernelGGL(k_final<false>, dim3(g), dim3(256), 0, stream, a);
        }
    }
    for (int L = levels - 2; L >= 1; L--)
    {
        a.level = L;
        a.cur = level_view(s->lv[L]);
        a.nxt = level_view(s->lv[L + 1]);
        const uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>(
            (a.cur.cap + 255) / 256, (uint32_t)s->grid));
        hipLaunchKernelGGL(k_fold, dim3(g), dim3(256), 0, stream, a);
    }
    a.level = 0;
    a.cur = level_view(s->lv[0]);
    a.nxt = levels > 1 ? level_view(s->lv[1]) : Level{};
    a.fd.x_row = fd->x_res;
    hipLaunchKernelGGL(k_resolve, dim3((fd->x_res + 15) / 16, (fd->y_res + 15) / 16),
                       dim3(256), 0, stream, a);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(s->cnt_host, s->cnt.ptr, sizeof(uint32_t) * CNT_N,
                          hipMemcpyDeviceToHost, stream);
}

int qr_wf_check(qr_wf_state *s, qr_wf_stats *stats)
{
    const uint32_t *c = s->cnt_host;
    if (stats != nullptr)
    {
        memset(stats, 0, sizeof(*stats));
        for (int L = 1; L < s->levels_used && L < MAX_LEVELS; L++)
        {
            stats->level_rays[L] = c[CNT_LVL + L];
        }
    }
    if (c[CNT_OVF] == 0) return 0;
    /* grow every queue that overflowed to its counted length */
    for (int L = 1; L < s->levels_used; L++)
    {
        if (c[CNT_LVL + L] > s->lv[L].cap) s->want[L] = round_cap(c[CNT_LVL + L]);
    }
    if (c[CNT_SHMAX] > s->sh_cap) s->sh_want = round_cap(c[CNT_SHMAX]);
    return 1;
}
