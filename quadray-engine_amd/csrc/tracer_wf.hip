/*
 * tracer_wf.hip -- level-by-level Whitted pipeline for MI355X (gfx950),
 * batched over frames.
 *
 * The Whitted tree of every camera sample is traced one tree level at a
 * time, over compacted queues of each level's rays.  Per level L:
 *
 *   nearest   nearest-hit walk of each ray; for a hit, the hit point, local
 *             hit and normal, one shadow job per light facing the hit
 *             (appended to that light's queue), and the hit's children (its
 *             refraction and reflection rays, which do not depend on its
 *             lighting), appended to level L+1's queue with their
 *             coefficients in the hit's fold record
 *   shadow    any-hit walk of each shadow job; an unoccluded light sets its
 *             bit in the hit's visibility mask
 *   shading   the hit's colour with the visibility mask (QuadRay's MT/LT
 *             arithmetic)
 *
 * Shadow walks of a level need its nearest walks, and its shading needs its
 * shadow walks, but nothing else orders the levels, so stage j is one launch
 * of the nearest walks of level j, the shadow walks of level j-1 and the
 * shading of level j-2: the small queues of the deeper levels run side by
 * side instead of one after another.  Shadow jobs go into three queue
 * sets: level 0 has one of its own (its shadow walks may run beside every
 * deeper stage), and deeper levels alternate between the other two (level
 * L's shadow walks read one while level L+1's nearest walks fill the other).
 *
 * Then the tree is folded bottom-up, child colour times its coefficient
 * plus the node's colour, refraction before reflection (QuadRay's TR_mix /
 * RF_mix order, core/tracer/tracer.cpp 4862-4936, 5277-5357), and level 0 is
 * resolved to pixels: levels deeper than 3 by a fold launch each, levels 3
 * to 1 by the resolve as it follows each camera sample's children.
 *
 * One launch of each stage serves a whole batch of frames, each with its own
 * scene tables, camera, resolution, anti-aliasing, depth and queues.  A
 * stage's workgroups form a global queue of (frame, range) tiles: a one-
 * workgroup plan kernel turns the frames' device-side queue lengths into
 * prefix sums of workgroups per frame, and the stage kernel, launched with a
 * fixed persistent grid, maps each virtual workgroup to its frame by binary
 * search.  Grids never depend on queue lengths, so a batch's launches can be
 * captured in a HIP graph and replayed.  Every workgroup belongs to one
 * frame, so scene tables stay wave-uniform scalar loads.
 *
 * Appends are counted per wave (ballot + prefix count) and per workgroup in
 * LDS, with one global atomic per workgroup and queue pair, so queues stay
 * dense and lanes stay busy between bounces.  Queues are preallocated; an
 * append beyond capacity only sets a flag and keeps counting, the host grows
 * the queue to the counted size and traces the frame again, so capacity never
 * changes a result.
 *
 * Two launch knobs, chosen by the batch autotuner with the tile shape:
 * rays per thread (a virtual workgroup takes that many consecutive blocks
 * of 256 queue entries, one after another), and the compaction cadence c:
 * the children of level L are compacted into dense queues only when L + 1
 * is a multiple of c; otherwise every ray of level L gets two slots of
 * level L + 1, live or dead, in its workgroup's order, and the dead ones
 * idle through the next level's walks.  Neither changes a result.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "qr_soa.h"
#include "tracer_dev.h"
#include "tracer_kernels.h"

namespace {

constexpr int TILE_PX = 64;         /* level-0 samples: tiles of 64 pixels */
constexpr int MAX_LEVELS = 16;      /* depth 0..14 plus one */
constexpr int MAX_LIGHTS = 32;      /* visibility mask bits */
constexpr int WG = 256;             /* threads per workgroup of every stage */
constexpr int WAVES = WG / 64;
/* ray flag of a slot with no ray, in a level that is not compacted */
constexpr int FLAG_DEAD = 8;
/* most lights whose visibility flags are bytes of the hit's word, written
 * with plain byte stores (measured per frame with 4 lights: 1.753 ms
 * against 1.760 with a device-scope atomicOr per visible light) */
constexpr int VIS_BYTE_LIGHTS = 4;

/* counter block per frame: [0] overflow flag, [1] largest shadow queue of the
 * frame, [2 + L] queue length of level L; the host reads it after a batch */
constexpr int CNT_OVF = 0;
constexpr int CNT_SHMAX = 1;
constexpr int CNT_LVL = 2;
constexpr int CNT_N = CNT_LVL + MAX_LEVELS;
/* shadow-job counters per frame: [L * MAX_LIGHTS + l] jobs of light l at
 * level L (device only) */
constexpr int SHC_N = MAX_LEVELS * MAX_LIGHTS;
static_assert(CNT_N % 2 == 0, "counter blocks stay 8-byte aligned");
static_assert(MAX_LIGHTS % 2 == 0 && SHC_N % 2 == 0,
              "shadow counters are updated in 64-bit pairs");

/* one tree level in HBM (structure of float4 arrays, one entry per ray) */
struct Level
{
    float4   *ra, *rb, *rc;     /* ray: org|origin, dir|flags, oloc (L >= 1) */
    float4   *na, *nb, *nc;     /* hit: point|srf<<1|side, loc|u, normal|v */
    uint32_t *vis;              /* visible lights: with at most VIS_BYTE_LIGHTS
                                   lights byte l of the word is light l's
                                   flag, with more bit l */
    float4   *fa, *fb;          /* fold: colour, c_rfl|tr|rf|c_trn */
    uint32_t  cap;
    uint32_t  pad;
};

/* shadow-job queue set of level L: level 0 has one of its own, so its
 * shadow walks may run until the resolve; deeper levels alternate between
 * two (level L's walks read one while level L+1's nearest walks fill the
 * other) */
__host__ __device__ constexpr int shadow_set(int L)
{
    return L == 0 ? 0 : 1 + (L & 1);
}

/* per-frame arguments of the stage kernels, in device memory */
struct Frame
{
    Scene          sc;
    qr_frame_desc  fd;
    Level          lv[MAX_LEVELS];  /* levels >= "levels" are empty */
    uint32_t      *cnt;             /* counter block, CNT_N words */
    uint32_t      *shc;             /* shadow-job counters, SHC_N words */
    uint32_t      *shq;             /* shadow jobs of level L, light l at
                                       [(shadow_set(L) * sc.n_lgt + l) * sh_cap] */
    uint32_t       sh_cap;
    uint32_t       n0;              /* level-0 samples: the chunk's tiles */
    uint32_t       tile0;           /* the chunk: tiles tile0 .. tile0 + ntiles - 1 */
    uint32_t       ntiles;
    uint32_t       tiles_x;
    int32_t        ns;              /* samples per pixel */
    int32_t        levels;          /* depth + 1 */
    int32_t        tw_log2;         /* tile width 2^tw_log2, height 64 / width */
    uint32_t       sparse;          /* bit L: level L is not compacted (two
                                       slots per ray of level L - 1) */
    int32_t       *ids;
    uint32_t      *frame;           /* ARGB8, row pitch fd.x_res */
};

/* kernel arguments: the batch and the stage's plan */
struct Batch
{
    const QC Frame    *fr;
    const QC uint32_t *plan;        /* nfr + 1 prefix sums of workgroups */
    int32_t            nfr;
    int32_t            level;       /* level of a fold or level plan */
    int32_t            wl, sl, fl;  /* a stage's levels of nearest walks,
                                       shadow walks and shading (-1: none) */
    int32_t            rpt;         /* blocks of WG entries per virtual
                                       workgroup (rays per thread) */
};

/* plan kinds */
constexpr int PLAN_LEVEL = 0;       /* rays of level L (level 0's walk and
                                       resolve, fold) */
constexpr int PLAN_STAGE = 1;       /* a stage: rays of level wl, shadow
                                       jobs of level sl, rays of level fl */
/* plan arrays: levels 0..MAX_LEVELS-1, then stages 0..MAX_STAGES-1 */
constexpr int MAX_STAGES = 2 * MAX_LEVELS + 2;
constexpr int PLAN_SLOTS = MAX_LEVELS + MAX_STAGES;

constexpr int plan_slot(int kind, int L)
{
    return kind == PLAN_LEVEL ? L : MAX_LEVELS + L;
}

/* copy of a record read through the constant address space (wave-uniform
 * address: scalar loads) */
template <class T>
__device__ __forceinline__ T ldc(const QC T &x)
{
    T v;
    __builtin_memcpy(&v, (const T *)&x, sizeof(T));
    return v;
}

__device__ __forceinline__ uint32_t wgs(uint32_t n)
{
    return (n + WG - 1) / WG;
}

/* virtual workgroups of n entries, rpt blocks of WG entries each */
__device__ __forceinline__ uint32_t vwgs(uint32_t n, int rpt)
{
    return (wgs(n) + (uint32_t)rpt - 1u) / (uint32_t)rpt;
}

__device__ __forceinline__ uint32_t level_len(const QC Frame &a, int L)
{
    if (L < 0 || L >= a.levels) return 0u;
    if (L == 0) return a.n0;
    return min(a.cnt[CNT_LVL + L], a.lv[L].cap);
}

/* virtual workgroups of the shadow jobs of level L, all lights */
__device__ __forceinline__ uint32_t shadow_wgs(const QC Frame &a, int L, int rpt)
{
    if (L < 0 || L >= a.levels) return 0u;
    uint32_t s = 0;
    for (int li = 0; li < a.sc.n_lgt; li++)
    {
        s += vwgs(min(a.shc[L * MAX_LIGHTS + li], a.sh_cap), rpt);
    }
    return s;
}

__device__ __forceinline__ uint32_t plan_count(const QC Frame &a, int kind, int L,
                                              const Batch &b)
{
    if (kind == PLAN_LEVEL) return vwgs(level_len(a, L), b.rpt);
    return shadow_wgs(a, b.sl, b.rpt) + vwgs(level_len(a, b.wl), b.rpt) +
           vwgs(level_len(a, b.fl), b.rpt);
}

/* prefix sums of workgroups per frame for one plan, by the whole workgroup */
__device__ void scan_plan(const Batch &b, int kind, int L, uint32_t *plan, uint32_t *part)
{
    const int t = threadIdx.x;
    const int per = (b.nfr + 1023) / 1024;
    const int f0 = min(t * per, b.nfr), f1 = min(f0 + per, b.nfr);
    uint32_t sum = 0;
    for (int f = f0; f < f1; f++) sum += plan_count(b.fr[f], kind, L, b);
    __syncthreads();    /* the previous scan has read "part" */
    part[t] = sum;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1)
    {
        const uint32_t v = t >= d ? part[t - d] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;           /* exclusive prefix */
    for (int f = f0; f < f1; f++)
    {
        plan[f] = run;
        run += plan_count(b.fr[f], kind, L, b);
    }
    if (t == 1023) plan[b.nfr] = part[1023];
}

/* the plans of one or more launches (one workgroup of 1024): "count" plans
 * of levels b.level, b.level + 1, ... at "stride" words from each other */
__global__ void __launch_bounds__(1024) k_plan(Batch b, int kind, uint32_t *plan, int count,
                                               uint32_t stride)
{
    __shared__ uint32_t part[1024];
    for (int k = 0; k < count; k++) scan_plan(b, kind, b.level + k, plan + k * stride, part);
}

/* frame of virtual workgroup vb: plan[f] <= vb < plan[f + 1] */
__device__ __forceinline__ int plan_frame(const QC uint32_t *plan, int nfr, uint32_t vb)
{
    int lo = 0, hi = nfr;
    while (hi - lo > 1)
    {
        const int mid = (lo + hi) >> 1;
        if (plan[mid] <= vb) lo = mid; else hi = mid;
    }
    return lo;
}

/* A launch's virtual workgroups: with a plan, its total and each frame's
 * first workgroup; without one (a batch of one frame), counted by each
 * workgroup from the frame's counters, which no kernel of the launch
 * changes. */
__device__ __forceinline__ uint32_t launch_total(const Batch &b, int kind)
{
    return b.plan != nullptr ? b.plan[b.nfr] : plan_count(b.fr[0], kind, b.level, b);
}
__device__ __forceinline__ int launch_frame(const Batch &b, uint32_t vb)
{
    return b.plan != nullptr ? plan_frame(b.plan, b.nfr, vb) : 0;
}
__device__ __forceinline__ uint32_t launch_first(const Batch &b, int f)
{
    return b.plan != nullptr ? b.plan[f] : 0u;
}

__device__ __forceinline__ void overflow(uint32_t *cnt)
{
    atomicOr(&cnt[CNT_OVF], 1u);
}

/* Workgroup-wide append to up to MAX_LIGHTS queues whose counters are
 * consecutive 32-bit words starting 8-byte aligned.  Waves count their lanes
 * in LDS; then one 64-bit atomic per pair of queues and workgroup reserves
 * the workgroup's slots (two 32-bit totals never carry into each other: each
 * stays below 2^32).  Every thread of the workgroup calls it with bit q of
 * "mask" set for one slot in queue q, and reads its slot with block_slot(). */
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

/* slot in queue q of a lane with bit q of mask set; every lane of the wave
 * calls it */
__device__ __forceinline__ uint32_t block_slot(const BlockAppend &b, uint32_t mask, int q)
{
    const uint64_t m = __ballot((mask >> q) & 1u);
    const int lane = __lane_id();
    const uint64_t below = lane == 0 ? 0ull : (m & (~0ull >> (64 - lane)));
    return b.base[q] + b.woff[threadIdx.x >> 6][q] + (uint32_t)__popcll(below);
}

/* Workgroup-wide append of up to two slots per lane (bit 0 and bit 1 of
 * mask) to one queue: one 32-bit atomic per workgroup.  The workgroup's
 * bit-0 slots come first, wave by wave, then its bit-1 slots, so the next
 * level's waves hold rays of one kind (refraction or reflection) more often.
 * Every thread of the workgroup calls it; pair_slot() gives a lane's slots. */
struct PairAppend
{
    uint32_t base;
    uint32_t off[2][WAVES];             /* wave's first slot per bit */
};

__device__ __forceinline__ void pair_append(PairAppend &b, uint32_t mask, uint32_t *counter)
{
    const int wv = threadIdx.x >> 6;
    __syncthreads();    /* the previous call's slots have been read */
    const uint64_t m0 = __ballot(mask & 1u), m1 = __ballot((mask >> 1) & 1u);
    if (__lane_id() == 0)
    {
        b.off[0][wv] = (uint32_t)__popcll(m0);
        b.off[1][wv] = (uint32_t)__popcll(m1);
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        uint32_t run = 0;
        for (int k = 0; k < 2; k++)
        {
            for (int v = 0; v < WAVES; v++)
            {
                const uint32_t c = b.off[k][v];
                b.off[k][v] = run;
                run += c;
            }
        }
        b.base = run != 0 ? atomicAdd(counter, run) : 0u;
    }
    __syncthreads();
}

/* slot of bit k (0 or 1) of a lane's mask; every lane of the wave calls it */
__device__ __forceinline__ uint32_t pair_slot(const PairAppend &b, uint32_t mask, int k)
{
    const uint64_t m = __ballot((mask >> k) & 1u);
    const int lane = __lane_id();
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    return b.base + b.off[k][threadIdx.x >> 6] + (uint32_t)__popcll(m & below);
}

/* level-0 sample i of the chunk -> pixel and sub-sample: tiles of 64 pixels
 * in row-major order, pixels row-major in a tile, samples innermost */
__device__ __forceinline__ bool sample_pixel(const QC Frame &a, uint32_t i, int &x,
                                             int &y, int &k)
{
    const uint32_t ns = (uint32_t)a.ns;
    const int tw = a.tw_log2;
    const int th = 6 - tw;
    const uint32_t per_tile = TILE_PX * ns;
    const uint32_t lt = i / per_tile;
    const uint32_t rem = i - lt * per_tile;
    const uint32_t tile = a.tile0 + lt;
    const uint32_t p = rem / ns;
    k = (int)(rem - p * ns);
    x = (int)(((tile % a.tiles_x) << tw) + (p & ((1u << tw) - 1u)));
    y = (int)(((tile / a.tiles_x) << th) + (p >> tw));
    return x < a.fd.x_res && y < a.fd.y_res;
}

/* camera ray of a sub-sample (render_slice, engine.cpp:3514-3697 and the
 * 2X / 4X offsets of 6942-7150) */
__device__ __forceinline__ Ray camera_ray(const QC qr_frame_desc &fd, int x, int y, int k)
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

/* ray i of level L; false for padding samples outside the frame and (with
 * SPARSE) for slots with no ray */
template <bool L0, bool SPARSE = false>
__device__ __forceinline__ bool level_ray(const QC Frame &a, const Level &cur, int L,
                                          uint32_t i, Ray &r)
{
    if (L0)
    {
        int x, y, k;
        const bool in = sample_pixel(a, i, x, y, k);
        r = camera_ray(a.fd, x, y, k);
        return in;
    }
    const float4 ra = cur.ra[i], rb = cur.rb[i], rc = cur.rc[i];
    r.org = V3{ra.x, ra.y, ra.z};
    r.origin = __float_as_int(ra.w);
    r.dir = V3{rb.x, rb.y, rb.z};
    r.flags = __float_as_int(rb.w);
    r.oloc = V3{rc.x, rc.y, rc.z};
    r.t_min = 0.0f;
    r.t_max = 3.402823466e+38f;
    r.depth = a.fd.depth - L;
    /* a slot of a level that is not compacted, with no ray */
    return !SPARSE || (r.flags & FLAG_DEAD) == 0;
}

__global__ void __launch_bounds__(WG) k_reset(Batch b)
{
    for (int f = blockIdx.x; f < b.nfr; f += gridDim.x)
    {
        uint32_t *cnt = b.fr[f].cnt;
        uint32_t *shc = b.fr[f].shc;
        for (int i = threadIdx.x; i < CNT_N; i += WG) cnt[i] = 0u;
        for (int i = threadIdx.x; i < SHC_N; i += WG) shc[i] = 0u;
    }
}

/* nearest hit of ray i of level L: hit record, shadow jobs, children and the
 * fold record's coefficients and child slots; every thread of the workgroup
 * calls it.  SPARSE: some levels may be left uncompacted (a.sparse). */
template <bool L0, bool SPARSE>
__device__ __forceinline__ void nearest(const QC Frame &a, int L, uint32_t i,
                                        BlockAppend &ba, PairAppend &pa)
{
    const Scene sc = ldc(a.sc);
    const Level cur = ldc(a.lv[L]);
    const uint32_t n = level_len(a, L);

    Ray r;
    bool valid = false;
    if (i < n) valid = level_ray<L0, SPARSE>(a, cur, L, i, r);

    Hit h;
    h.srf = -1;
    bool hit = false;
    if (valid) hit = walk<false>(sc, r, h);

    V3 nrm = V3{0.0f, 0.0f, 0.0f};
    if (hit)
    {
        hit_info(sc, r, h);
        nrm = normal_at(sc, sc.srf[h.srf], h);
        cur.na[i] = make_float4(h.hit.x, h.hit.y, h.hit.z,
                                __int_as_float((h.srf << 1) | h.side));
        cur.nb[i] = make_float4(h.loc.x, h.loc.y, h.loc.z, h.u);
        cur.nc[i] = make_float4(nrm.x, nrm.y, nrm.z, h.v);
        cur.vis[i] = 0u;
    }
    else if (i < n)
    {
        cur.na[i] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
    }
    if (L0 && a.ids != nullptr && valid)
    {
        int x, y, k;
        sample_pixel(a, i, x, y, k);
        a.ids[((size_t)y * a.fd.x_res + x) * a.ns + k] =
            hit ? (sc.srf[h.srf].id << 1) | h.side : -1;
    }

    /* one shadow job per light in front of a lit hit (shade's gate) */
    const bool lit = hit && !(sc.srf[h.srf].props[h.side] & QR_PROP_LIGHT);
    uint32_t need = 0u;
    if (lit)
    {
        for (int li = 0; li < sc.n_lgt; li++)
        {
            V3 Lv;
            if (0.0f < light_vec(sc.lgt[li], h.hit, nrm, Lv)) need |= 1u << li;
        }
    }
    block_append(ba, need, sc.n_lgt, &a.shc[L * MAX_LIGHTS]);
    uint32_t *shq = a.shq + (size_t)shadow_set(L) * sc.n_lgt * a.sh_cap;
    for (int li = 0; li < sc.n_lgt; li++)
    {
        const uint32_t slot = block_slot(ba, need, li);
        if ((need >> li) & 1u)
        {
            if (slot < a.sh_cap) shq[(size_t)li * a.sh_cap + slot] = i;
            else overflow(a.cnt);
        }
    }

    /* the children: rays of level L+1 */
    Node nd;
    nd.todo = 0;
    if (hit) node_children(sc, r, h, nrm, nd);
    const bool tr = nd.todo & 1;
    const bool rf = nd.todo & 2;
    const uint32_t kids = (tr ? 1u : 0u) | (rf ? 2u : 0u);
    /* a level that is not compacted gives every ray two slots */
    const bool dense = !SPARSE || ((a.sparse >> (L + 1)) & 1u) == 0u;
    const uint32_t slots = dense ? kids : (i < n ? 3u : 0u);
    pair_append(pa, slots, &a.cnt[CNT_LVL + L + 1]);
    int s_tr = (int)pair_slot(pa, slots, 0);
    int s_rf = (int)pair_slot(pa, slots, 1);
    if (slots != 0u)
    {
        const Level nxt = ldc(a.lv[L + 1]);
        if (slots & 1u)
        {
            if ((uint32_t)s_tr >= nxt.cap)
            {
                overflow(a.cnt);
                s_tr = -1;
            }
            else if (tr)
            {
                nxt.ra[s_tr] = make_float4(nd.org.x, nd.org.y, nd.org.z,
                                           __int_as_float(nd.srf));
                nxt.rb[s_tr] = make_float4(nd.tdir.x, nd.tdir.y, nd.tdir.z,
                                           __int_as_float(nd.side | FLAG_PASS));
                nxt.rc[s_tr] = make_float4(nd.loc.x, nd.loc.y, nd.loc.z, 0.0f);
            }
            else
            {
                nxt.rb[s_tr] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(FLAG_DEAD));
            }
        }
        if (slots & 2u)
        {
            if ((uint32_t)s_rf >= nxt.cap)
            {
                overflow(a.cnt);
                s_rf = -1;
            }
            else if (rf)
            {
                nxt.ra[s_rf] = make_float4(nd.org.x, nd.org.y, nd.org.z,
                                           __int_as_float(nd.srf));
                nxt.rb[s_rf] = make_float4(nd.rdir.x, nd.rdir.y, nd.rdir.z,
                                           __int_as_float(nd.side));
                nxt.rc[s_rf] = make_float4(nd.loc.x, nd.loc.y, nd.loc.z, 0.0f);
            }
            else
            {
                nxt.rb[s_rf] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(FLAG_DEAD));
            }
        }
    }
    if (i < n)
    {
        if (hit)
        {
            cur.fb[i] = make_float4(nd.c_rfl, __int_as_float(tr ? s_tr : -1),
                                    __int_as_float(rf ? s_rf : -1), nd.c_trn);
        }
        else
        {
            cur.fb[i] = make_float4(0.0f, __int_as_float(-1), __int_as_float(-1), 0.0f);
        }
    }
}

/* any-hit walk of shadow job j of light li at level L: visibility bit */
__device__ __forceinline__ void shadow_job(const QC Frame &a, int L, int li, uint32_t j)
{
    const Scene sc = ldc(a.sc);
    const Level cur = ldc(a.lv[L]);
    const QC qr_light &l = sc.lgt[li];
    const uint32_t i = a.shq[((size_t)shadow_set(L) * sc.n_lgt + li) * a.sh_cap + j];
    const float4 na = cur.na[i], nb = cur.nb[i], nc = cur.nc[i];
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
    if (!walk<true>(sc, sr, sh))
    {
        /* a (hit, light) pair has one shadow job, so with a byte per light
         * each byte has one writer and needs no atomic */
        if (sc.n_lgt <= VIS_BYTE_LIGHTS) ((uint8_t *)&cur.vis[i])[li] = 1u;
        else atomicOr(&cur.vis[i], 1u << li);
    }
}

/* any-hit walks of shadow jobs "(local * rpt + k) * WG + thread", k < rpt,
 * of level L: a frame's virtual workgroups cover its lights' queues one
 * after the other */
__device__ __forceinline__ void shadow(const QC Frame &a, int L, uint32_t local, int rpt)
{
    const int nl = a.sc.n_lgt;
    int li = 0;
    uint32_t n = 0, raw = 0;
    for (; li < nl; li++)
    {
        raw = a.shc[L * MAX_LIGHTS + li];
        n = min(raw, a.sh_cap);
        const uint32_t w = vwgs(n, rpt);
        if (local < w) break;
        local -= w;
    }
    if (li >= nl) return;
    /* the largest queue, for the host's growth of the queues */
    if (local == 0 && threadIdx.x == 0) atomicMax(&a.cnt[CNT_SHMAX], raw);
    for (int k = 0; k < rpt; k++)
    {
        const uint32_t j = (local * (uint32_t)rpt + (uint32_t)k) * WG + threadIdx.x;
        if (j < n) shadow_job(a, L, li, j);
    }
}

/* colour of ray i of level L with its visibility mask (0 for a miss) */
template <bool L0>
__device__ __forceinline__ V3 hit_color(const QC Frame &a, const Level &cur, int L, uint32_t i)
{
    const float4 na = cur.na[i];
    const int code = __float_as_int(na.w);
    V3 col = V3{0.0f, 0.0f, 0.0f};
    if (code >= 0)
    {
        const Scene sc = ldc(a.sc);
        Ray r;
        level_ray<L0>(a, cur, L, i, r);
        const float4 nb = cur.nb[i], nc = cur.nc[i];
        Hit h;
        h.srf = code >> 1;
        h.side = code & 1;
        h.hit = V3{na.x, na.y, na.z};
        h.loc = V3{nb.x, nb.y, nb.z};
        h.u = nb.w;
        h.v = nc.w;
        const V3 nrm = V3{nc.x, nc.y, nc.z};
        const uint32_t vis = cur.vis[i];
        const qr_frame_desc fd = ldc(a.fd);
        col = shade_color(sc, fd, r, h, nrm, [&](int li, const Ray &) {
            const int bit = sc.n_lgt <= VIS_BYTE_LIGHTS ? 8 * li : li;
            return ((vis >> bit) & 1u) == 0u;
        });
    }
    return col;
}

/* shading of ray i of level L >= 1 into its fold record (level 0 is shaded
 * by the resolve) */
__device__ __forceinline__ void shading(const QC Frame &a, int L, uint32_t i)
{
    if (i >= level_len(a, L)) return;
    const Level cur = ldc(a.lv[L]);
    const V3 col = hit_color<false>(a, cur, L, i);
    cur.fa[i] = make_float4(col.x, col.y, col.z, 0.0f);
}

/* stage 0: nearest walks of the camera samples */
/* six waves per SIMD (80 VGPRs, 6 spilled; 1% faster per frame than five) */
template <int R>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(6, 6))) k_walk0(Batch b)
{
    const int rpt = R != 0 ? R : b.rpt;    /* rays per thread; R = 1: every level dense */
    __shared__ BlockAppend ba;
    __shared__ PairAppend pa;
    const uint32_t total = launch_total(b, PLAN_LEVEL);
    for (uint32_t vb = blockIdx.x; vb < total; vb += gridDim.x)
    {
        const int f = launch_frame(b, vb);
        const QC Frame &a = b.fr[f];
        const uint32_t blk = (vb - launch_first(b, f)) * (uint32_t)rpt;
        for (int k = 0; k < rpt; k++)
        {
            nearest<true, R == 0>(a, 0, (blk + (uint32_t)k) * WG + threadIdx.x, ba, pa);
        }
    }
}

/* a stage of shadow walks alone (level b.sl) */
/* eight waves per SIMD (64 VGPRs, 3 spilled outside the inner loops): the
 * extra waves hide more of the walk's scalar-load latency, and leave room
 * for level 1's nearest walks beside it (measured per frame: 1.2% faster
 * than seven waves, 2.5% faster than six) */
template <int R>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(8, 8))) k_shadow(Batch b)
{
    const int rpt = R != 0 ? R : b.rpt;    /* rays per thread; R = 1: every level dense */
    const uint32_t total = launch_total(b, PLAN_STAGE);
    for (uint32_t vb = blockIdx.x; vb < total; vb += gridDim.x)
    {
        const int f = launch_frame(b, vb);
        shadow(b.fr[f], b.sl, vb - launch_first(b, f), rpt);
    }
}

/* a stage of nearest walks alone (level b.wl >= 1) */
/* five waves per SIMD (93 VGPRs) */
template <int R>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(5, 5))) k_walk(Batch b)
{
    const int rpt = R != 0 ? R : b.rpt;    /* rays per thread; R = 1: every level dense */
    __shared__ BlockAppend ba;
    __shared__ PairAppend pa;
    const uint32_t total = launch_total(b, PLAN_STAGE);
    for (uint32_t vb = blockIdx.x; vb < total; vb += gridDim.x)
    {
        const int f = launch_frame(b, vb);
        const uint32_t blk = (vb - launch_first(b, f)) * (uint32_t)rpt;
        for (int k = 0; k < rpt; k++)
        {
            nearest<false, R == 0>(b.fr[f], b.wl, (blk + (uint32_t)k) * WG + threadIdx.x, ba, pa);
        }
    }
}

/* a stage of shading alone (level b.fl >= 1) */
template <int R>
__global__ void __launch_bounds__(WG) k_shading(Batch b)
{
    const int rpt = R != 0 ? R : b.rpt;    /* rays per thread; R = 1: every level dense */
    const uint32_t total = launch_total(b, PLAN_STAGE);
    for (uint32_t vb = blockIdx.x; vb < total; vb += gridDim.x)
    {
        const int f = launch_frame(b, vb);
        const uint32_t blk = (vb - launch_first(b, f)) * (uint32_t)rpt;
        for (int k = 0; k < rpt; k++)
        {
            shading(b.fr[f], b.fl, (blk + (uint32_t)k) * WG + threadIdx.x);
        }
    }
}

/* a stage of two or three kinds of work: per frame, the shadow walks of
 * level b.sl, then the nearest walks of level b.wl >= 1, then the shading of
 * level b.fl >= 1 (shadow walks first measured 1.7% faster per frame) */
/* seven waves per SIMD (72 VGPRs, 11 spilled): the deeper levels' short
 * queues are latency bound (measured per frame: 1.775 ms at seven, 1.816
 * at six, 1.807 at eight with 29 spilled; five was 4% slower than six) */
template <int R>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(7, 7))) k_stage(Batch b)
{
    const int rpt = R != 0 ? R : b.rpt;    /* rays per thread; R = 1: every level dense */
    __shared__ BlockAppend ba;
    __shared__ PairAppend pa;
    const uint32_t total = launch_total(b, PLAN_STAGE);
    for (uint32_t vb = blockIdx.x; vb < total; vb += gridDim.x)
    {
        const int f = launch_frame(b, vb);
        const QC Frame &a = b.fr[f];
        uint32_t local = vb - launch_first(b, f);
        const uint32_t ns = shadow_wgs(a, b.sl, rpt);
        if (local < ns)
        {
            shadow(a, b.sl, local, rpt);
            continue;
        }
        local -= ns;
        const uint32_t nw = vwgs(level_len(a, b.wl), rpt);
        const bool walks = local < nw;
        if (!walks) local -= nw;
        for (int k = 0; k < rpt; k++)
        {
            const uint32_t i = (local * (uint32_t)rpt + (uint32_t)k) * WG + threadIdx.x;
            if (walks) nearest<false, R == 0>(a, b.wl, i, ba, pa);
            else shading(a, b.fl, i);
        }
    }
}

/* node colour with its children's results folded in */
__device__ __forceinline__ V3 fold_children(V3 col, const float4 &fb, const Level &nxt)
{
    const int tr = __float_as_int(fb.y), rf = __float_as_int(fb.z);
    if (tr >= 0)
    {
        const float4 c = nxt.fa[tr];
        col.x = c.x * fb.w + col.x;
        col.y = c.y * fb.w + col.y;
        col.z = c.z * fb.w + col.z;
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

__device__ __forceinline__ V3 fold(const Level &cur, const Level &nxt, uint32_t i)
{
    const float4 fa = cur.fa[i];
    return fold_children(V3{fa.x, fa.y, fa.z}, cur.fb[i], nxt);
}

/* levels below level 0 that the resolve folds itself, reading the level
 * below them as folded by k_fold (every level of a frame of depth 4 or
 * less, so no fold launch; measured per frame at depth 4: 1.787 ms with a
 * fold launch per level, 1.775 with two levels folded here, 1.766 with
 * three) */
constexpr int FOLD_INLINE = 3;

/* node colour "col" of level L with fold record fb and its subtree folded
 * in: its children at level L + 1 folded here down to N levels below L, the
 * levels under those read as folded (the same operations, in the same
 * order, as k_fold's level by level) */
template <int N>
__device__ __forceinline__ V3 fold_tree(const QC Frame &a, int L, V3 col, const float4 &fb)
{
    if constexpr (N <= 1)
    {
        return fold_children(col, fb, ldc(a.lv[L + 1]));
    }
    else
    {
        const int tr = __float_as_int(fb.y), rf = __float_as_int(fb.z);
        if (tr >= 0 || rf >= 0)
        {
            const Level nxt = ldc(a.lv[L + 1]);
            if (tr >= 0)
            {
                const float4 fa = nxt.fa[tr];
                const V3 c = fold_tree<N - 1>(a, L + 1, V3{fa.x, fa.y, fa.z}, nxt.fb[tr]);
                col.x = c.x * fb.w + col.x;
                col.y = c.y * fb.w + col.y;
                col.z = c.z * fb.w + col.z;
            }
            if (rf >= 0)
            {
                const float4 fa = nxt.fa[rf];
                const V3 c = fold_tree<N - 1>(a, L + 1, V3{fa.x, fa.y, fa.z}, nxt.fb[rf]);
                col.x = c.x * fb.x + col.x;
                col.y = c.y * fb.x + col.y;
                col.z = c.z * fb.x + col.z;
            }
        }
        return col;
    }
}

template <int R>
__global__ void __launch_bounds__(WG) k_fold(Batch b)
{
    const int rpt = R != 0 ? R : b.rpt;    /* rays per thread; R = 1: every level dense */
    const int L = b.level;
    const uint32_t total = launch_total(b, PLAN_LEVEL);
    for (uint32_t vb = blockIdx.x; vb < total; vb += gridDim.x)
    {
        const int f = launch_frame(b, vb);
        const QC Frame &a = b.fr[f];
        const uint32_t n = level_len(a, L);
        const uint32_t blk = (vb - launch_first(b, f)) * (uint32_t)rpt;
        const Level cur = ldc(a.lv[L]), nxt = ldc(a.lv[L + 1]);
        for (int k = 0; k < rpt; k++)
        {
            const uint32_t i = (blk + (uint32_t)k) * WG + threadIdx.x;
            if (i >= n) continue;
            const V3 c = fold(cur, nxt, i);
            cur.fa[i] = make_float4(c.x, c.y, c.z, 0.0f);
        }
    }
}

/* level 0: shading of the camera samples, fold, clamp, anti-aliasing
 * resolve and ARGB8 packing (6942-7150, 1922-1945): one sample per thread, a
 * pixel's samples in neighbouring lanes (samples are innermost) */
template <int R>
__global__ void __launch_bounds__(WG) k_resolve(Batch b)
{
    const int rpt = R != 0 ? R : b.rpt;    /* rays per thread; R = 1: every level dense */
    const uint32_t total = launch_total(b, PLAN_LEVEL);
    for (uint32_t vb = blockIdx.x; vb < total; vb += gridDim.x)
    {
        const int f = launch_frame(b, vb);
        const QC Frame &a = b.fr[f];
        const uint32_t blk = (vb - launch_first(b, f)) * (uint32_t)rpt;
        for (int kb = 0; kb < rpt; kb++)
        {
            const uint32_t i = (blk + (uint32_t)kb) * WG + threadIdx.x;
            int x = 0, y = 0, k = 0;
            bool in = false;
            V3 c = V3{0.0f, 0.0f, 0.0f};
            if (i < a.n0)
            {
                in = sample_pixel(a, i, x, y, k);
                const Level cur = ldc(a.lv[0]);
                c = fold_tree<FOLD_INLINE + 1>(a, 0, hit_color<true>(a, cur, 0, i), cur.fb[i]);
                c.x = c.x < 1.0f ? c.x : 1.0f;
                c.y = c.y < 1.0f ? c.y : 1.0f;
                c.z = c.z < 1.0f ? c.z : 1.0f;
            }
            const qr_frame_desc fd = ldc(a.fd);
            const uint32_t px = resolve_lanes(fd, a.ns, c);
            if (in && k == 0) a.frame[(ptrdiff_t)y * fd.x_res + x] = px;
        }
    }
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

uint32_t round_cap(uint64_t n)
{
    uint64_t c = 256;
    while (c < n) c <<= 1;
    return (uint32_t)std::min<uint64_t>(c, 0x7fffff00u);
}

hipError_t level_reserve(LevelBufs &l, uint32_t cap, bool rays)
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

Level level_view(const LevelBufs &l)
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
    v.pad = 0;
    return v;
}

int device_cus()
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) ==
            hipSuccess && cus > 0)
    {
        return cus;
    }
    return 256;
}

} /* namespace */

/* queues of one frame, reused by every frame traced with it */
struct qr_wf_state
{
    LevelBufs lv[MAX_LEVELS];
    uint32_t  want[MAX_LEVELS] = {};   /* capacity wanted per level */
    Buf       shq;
    uint32_t  sh_cap = 0, sh_want = 0;
    int       n_lgt = 0;
    qr_wf_batch  *own = nullptr;       /* batch of one for qr_wf_launch */
    qr_frame_desc last_fd;             /* its frame */
};

/* launch context of a batch: frame arguments, plans, counters */
struct qr_wf_batch
{
    Buf       frames;                  /* Frame[nfr] on the device */
    Frame    *host = nullptr;          /* pinned staging of the frame arguments */
    int       host_cap = 0;
    std::vector<Frame> last;           /* arguments on the device now */
    hipEvent_t staged = nullptr;       /* the staging buffer's last copy */
    Buf       plans;
    Buf       cnt;                     /* CNT_N words per frame */
    Buf       shc;                     /* SHC_N words per frame */
    uint32_t *cnt_host = nullptr;      /* pinned copy after the batch */
    int       cnt_cap = 0;
    int       nfr = 0;
    int       levels = 0;
    int       grid = 16384;            /* persistent workgroups per stage */
    int       tw_log2 = 3;             /* level-0 tile width 8, height 8 */
    int       grid_per_cu = 64;        /* measured on a 1080p frame: 2.28 ms at
                                          64, 2.34 at 32, 2.32 at 128 */
    int       rpt = 1;                 /* blocks per virtual workgroup */
    int       cadence = 1;             /* levels compacted: multiples of it */
    int       cus = 256;
    hipStream_t side = nullptr;        /* the device's second stream (shared),
                                          set at the first overlapped launch */
    hipEvent_t fork = nullptr, join = nullptr;
};

/* One second stream per device, shared by every batch on it, made when a
 * batch first overlaps stages: launches are enqueued in order on the device
 * streams, and each batch orders its side work with events of its own, so
 * sharing only serialises side work that the device would run in that order
 * anyway.  The streams live until the process ends. */
static std::mutex g_side_mu;
static hipStream_t g_side[64];

static hipError_t side_stream(qr_wf_batch *b)
{
    if (b->side != nullptr) return hipSuccess;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    if (b->fork == nullptr &&
        (e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming)) != hipSuccess)
    {
        return e;
    }
    if (b->join == nullptr &&
        (e = hipEventCreateWithFlags(&b->join, hipEventDisableTiming)) != hipSuccess)
    {
        return e;
    }
    std::lock_guard<std::mutex> lk(g_side_mu);
    if (g_side[dev] == nullptr &&
        (e = hipStreamCreateWithFlags(&g_side[dev], hipStreamNonBlocking)) != hipSuccess)
    {
        return e;
    }
    b->side = g_side[dev];
    return hipSuccess;
}

int qr_wf_side_streams(void)
{
    std::lock_guard<std::mutex> lk(g_side_mu);
    int n = 0;
    for (hipStream_t st : g_side) n += st != nullptr;
    return n;
}

int qr_wf_max_lights(void)
{
    return MAX_LIGHTS;
}

hipError_t qr_wf_batch_create(qr_wf_batch **out)
{
    *out = nullptr;
    qr_wf_batch *b = new (std::nothrow) qr_wf_batch();
    if (b == nullptr) return hipErrorOutOfMemory;
    hipError_t e = hipEventCreateWithFlags(&b->staged, hipEventDisableTiming);
    if (e != hipSuccess)
    {
        delete b;
        return e;
    }
    b->cus = device_cus();
    b->grid = b->cus * b->grid_per_cu;
    *out = b;
    return hipSuccess;
}

hipError_t qr_wf_batch_config(qr_wf_batch *b, int grid_per_cu, int tile_w, int rpt,
                              int cadence)
{
    int tw = 0;
    while ((1 << tw) < tile_w) tw++;
    if (grid_per_cu < 1 || grid_per_cu > 64 || (1 << tw) != tile_w || tw < 2 || tw > 5 ||
        rpt < 1 || rpt > QR_WF_MAX_RPT || cadence < 1 || cadence > QR_WF_MAX_CADENCE)
    {
        return hipErrorInvalidValue;
    }
    b->grid_per_cu = grid_per_cu;
    b->grid = b->cus * grid_per_cu;
    b->tw_log2 = tw;
    b->rpt = rpt;
    b->cadence = cadence;
    return hipSuccess;
}

void qr_wf_batch_destroy(qr_wf_batch *b)
{
    if (b == nullptr) return;
    if (b->staged != nullptr) (void)hipEventSynchronize(b->staged);
    /* side work may still read this batch's buffers */
    if (b->join != nullptr) (void)hipEventSynchronize(b->join);
    b->frames.release();
    b->plans.release();
    b->cnt.release();
    b->shc.release();
    if (b->host != nullptr) (void)hipHostFree(b->host);
    if (b->cnt_host != nullptr) (void)hipHostFree(b->cnt_host);
    if (b->staged != nullptr) (void)hipEventDestroy(b->staged);
    if (b->fork != nullptr) (void)hipEventDestroy(b->fork);
    if (b->join != nullptr) (void)hipEventDestroy(b->join);
    delete b;
}

uint32_t qr_wf_batch_tiles(const qr_wf_batch *b, int x_res, int y_res)
{
    const int tw = b->tw_log2, th = 6 - tw;
    const uint32_t tx = ((uint32_t)x_res + (1u << tw) - 1) >> tw;
    const uint32_t ty = ((uint32_t)y_res + (1u << th) - 1) >> th;
    return tx * ty;
}

hipError_t qr_wf_create(qr_wf_state **out)
{
    *out = nullptr;
    qr_wf_state *s = new (std::nothrow) qr_wf_state();
    if (s == nullptr) return hipErrorOutOfMemory;
    memset(&s->last_fd, 0, sizeof(s->last_fd));
    *out = s;
    return hipSuccess;
}

void qr_wf_destroy(qr_wf_state *s)
{
    if (s == nullptr) return;
    qr_wf_batch_destroy(s->own);
    for (LevelBufs &l : s->lv)
    {
        l.ra.release(); l.rb.release(); l.rc.release();
        l.na.release(); l.nb.release(); l.nc.release();
        l.vis.release(); l.fa.release(); l.fb.release();
    }
    s->shq.release();
    delete s;
}

/* queue capacities of one frame: level 0 exact, deeper levels as last
 * measured (at least an eighth of level 0; a queue that overflows grows to
 * its count and the frame is traced again), shadow jobs per light as large
 * as level 0, in three sets (level 0's own, then odd and even deeper
 * levels) */
static hipError_t frame_reserve(qr_wf_state *s, const qr_device_tables *t, uint64_t n0,
                                int levels)
{
    hipError_t e;
    if ((e = level_reserve(s->lv[0], (uint32_t)n0, false)) != hipSuccess) return e;
    for (int L = 1; L < levels; L++)
    {
        const uint32_t want = std::max(s->want[L], round_cap(n0 / 8));
        if (s->lv[L].cap < want)
        {
            if ((e = level_reserve(s->lv[L], want, true)) != hipSuccess) return e;
        }
    }
    const uint32_t want = std::max(s->sh_want, (uint32_t)n0);
    if (s->sh_cap < want || s->n_lgt < t->n_lgt)
    {
        const uint32_t cap = std::max(want, s->sh_cap);
        const int nl = std::max(s->n_lgt, t->n_lgt);
        const size_t per = (size_t)std::max(1, nl);
        if ((e = s->shq.reserve(sizeof(uint32_t) * 3 * per * cap)) != hipSuccess) return e;
        s->sh_cap = cap;
        s->n_lgt = nl;
    }
    return hipSuccess;
}

/* one launch after the camera samples' nearest walks: its levels of nearest
 * walks, shadow walks and shading (-1: none) */
struct Stage
{
    int wl, sl, fl;
};

/* the stages of a batch whose deepest frame has "levels" levels.  Each
 * level's shadow walks follow its nearest walks and precede its shading
 * (level 0's shading is part of the resolve).  The camera samples' shadow
 * walks and the nearest walks of level 1 fill the chip by themselves and
 * run alone; from then on stage j runs the nearest walks of level j, the
 * shadow walks of level j-1 and the shading of level j-2 side by side */
static std::vector<Stage> stage_list(int levels)
{
    std::vector<Stage> st;
    auto lv = [&](int L) { return L >= 0 && L < levels ? L : -1; };
    st.push_back({-1, 0, -1});
    st.push_back({lv(1), -1, -1});
    for (int j = 2; j <= levels + 1; j++) st.push_back({lv(j), lv(j - 1), j >= 3 ? lv(j - 2) : -1});
    st.erase(std::remove_if(st.begin(), st.end(),
                            [](const Stage &x) { return x.wl < 0 && x.sl < 0 && x.fl < 0; }),
             st.end());
    return st;
}

hipError_t qr_wf_batch_launch(qr_wf_batch *b, const qr_wf_item *items, int n,
                              hipStream_t stream)
{
    if (n <= 0 || n > QR_WF_MAX_BATCH) return hipErrorInvalidValue;
    hipError_t e;
    int levels = 1;
    for (int k = 0; k < n; k++)
    {
        const qr_frame_desc *fd = items[k].fd;
        if (fd->depth < 0 || fd->depth > MAX_LEVELS - 2 ||
            items[k].t->n_lgt > MAX_LIGHTS || fd->x_res <= 0 || fd->y_res <= 0)
        {
            return hipErrorInvalidValue;
        }
        levels = std::max(levels, fd->depth + 1);
    }

    /* counters and plans */
    if (b->cnt_cap < n)
    {
        if ((e = b->cnt.reserve(sizeof(uint32_t) * CNT_N * (size_t)n)) != hipSuccess) return e;
        if ((e = b->shc.reserve(sizeof(uint32_t) * SHC_N * (size_t)n)) != hipSuccess) return e;
        if (b->cnt_host != nullptr) (void)hipHostFree(b->cnt_host);
        b->cnt_host = nullptr;
        b->cnt_cap = 0;
        e = hipHostMalloc((void **)&b->cnt_host, sizeof(uint32_t) * CNT_N * (size_t)n,
                          hipHostMallocDefault);
        if (e != hipSuccess) return e;
        memset(b->cnt_host, 0, sizeof(uint32_t) * CNT_N * (size_t)n);
        b->cnt_cap = n;
    }
    const size_t plan_words = (size_t)(n + 1);
    if ((e = b->plans.reserve(sizeof(uint32_t) * plan_words * PLAN_SLOTS)) !=
        hipSuccess)
    {
        return e;
    }

    /* frame arguments */
    std::vector<Frame> fr((size_t)n);
    for (int k = 0; k < n; k++)
    {
        const qr_wf_item &it = items[k];
        const qr_frame_desc *fd = it.fd;
        const int ns = fd->fsaa == 2 ? 4 : fd->fsaa == 1 ? 2 : 1;
        const int tw = b->tw_log2, th = 6 - tw;
        const uint32_t tiles_x = ((uint32_t)fd->x_res + (1u << tw) - 1) >> tw;
        const uint32_t tiles_y = ((uint32_t)fd->y_res + (1u << th) - 1) >> th;
        const uint64_t all = (uint64_t)tiles_x * tiles_y;
        uint64_t tile0 = it.tile0, ntiles = it.ntiles;
        if (ntiles == 0)
        {
            tile0 = 0;
            ntiles = all;
        }
        if (tile0 + ntiles > all) return hipErrorInvalidValue;
        const uint64_t n0 = ntiles * TILE_PX * ns;
        if (n0 > 0x7fffff00ull) return hipErrorInvalidValue;
        const int lv = fd->depth + 1;
        if ((e = frame_reserve(it.s, it.t, n0, lv)) != hipSuccess) return e;

        Frame &a = fr[k];
        memset(&a, 0, sizeof(a));
        a.sc = scene_of(it.t);
        a.fd = *fd;
        a.fd.x_row = fd->x_res;
        for (int L = 0; L < lv; L++) a.lv[L] = level_view(it.s->lv[L]);
        a.lv[0].cap = (uint32_t)n0;
        a.cnt = (uint32_t *)b->cnt.ptr + (size_t)CNT_N * k;
        a.shc = (uint32_t *)b->shc.ptr + (size_t)SHC_N * k;
        a.shq = (uint32_t *)it.s->shq.ptr;
        a.sh_cap = it.s->sh_cap;
        a.n0 = (uint32_t)n0;
        a.tile0 = (uint32_t)tile0;
        a.ntiles = (uint32_t)ntiles;
        a.tiles_x = tiles_x;
        a.ns = ns;
        a.levels = lv;
        a.tw_log2 = tw;
        for (int L = 1; L < lv; L++)
        {
            if (L % b->cadence != 0) a.sparse |= 1u << L;
        }
        a.ids = it.ids;
        a.frame = it.frame;
    }
    const size_t bytes = sizeof(Frame) * (size_t)n;
    if (b->last.size() != (size_t)n || memcmp(b->last.data(), fr.data(), bytes) != 0)
    {
        /* the pinned staging buffer may still feed the previous copy */
        if ((e = hipEventSynchronize(b->staged)) != hipSuccess) return e;
        if (b->host_cap < n)
        {
            if (b->host != nullptr) (void)hipHostFree(b->host);
            b->host = nullptr;
            b->host_cap = 0;
            if ((e = hipHostMalloc((void **)&b->host, bytes, hipHostMallocDefault)) !=
                hipSuccess)
            {
                return e;
            }
            b->host_cap = n;
        }
        if ((e = b->frames.reserve(bytes)) != hipSuccess) return e;
        memcpy(b->host, fr.data(), bytes);
        e = hipMemcpyAsync(b->frames.ptr, b->host, bytes, hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) return e;
        if ((e = hipEventRecord(b->staged, stream)) != hipSuccess) return e;
        b->last = fr;
    }
    b->nfr = n;
    b->levels = levels;

    /* the stages */
    uint32_t *plans = (uint32_t *)b->plans.ptr;
    auto plan_of = [&](int kind, int L) -> uint32_t *
    {
        return plans + plan_words * (size_t)plan_slot(kind, L);
    };
    Batch bt;
    bt.fr = (const QC Frame *)b->frames.ptr;
    bt.plan = nullptr;
    bt.nfr = n;
    bt.level = 0;
    bt.wl = bt.sl = bt.fl = -1;
    bt.rpt = b->rpt;
    /* the persistent grid, but no more workgroups than the camera samples
     * fill: a small batch (a chunk of a served step, a thumbnail) does not
     * pay for thousands of empty workgroups per launch */
    uint64_t wg0 = 0;
    for (const Frame &a : fr) wg0 += ((uint64_t)a.n0 + WG - 1) / WG;
    const unsigned gfit = (unsigned)std::min<uint64_t>(
        (uint64_t)b->grid, std::max<uint64_t>(wg0, (uint64_t)b->cus));
    const dim3 g(gfit), w(WG);
    /* kernels for one block of rays per thread and dense queues at every
     * level (the default shape), or for any shape */
    const bool one = b->rpt == 1 && b->cadence == 1;
    /* a batch of one frame needs no plans: its kernels count their own work */
    const bool planned = n > 1;
    auto plan = [&](int kind, int L) -> Batch
    {
        Batch p = bt;
        p.level = L;
        if (!planned) return p;
        hipLaunchKernelGGL(k_plan, dim3(1), dim3(1024), 0, stream, p, kind, plan_of(kind, L),
                           1, 0u);
        p.plan = (const QC uint32_t *)plan_of(kind, L);
        return p;
    };
    hipLaunchKernelGGL(k_reset, dim3((unsigned)std::min(n, 1024)), w, 0, stream, bt);
    hipLaunchKernelGGL(one ? k_walk0<1> : k_walk0<0>, g, w, 0, stream, plan(PLAN_LEVEL, 0));
    const std::vector<Stage> stages = stage_list(levels);
    /* Only the resolve reads the camera samples' shadow walks (stage 0),
     * whose queue set, visibility masks and counters no other stage
     * touches: outside graph capture they run on a second stream beside
     * all the deeper stages and folds, each kernel with its own register
     * budget, and the resolve waits for them. */
    hipStreamCaptureStatus cap = hipStreamCaptureStatusActive;
    const bool overlap = !stages.empty() && stages[0].sl == 0 && stages[0].wl < 0 &&
                         stages[0].fl < 0 &&
                         hipStreamIsCapturing(stream, &cap) == hipSuccess &&
                         cap == hipStreamCaptureStatusNone;
    for (size_t k = 0; k < stages.size(); k++)
    {
        const Stage &st = stages[k];
        Batch p = bt;
        p.wl = st.wl;
        p.sl = st.sl;
        p.fl = st.fl;
        if (planned)
        {
            hipLaunchKernelGGL(k_plan, dim3(1), dim3(1024), 0, stream, p, PLAN_STAGE,
                               plan_of(PLAN_STAGE, (int)k), 1, 0u);
            p.plan = (const QC uint32_t *)plan_of(PLAN_STAGE, (int)k);
        }
        if (overlap && k == 0)
        {
            /* the full grid: one that holds only part of each CU (to leave
             * more room to the deeper stages) measured slower, 2.16 ms per
             * frame at 4 workgroups per CU against 2.11 */
            if ((e = side_stream(b)) != hipSuccess) return e;
            if ((e = hipEventRecord(b->fork, stream)) != hipSuccess) return e;
            if ((e = hipStreamWaitEvent(b->side, b->fork, 0)) != hipSuccess) return e;
            hipLaunchKernelGGL(one ? k_shadow<1> : k_shadow<0>, g, w, 0, b->side, p);
            continue;
        }
        const int roles = (st.wl >= 0) + (st.sl >= 0) + (st.fl >= 0);
        if (roles > 1) hipLaunchKernelGGL(one ? k_stage<1> : k_stage<0>, g, w, 0, stream, p);
        else if (st.wl >= 0) hipLaunchKernelGGL(one ? k_walk<1> : k_walk<0>, g, w, 0, stream, p);
        else if (st.sl >= 0) hipLaunchKernelGGL(one ? k_shadow<1> : k_shadow<0>, g, w, 0, stream, p);
        else hipLaunchKernelGGL(one ? k_shading<1> : k_shading<0>, g, w, 0, stream, p);
    }
    /* levels 1 .. FOLD_INLINE are folded by the resolve, the levels from
     * FOLD_INLINE + 1 to levels - 2 by a launch each, deepest first */
    const int fold_lo = 1 + FOLD_INLINE;
    if (planned && levels - 2 >= fold_lo)
    {
        /* the fold launches' plans */
        Batch p = bt;
        p.level = fold_lo;
        hipLaunchKernelGGL(k_plan, dim3(1), dim3(1024), 0, stream, p, PLAN_LEVEL,
                           plan_of(PLAN_LEVEL, fold_lo), levels - 1 - fold_lo,
                           (uint32_t)plan_words);
    }
    for (int L = levels - 2; L >= fold_lo; L--)
    {
        Batch pl = bt;
        pl.level = L;
        if (planned) pl.plan = (const QC uint32_t *)plan_of(PLAN_LEVEL, L);
        hipLaunchKernelGGL(one ? k_fold<1> : k_fold<0>, g, w, 0, stream, pl);
    }
    if (overlap)
    {
        /* the resolve shades level 0 with its shadow walks' visibility */
        if ((e = hipEventRecord(b->join, b->side)) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(stream, b->join, 0)) != hipSuccess) return e;
    }
    Batch pr = bt;
    if (planned) pr.plan = (const QC uint32_t *)plan_of(PLAN_LEVEL, 0);
    hipLaunchKernelGGL(one ? k_resolve<1> : k_resolve<0>, g, w, 0, stream, pr);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return hipMemcpyAsync(b->cnt_host, b->cnt.ptr, sizeof(uint32_t) * CNT_N * (size_t)n,
                          hipMemcpyDeviceToHost, stream);
}

int qr_wf_batch_check(qr_wf_batch *b, const qr_wf_item *items, int n)
{
    int grow = 0;
    for (int k = 0; k < n && k < b->nfr; k++)
    {
        const uint32_t *c = b->cnt_host + (size_t)CNT_N * k;
        if (c[CNT_OVF] == 0) continue;
        /* grow every queue that overflowed to its counted length */
        qr_wf_state *s = items[k].s;
        const int lv = items[k].fd->depth + 1;
        for (int L = 1; L < lv; L++)
        {
            if (c[CNT_LVL + L] > s->lv[L].cap) s->want[L] = round_cap(c[CNT_LVL + L]);
        }
        if (c[CNT_SHMAX] > s->sh_cap) s->sh_want = round_cap(c[CNT_SHMAX]);
        grow = 1;
    }
    return grow;
}

void qr_wf_batch_stats(const qr_wf_batch *b, int k, qr_wf_stats *stats)
{
    memset(stats, 0, sizeof(*stats));
    if (k < 0 || k >= b->nfr) return;
    const uint32_t *c = b->cnt_host + (size_t)CNT_N * k;
    for (int L = 1; L < b->levels && L < MAX_LEVELS; L++) stats->level_rays[L] = c[CNT_LVL + L];
}

hipError_t qr_wf_launch(qr_wf_state *s, const qr_device_tables *t,
                        const qr_frame_desc *fd, uint32_t *frame, int32_t *ids,
                        hipStream_t stream)
{
    hipError_t e;
    if (s->own == nullptr && (e = qr_wf_batch_create(&s->own)) != hipSuccess) return e;
    const qr_wf_item it = {s, t, fd, frame, ids};
    s->last_fd = *fd;
    return qr_wf_batch_launch(s->own, &it, 1, stream);
}

int qr_wf_check(qr_wf_state *s, qr_wf_stats *stats)
{
    if (s->own == nullptr) return 0;
    const qr_wf_item it = {s, nullptr, &s->last_fd, nullptr, nullptr};
    if (stats != nullptr) qr_wf_batch_stats(s->own, 0, stats);
    return qr_wf_batch_check(s->own, &it, 1);
}
