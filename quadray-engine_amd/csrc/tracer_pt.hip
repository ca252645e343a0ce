/*
 * tracer_pt.hip -- QuadRay's path tracer (PT domain of render0) for MI355X.
 *
 * One work-item per sample slot (pixel x anti-aliasing sample), 8x8-pixel
 * tiles so a wave's paths start close together.  A pass draws one sample per
 * slot and folds it into the slot's running-mean colour planes, which stay
 * resident in HBM between passes with the slots' random seeds (QuadRay keeps
 * the same planes and seed plane per framebuffer, core/engine/engine.cpp
 * 2910-2935, 3719-3756).
 *
 * A path's random numbers come from its slot's 32-bit LCG in the order
 * QuadRay's buffered tracer draws them along the path: two for the tent
 * jitter of the camera ray, then per hit one for Russian roulette, two for
 * the cosine-distributed diffuse direction and one for the refraction /
 * reflection pick; a diffuse child continues the state after its two draws,
 * refraction and reflection children after all draws of the hit.  The slot's
 * seed itself steps once per pass.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qr_soa.h"
#include "tracer_dev.h"
#include "tracer_kernels.h"

namespace {

constexpr int PT_TILE = 8;
constexpr int PT_STACK = 32;            /* > 1 + 2 * depth for depth <= 14 */
constexpr int PT_ROULETTE_DEPTH = 5;    /* RT_STACK_DEPTH - 5 (tracer.cpp:3488) */
constexpr int PT_SPLIT_DEPTH = 8;       /* RT_STACK_DEPTH - 2 (tracer.cpp:4690) */

struct PtArgs
{
    Scene          sc;
    qr_frame_desc  fd;
    float         *pr, *pg, *pb;        /* colour planes, one value per slot */
    uint32_t      *seeds;
    uint32_t       i0, i1;              /* tile-ordered slots of the pass: the
                                           tiles tile0 .. tile0 + ntiles - 1 */
    uint32_t       tile0, ntiles;
    uint32_t       tiles_x;
    int            ns;
    float          pts_o, pts_u;        /* 1/pass count, 1 - that */
    uint32_t      *frame;
};

/* next number of the slot's LCG as a float in [0, 1): top 24 bits over 2^24
 * (GET_RANDOM_F tracer.cpp:1949-1980, LCG24 tracer.h:58) */
__device__ __forceinline__ float pt_random(uint32_t &st)
{
    st = st * 214013u + 2531011u;
    return (float)((st >> 8) & 0xFFFFFFu) / 16777216.0f;
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

struct PtJob
{
    Ray      r;
    V3       mul;       /* path throughput (MUL_R/G/B) */
    uint32_t st;
};

__device__ __forceinline__ Ray pt_child(const Ray &r, const Hit &h, const V3 &dir,
                                        int flags)
{
    Ray c;
    c.org = h.hit;
    c.dir = dir;
    c.t_min = 0.0f;
    c.t_max = 3.402823466e+38f;
    c.flags = flags;
    c.origin = h.srf;
    c.oloc = h.loc;
    c.depth = r.depth - 1;
    return c;
}

/* one sample: each hit adds its emission times the path throughput; diffuse
 * hits continue along a cosine-distributed direction (3479-3800), Russian
 * roulette by the colour's maximum at remaining depth <= 5; transparent
 * Fresnel surfaces pick refraction or reflection at remaining depth <= 8
 * with smallpt's weights (4686-4752) */
__device__ void pt_sample(const PtArgs &a, const Ray &primary, uint32_t st, V3 &acc)
{
    PtJob stk[PT_STACK];
    int sp = 0;
    stk[0].r = primary;
    stk[0].mul = V3{1.0f, 1.0f, 1.0f};
    stk[0].st = st;
    sp = 1;
    while (sp > 0)
    {
        PtJob j = stk[--sp];
        Hit h;
        const bool rev = j.r.origin < 0 || own_list(a.sc.srf[j.r.origin]);
        if (!walk<false>(a.sc, j.r, h, rev)) continue;
        hit_info(a.sc, j.r, h);
        const QC qr_surface &s = a.sc.srf[h.srf];
        const uint32_t props = s.props[h.side];
        const QC qr_material &m = a.sc.mat[s.mat[h.side]];
        const V3 nrm = normal_at(a.sc, s, h);
        V3 tex = texel_color(a.sc, m, props, h);

        float c;
        c = m.col[0] * a.pts_o; c = c * j.mul.x; acc.x = acc.x + c;
        c = m.col[1] * a.pts_o; c = c * j.mul.y; acc.y = acc.y + c;
        c = m.col[2] * a.pts_o; c = c * j.mul.z; acc.z = acc.z + c;

        if (props & QR_PROP_DIFFUSE)
        {
            bool alive = true;
            if (j.r.depth <= PT_ROULETTE_DEPTH)
            {
                const float p = fmaxf(fmaxf(tex.x, tex.y), tex.z);
                const float rr = pt_random(j.st);
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
                const V3 &ry = j.r.dir;
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
                const float r1 = pt_random(j.st);
                const float sn = sqrtf(r1);
                const float cs = sqrtf(1.0f - r1);
                const float r2 = pt_random(j.st);
                float phi = (r2 + r2) * 3.14159265358979323846f;
                phi = phi - 3.14159265358979323846f;
                const float cu = pt_cos(phi) * sn, sv = pt_sin(phi) * sn;
                V3 dir;
                dir.x = nrm.x * cs; dir.x = dir.x + u.x * cu; dir.x = dir.x + v.x * sv;
                dir.y = nrm.y * cs; dir.y = dir.y + u.y * cu; dir.y = dir.y + v.y * sv;
                dir.z = nrm.z * cs; dir.z = dir.z + u.z * cu; dir.z = dir.z + v.z * sv;
                if (j.r.depth != 0 && sp < PT_STACK)
                {
                    PtJob &cj = stk[sp++];
                    cj.r = pt_child(j.r, h, dir, h.side);
                    float w;
                    w = j.mul.x * m.l_dff; cj.mul.x = w * tex.x;
                    w = j.mul.y * m.l_dff; cj.mul.y = w * tex.y;
                    w = j.mul.z * m.l_dff; cj.mul.z = w * tex.z;
                    cj.st = j.st;
                }
            }
        }

        TrRf t = trrf_setup(m, props, j.r.dir, nrm);
        if (t.fresnel_tr && j.r.depth <= PT_SPLIT_DEPTH)
        {
            const float rr = pt_random(j.st);
            float q = t.c_rfl / (t.c_trn + t.c_rfl);
            q = q * 0.5f;
            const float pr = 0.25f + q;
            const bool pick_tr = rr >= pr;
            t.c_trn = pick_tr ? t.c_trn / (1.0f - pr) : 0.0f;
            t.c_rfl = pick_tr ? 0.0f : t.c_rfl / pr;
            t.tr = t.tr && pick_tr;
            t.rf = t.rf && !pick_tr;
        }
        if (t.tr && j.r.depth != 0 && sp < PT_STACK)
        {
            PtJob &cj = stk[sp++];
            cj.r = pt_child(j.r, h, t.tdir, h.side | FLAG_PASS);
            cj.mul = V3{j.mul.x * t.c_trn, j.mul.y * t.c_trn, j.mul.z * t.c_trn};
            cj.st = j.st;
        }
        if (t.rf && j.r.depth != 0 && sp < PT_STACK)
        {
            PtJob &cj = stk[sp++];
            cj.r = pt_child(j.r, h, t.rdir, h.side);
            cj.mul = V3{j.mul.x * t.c_rfl, j.mul.y * t.c_rfl, j.mul.z * t.c_rfl};
            cj.st = j.st;
        }
    }
}

__global__ void __launch_bounds__(256) k_pt(PtArgs a)
{
    const uint32_t i = a.i0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.i1) return;
    const uint32_t per_tile = PT_TILE * PT_TILE * a.ns;
    const uint32_t tile = i / per_tile;
    const uint32_t rem = i - tile * per_tile;
    const uint32_t p = rem / a.ns;
    const int k = (int)(rem - p * a.ns);
    const int x = (int)((tile % a.tiles_x) * PT_TILE + p % PT_TILE);
    const int y = (int)((tile / a.tiles_x) * PT_TILE + p / PT_TILE);
    if (x >= a.fd.x_res || y >= a.fd.y_res) return;

    const size_t slot = ((size_t)y * a.fd.x_res + x) * a.ns + k;
    uint32_t st = a.seeds[slot];
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

    V3 acc = V3{a.pr[slot] * a.pts_u, a.pg[slot] * a.pts_u, a.pb[slot] * a.pts_u};
    pt_sample(a, r, st, acc);
    a.pr[slot] = acc.x;
    a.pg[slot] = acc.y;
    a.pb[slot] = acc.z;
    a.seeds[slot] = a.seeds[slot] * 214013u + 2531011u;
}

/* the pass's pixels from the planes: clamp, anti-aliasing resolve, ARGB8;
 * four tiles of 8x8 pixels per workgroup */
__global__ void __launch_bounds__(256) k_pt_resolve(PtArgs a)
{
    const uint32_t lt = blockIdx.x * 4 + threadIdx.x / 64;
    if (lt >= a.ntiles) return;
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

} /* namespace */

uint32_t qr_pt_tiles(int x_res, int y_res)
{
    return (uint32_t)((x_res + PT_TILE - 1) / PT_TILE) * (uint32_t)((y_res + PT_TILE - 1) / PT_TILE);
}

hipError_t qr_launch_pt_tiles(const qr_device_tables *t, const qr_frame_desc *fd,
                              float *planes, uint32_t *seeds, uint32_t *frame,
                              uint32_t tile0, uint32_t ntiles, hipStream_t stream)
{
    if (fd->depth < 0 || fd->depth > 14 || !(fd->pts_c >= 1.0f))
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
    a.tiles_x = (uint32_t)(fd->x_res + PT_TILE - 1) / PT_TILE;
    const uint32_t per_tile = PT_TILE * PT_TILE * (uint32_t)a.ns;
    a.tile0 = tile0;
    a.ntiles = ntiles;
    a.i0 = tile0 * per_tile;
    a.i1 = (tile0 + ntiles) * per_tile;
    a.pts_o = 1.0f / fd->pts_c;
    a.pts_u = 1.0f - a.pts_o;
    a.frame = frame;
    hipLaunchKernelGGL(k_pt, dim3((ntiles * per_tile + 255) / 256), dim3(256), 0, stream, a);
    hipLaunchKernelGGL(k_pt_resolve, dim3((ntiles + 3) / 4), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t qr_launch_pt(const qr_device_tables *t, const qr_frame_desc *fd,
                        float *planes, uint32_t *seeds, uint32_t *frame,
                        hipStream_t stream)
{
    return qr_launch_pt_tiles(t, fd, planes, seeds, frame, 0, 0, stream);
}
