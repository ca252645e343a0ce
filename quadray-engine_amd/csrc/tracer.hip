/*
 * tracer.hip -- one-kernel Whitted tracer for MI355X (gfx950).
 *
 * One work-item per pixel, looping over its anti-aliasing sub-samples; the
 * whole reflection / refraction tree of a sample is traced inside the
 * kernel.  QuadRay's recursive context stack (shadow, refraction and
 * reflection rays re-entering the list walk, core/tracer/tracer.cpp
 * 4046-4077, 4770-4862, 5160-5277) becomes an explicit per-lane stack of
 * pending tree nodes; colours are accumulated bottom-up in QuadRay's order.
 * The level-by-level pipeline (tracer_wf.hip) is the default path; this
 * kernel serves scenes beyond that pipeline's light-mask limit.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qr_soa.h"
#include "tracer_dev.h"
#include "tracer_kernels.h"

namespace {

__device__ __forceinline__ Ray child(const Node &nd, int which)
{
    Ray c;
    c.org = nd.org;
    c.dir = which == 1 ? nd.tdir : nd.rdir;
    c.t_min = 0.0f;
    c.t_max = 3.402823466e+38f;
    c.flags = nd.side | (which == 1 ? FLAG_PASS : 0);
    c.origin = nd.srf;
    c.oloc = nd.loc;
    c.depth = nd.depth - 1;
    return c;
}

/* one camera sample through the whole Whitted tree: a node waits on the
 * stack for its refraction child, then its reflection child, and folds each
 * result into its colour as QuadRay's TR_mix / RF_mix do */
__device__ V3 trace_sample(const Scene &sc, const qr_frame_desc &fd, Ray ray,
                           int32_t *hit_id)
{
    Node stk[MAX_STACK];
    int sp = 0;
    V3 ret = {0.0f, 0.0f, 0.0f};
    bool primary = true;

    for (;;)
    {
        /* trace "ray" into node sp */
        Hit h;
        bool descend = false;
        if (!walk<false>(sc, ray, h))
        {
            ret = V3{0.0f, 0.0f, 0.0f};
            if (primary) *hit_id = -1;
        }
        else
        {
            if (primary) *hit_id = (sc.srf[h.srf].id << 1) | h.side;
            hit_info(sc, ray, h);
            const V3 nrm = normal_at(sc, sc.srf[h.srf], h);
            Node &nd = stk[sp];
            shade(sc, fd, ray, h, nrm, nd, [&](int, const Ray &sr) {
                Hit sh;
                return walk<true>(sc, sr, sh);
            });
            if (nd.todo & 1)
            {
                nd.todo |= 4;
                ray = child(nd, 1);
                descend = true;
            }
            else if (nd.todo & 2)
            {
                nd.todo |= 8;
                ray = child(nd, 2);
                descend = true;
            }
            else
            {
                ret = nd.col;
            }
        }
        primary = false;
        if (descend)
        {
            sp++;
            continue;
        }

        /* fold completed results into their parents */
        for (;;)
        {
            if (sp == 0) return ret;
            sp--;
            Node &nd = stk[sp];
            if (nd.todo & 4)
            {
                nd.col.x = ret.x * nd.c_trn + nd.col.x;
                nd.col.y = ret.y * nd.c_trn + nd.col.y;
                nd.col.z = ret.z * nd.c_trn + nd.col.z;
                nd.todo &= ~4;
                if (nd.todo & 2)
                {
                    nd.todo |= 8;
                    ray = child(nd, 2);
                    descend = true;
                    break;
                }
                ret = nd.col;
                continue;
            }
            nd.col.x = ret.x * nd.c_rfl + nd.col.x;
            nd.col.y = ret.y * nd.c_rfl + nd.col.y;
            nd.col.z = ret.z * nd.c_rfl + nd.col.z;
            ret = nd.col;
        }
        sp++;
    }
}

__global__ void __launch_bounds__(256)
whitted_kernel(Scene sc, qr_frame_desc fd, uint32_t *__restrict__ frame,
               int32_t *__restrict__ ids)
{
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= fd.x_res || y >= fd.y_res) return;

    const int ns = fd.fsaa == 2 ? 4 : fd.fsaa == 1 ? 2 : 1;
    V3 acc[4];
    for (int k = 0; k < ns; k++)
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
        int32_t id;
        V3 c = trace_sample(sc, fd, r, &id);
        acc[k].x = c.x < 1.0f ? c.x : 1.0f;
        acc[k].y = c.y < 1.0f ? c.y : 1.0f;
        acc[k].z = c.z < 1.0f ? c.z : 1.0f;
        if (ids != nullptr)
        {
            ids[((size_t)y * fd.x_res + x) * ns + k] = id;
        }
    }

    V3 c;
    if (ns == 1)
    {
        c = acc[0];
    }
    else if (ns == 2)
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

    float v[3] = {c.x, c.y, c.z};
    uint32_t px = 0;
    #pragma unroll
    for (int a = 0; a < 3; a++)
    {
        float f = v[a];
        if (fd.gamma & QR_PROP_GAMMA) f = sqrtf(f);
        f = f * 255.0f;
        px |= ((uint32_t)cvt_rne(f) & 0xFFu) << (16 - 8 * a);
    }
    frame[(ptrdiff_t)y * fd.x_row + x] = px;
}

} /* namespace */

hipError_t qr_launch_whitted(const qr_device_tables *t, const qr_frame_desc *fd,
                             uint32_t *frame, int32_t *ids, hipStream_t stream)
{
    const Scene sc = scene_of(t);
    if (fd->depth < 0 || fd->depth > MAX_STACK - 2)
    {
        return hipErrorInvalidValue;
    }
    dim3 grid((fd->x_res + 15) / 16, (fd->y_res + 15) / 16);
    hipLaunchKernelGGL(whitted_kernel, grid, dim3(256), 0, stream, sc, *fd,
                       frame, ids);
    return hipGetLastError();
}
