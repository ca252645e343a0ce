/*
 * update.hip -- the device side of the update stage.  (1) Every record of
 * the surface table from its inputs (qr_geom: the host's matrices and axis
 * maps, static shape and clip data): coefficients, clip box and bounds,
 * with custom clippers' tightening, the record itself, a plane's texture
 * transforms, and its element's inputs of the culling programs, one
 * work-item per record (fields.h, the functions the host runs).  (2) The
 * culling programs' boxes refitted from those element inputs, with the
 * programs' steps (their ranges and elements) unchanged.
 *
 * No counterpart in QuadRay, which rebuilds its tile lists every frame on
 * the CPU (core/engine/engine.cpp:991-1139, 1985-2160).  The host keeps the
 * same boxes in its tables (Engine::refit_cull_program); both sides use the
 * functions of refit.h, so they agree bit for bit.
 */
#include <hip/hip_runtime.h>

#include "fields.h"
#include "refit.h"
#include "tracer_kernels.h"

namespace {

constexpr int RF_WG = 256;

/* (1) each element's world box into eb (8 floats: low, high, has-box) */
__global__ void __launch_bounds__(RF_WG) k_world(const qr_refit_elem *el, int n_elem, float *eb)
{
    const int e = blockIdx.x * RF_WG + threadIdx.x;
    if (e >= n_elem) return;
    float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {0.0f, 0.0f, 0.0f};
    qr_refit_elem r = el[e];
    const int ok = refit_world_box(&r, lo, hi);
    float *o = eb + 8 * (size_t)e;
    for (int a = 0; a < 3; a++)
    {
        o[a] = lo[a];
        o[3 + a] = hi[a];
    }
    o[6] = ok ? 1.0f : 0.0f;
    o[7] = 0.0f;
}

/* (2) one wave per step of both programs (the second's ids offset by n0):
 * a box step's box is the union of the world boxes of every element in its
 * range -- the union of the nested boxes' unions, since min and max are
 * exact -- in the kernels' form; lanes take every 64th step of the range
 * and the wave reduces.  (3) An element step with a local box takes its
 * element's. */
__global__ void __launch_bounds__(RF_WG) k_boxes(qr_dev_step *prog0, int n0, qr_dev_step *prog1,
                                                 int n1, const qr_refit_elem *el, int n_elem,
                                                 const qr_surface *srf, float rx, float ry,
                                                 float rz, const float *eb)
{
    const int g = (int)((blockIdx.x * RF_WG + threadIdx.x) >> 6);
    const int lane = (int)(threadIdx.x & 63);
    if (g >= n0 + n1) return;
    const bool second = g >= n0;
    qr_dev_step *ps = second ? prog1 : prog0;
    const int n = second ? n1 : n0;
    const int p = second ? g - n0 : g;
    const float ref[3] = {rx, ry, rz};
    const int elem = ps[p].elem;
    if (elem >= 0)
    {
        if (qr_dev_step_local(&ps[p]) == QR_LOCAL_NONE || elem >= n_elem || lane != 0) return;
        float c[3], h[3], lo[3], hi[3];
        for (int a = 0; a < 3; a++)
        {
            c[a] = el[elem].lc[a];
            h[a] = el[elem].lh[a];
        }
        qr_local_bounds(srf, qr_dev_step_surf(&ps[p]), c, h, lo, hi);
        for (int a = 0; a < 3; a++)
        {
            ps[p].lo[a] = lo[a];
            ps[p].hi[a] = hi[a];
        }
        return;
    }
    float lo[3], hi[3];
    for (int a = 0; a < 3; a++)
    {
        lo[a] = 3.402823466e+38f;
        hi[a] = -3.402823466e+38f;
    }
    const int skip = qr_dev_box_end(&ps[p]);
    const int end = skip < n ? skip : n;
    for (int q = p + 1 + lane; q < end; q += 64)
    {
        const int e = ps[q].elem;
        if (e < 0 || e >= n_elem) continue;
        const float *w = eb + 8 * (size_t)e;
        if (w[6] == 0.0f) continue;
        for (int a = 0; a < 3; a++)
        {
            lo[a] = w[a] < lo[a] ? w[a] : lo[a];
            hi[a] = hi[a] < w[3 + a] ? w[3 + a] : hi[a];
        }
    }
    for (int d = 32; d >= 1; d >>= 1)
    {
        for (int a = 0; a < 3; a++)
        {
            const float ol = __shfl_xor(lo[a], d), oh = __shfl_xor(hi[a], d);
            lo[a] = ol < lo[a] ? ol : lo[a];
            hi[a] = hi[a] < oh ? oh : hi[a];
        }
    }
    if (lane != 0) return;
    float c[3], h[3], blo[3], bhi[3];
    refit_centre(lo, hi, c, h);
    refit_bounds(c, h, 1, ref, blo, bhi);
    for (int a = 0; a < 3; a++)
    {
        ps[p].lo[a] = blo[a];
        ps[p].hi[a] = bhi[a];
    }
}

/* every record of the table: one work-item each.  expect holds, per
 * element, the world and local box flags the programs were built for; an
 * element whose flags differ sets *mismatch (the programs' structure no
 * longer fits: the host rebuilds them) */
__global__ void __launch_bounds__(64) k_records(const qr_geom *geoms, int n_rec,
                                                 const qr_clip *clips, qr_surface *srf,
                                                 qr_material *mat, qr_refit_elem *refit_in,
                                                 const int32_t *expect, uint32_t *mismatch)
{
    const int r = blockIdx.x * 64 + threadIdx.x;
    if (r >= n_rec) return;
    const qr_geom &g = geoms[r];
    if (g.tag == RT_TAG_ARRAY)
    {
        qr_surface out;
        qf_array_record(g, out);
        srf[r] = out;
        if (g.elem >= 0)
        {
            qr_refit_elem z;
            memset(&z, 0, sizeof(z));
            refit_in[g.elem] = z;
            if (expect[2 * g.elem] != 0 || expect[2 * g.elem + 1] != 0) atomicOr(mismatch, 1u);
        }
        return;
    }
    float sci[4], scj[4];
    float bmin[4], bmax[4], cmin[4], cmax[4];
    qf_shape(g, sci, scj);
    qf_bounds(geoms, clips, r, bmin, bmax, cmin, cmax);
    qr_surface out;
    qf_surface_record(g, sci, scj, bmin, bmax, cmin, cmax, out);
    srf[r] = out;
    if (g.tag == RT_TAG_PLANE)
    {
        for (int side = 0; side < 2; side++)
        {
            float tx[4];
            qf_plane_tex(g, side, tx);
            qr_material &m = mat[g.mat[side]];
            m.xscal = tx[0];
            m.yscal = tx[1];
            m.xoffs = tx[2];
            m.yoffs = tx[3];
        }
    }
    if (g.elem >= 0)
    {
        qr_refit_elem e;
        qf_refit_elem(g, bmin, bmax, cmin, cmax, e);
        refit_in[g.elem] = e;
        if (expect[2 * g.elem] != e.world || expect[2 * g.elem + 1] != e.local)
        {
            atomicOr(mismatch, 1u);
        }
    }
}

}  /* namespace */

hipError_t qr_update_records(const qr_geom *geoms, int n_rec, const qr_clip *clips,
                             qr_surface *srf, qr_material *mat, qr_refit_elem *refit_in,
                             const int32_t *expect, uint32_t *mismatch, hipStream_t stream)
{
    if (n_rec <= 0) return hipSuccess;
    /* one wave per workgroup: a record's work is a long serial chain, so
     * the records spread over as many compute units as they can */
    hipLaunchKernelGGL(k_records, dim3((unsigned)((n_rec + 63) / 64)), dim3(64), 0, stream,
                       geoms, n_rec, clips, srf, mat, refit_in, expect, mismatch);
    return hipGetLastError();
}

hipError_t qr_refit_programs(qr_dev_step *steps, int n_step, qr_dev_step *steps_any,
                             int n_step_any, const qr_refit_elem *elems, int n_elem,
                             const qr_surface *srf, const float ref[4], float *scratch,
                             hipStream_t stream)
{
    if (n_step < 0 || n_step_any < 0 || n_elem < 0 || (n_step_any > 0 && steps_any == nullptr))
    {
        return hipErrorInvalidValue;
    }
    if (n_elem > 0)
    {
        hipLaunchKernelGGL(k_world, dim3((unsigned)((n_elem + RF_WG - 1) / RF_WG)), dim3(RF_WG),
                           0, stream, elems, n_elem, scratch);
    }
    const int n = n_step + n_step_any;
    if (n > 0)
    {
        const unsigned waves_per_wg = RF_WG / 64;
        hipLaunchKernelGGL(k_boxes, dim3((unsigned)((n + waves_per_wg - 1) / waves_per_wg)),
                           dim3(RF_WG), 0, stream, steps, n_step, steps_any, n_step_any, elems,
                           n_elem, srf, ref[0], ref[1], ref[2], (const float *)scratch);
    }
    return hipGetLastError();
}
