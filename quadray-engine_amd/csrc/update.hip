/*
 * update.hip -- the device side of the update stage: the culling programs'
 * boxes refitted from the element inputs of this update (qr_refit_elem),
 * with the programs' steps (their ranges and elements) unchanged.
 *
 * No counterpart in QuadRay, which rebuilds its tile lists every frame on
 * the CPU (core/engine/engine.cpp:991-1139, 1985-2160).  The host keeps the
 * same boxes in its tables (Engine::refit_cull_program); both sides use the
 * functions of refit.h, so they agree bit for bit.
 */
#include <hip/hip_runtime.h>

#include "refit.h"
#include "tracer_kernels.h"

namespace {

constexpr int RF_WG = 1024;

/* one workgroup: the elements' world boxes into wb (8 floats each: low,
 * high, has-box), then every box step of both programs as the union of the
 * world boxes of the elements it ranges over, and every element step with a
 * local box from its element's, each in the kernels' form (refit_bounds) */
__global__ void __launch_bounds__(RF_WG) k_refit(qr_dev_step *prog0, int n0, qr_dev_step *prog1,
                                                 int n1, const qr_refit_elem *el, int n_elem,
                                                 float rx, float ry, float rz, float *wb)
{
    for (int e = threadIdx.x; e < n_elem; e += RF_WG)
    {
        float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {0.0f, 0.0f, 0.0f};
        qr_refit_elem r = el[e];
        const int ok = refit_world_box(&r, lo, hi);
        float *o = wb + 8 * (size_t)e;
        for (int a = 0; a < 3; a++)
        {
            o[a] = lo[a];
            o[3 + a] = hi[a];
        }
        o[6] = ok ? 1.0f : 0.0f;
        o[7] = 0.0f;
    }
    __syncthreads();

    const float ref[3] = {rx, ry, rz};
    for (int k = 0; k < 2; k++)
    {
        qr_dev_step *ps = k == 0 ? prog0 : prog1;
        const int n = k == 0 ? n0 : n1;
        for (int p = threadIdx.x; p < n; p += RF_WG)
        {
            const int elem = ps[p].elem, skip = ps[p].skip;
            float c[3], h[3], lo[3], hi[3];
            if (elem >= 0)
            {
                if (skip != -2 || elem >= n_elem) continue;
                for (int a = 0; a < 3; a++)
                {
                    c[a] = el[elem].lc[a];
                    h[a] = el[elem].lh[a];
                }
                refit_bounds(c, h, 0, ref, lo, hi);
            }
            else
            {
                for (int a = 0; a < 3; a++)
                {
                    lo[a] = 3.402823466e+38f;
                    hi[a] = -3.402823466e+38f;
                }
                const int end = skip < n ? skip : n;
                for (int q = p + 1; q < end; q++)
                {
                    const int e = ps[q].elem;
                    if (e < 0 || e >= n_elem) continue;
                    const float *w = wb + 8 * (size_t)e;
                    if (w[6] == 0.0f) continue;
                    for (int a = 0; a < 3; a++)
                    {
                        lo[a] = w[a] < lo[a] ? w[a] : lo[a];
                        hi[a] = hi[a] < w[3 + a] ? w[3 + a] : hi[a];
                    }
                }
                refit_centre(lo, hi, c, h);
                float blo[3], bhi[3];
                refit_bounds(c, h, 1, ref, blo, bhi);
                for (int a = 0; a < 3; a++)
                {
                    lo[a] = blo[a];
                    hi[a] = bhi[a];
                }
            }
            for (int a = 0; a < 3; a++)
            {
                ps[p].lo[a] = lo[a];
                ps[p].hi[a] = hi[a];
            }
        }
    }
}

}  /* namespace */

hipError_t qr_refit_programs(qr_dev_step *steps, int n_step, qr_dev_step *steps_any,
                             int n_step_any, const qr_refit_elem *elems, int n_elem,
                             const float ref[4], float *scratch, hipStream_t stream)
{
    if (n_step < 0 || n_step_any < 0 || n_elem < 0 || (n_step_any > 0 && steps_any == nullptr))
    {
        return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(k_refit, dim3(1), dim3(RF_WG), 0, stream, steps, n_step, steps_any,
                       n_step_any, elems, n_elem, ref[0], ref[1], ref[2], scratch);
    return hipGetLastError();
}
