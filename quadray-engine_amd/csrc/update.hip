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

/* one workgroup.  (1) The elements' world boxes into eb (8 floats each:
 * low, high, has-box).  (2) The box steps level by level, deepest first
 * (order: step ids of both programs, the second's offset by n0, grouped by
 * nesting depth; bounds[d] .. bounds[d + 1] - 1 hold depth d): a box's raw
 * union (ub, 6 floats per step) is the union of the world boxes of the
 * elements directly in its range and the unions of the boxes directly in
 * it -- the union over every element of its range, since min and max are
 * exact -- and then its bounds in the kernels' form.  (3) The element steps
 * with a local box take their element's. */
__global__ void __launch_bounds__(RF_WG) k_refit(qr_dev_step *prog0, int n0, qr_dev_step *prog1,
                                                 int n1, const qr_refit_elem *el, int n_elem,
                                                 float rx, float ry, float rz, float *eb,
                                                 float *ub, const int32_t *order,
                                                 const int32_t *bounds, int depths)
{
    for (int e = threadIdx.x; e < n_elem; e += RF_WG)
    {
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
    __syncthreads();

    const float ref[3] = {rx, ry, rz};
    for (int d = depths - 1; d >= 0; d--)
    {
        for (int k = bounds[d] + threadIdx.x; k < bounds[d + 1]; k += RF_WG)
        {
            const int g = order[k];
            const bool second = g >= n0;
            qr_dev_step *ps = second ? prog1 : prog0;
            const int n = second ? n1 : n0;
            const int p = second ? g - n0 : g;
            const int off = second ? n0 : 0;
            float lo[3], hi[3];
            for (int a = 0; a < 3; a++)
            {
                lo[a] = 3.402823466e+38f;
                hi[a] = -3.402823466e+38f;
            }
            const int end = ps[p].skip < n ? ps[p].skip : n;
            for (int q = p + 1; q < end;)
            {
                const int e = ps[q].elem;
                const float *w = nullptr;
                int next = q + 1;
                if (e < 0)
                {
                    w = ub + 6 * (size_t)(off + q);
                    next = ps[q].skip > q ? ps[q].skip : q + 1;
                }
                else if (e < n_elem && eb[8 * (size_t)e + 6] != 0.0f)
                {
                    w = eb + 8 * (size_t)e;
                }
                if (w != nullptr)
                {
                    for (int a = 0; a < 3; a++)
                    {
                        lo[a] = w[a] < lo[a] ? w[a] : lo[a];
                        hi[a] = hi[a] < w[3 + a] ? w[3 + a] : hi[a];
                    }
                }
                q = next;
            }
            float *u = ub + 6 * (size_t)g;
            for (int a = 0; a < 3; a++)
            {
                u[a] = lo[a];
                u[3 + a] = hi[a];
            }
            float c[3], h[3], blo[3], bhi[3];
            refit_centre(lo, hi, c, h);
            refit_bounds(c, h, 1, ref, blo, bhi);
            for (int a = 0; a < 3; a++)
            {
                ps[p].lo[a] = blo[a];
                ps[p].hi[a] = bhi[a];
            }
        }
        __syncthreads();
    }

    for (int g = threadIdx.x; g < n0 + n1; g += RF_WG)
    {
        qr_dev_step &st = g < n0 ? prog0[g] : prog1[g - n0];
        const int elem = st.elem;
        if (elem < 0 || st.skip != -2 || elem >= n_elem) continue;
        float c[3], h[3], lo[3], hi[3];
        for (int a = 0; a < 3; a++)
        {
            c[a] = el[elem].lc[a];
            h[a] = el[elem].lh[a];
        }
        refit_bounds(c, h, 0, ref, lo, hi);
        for (int a = 0; a < 3; a++)
        {
            st.lo[a] = lo[a];
            st.hi[a] = hi[a];
        }
    }
}

}  /* namespace */

hipError_t qr_refit_programs(qr_dev_step *steps, int n_step, qr_dev_step *steps_any,
                             int n_step_any, const qr_refit_elem *elems, int n_elem,
                             const float ref[4], float *scratch, const int32_t *order,
                             const int32_t *bounds, int depths, hipStream_t stream)
{
    if (n_step < 0 || n_step_any < 0 || n_elem < 0 || depths < 0 ||
        (n_step_any > 0 && steps_any == nullptr))
    {
        return hipErrorInvalidValue;
    }
    float *eb = scratch;
    float *ub = scratch + 8 * (size_t)n_elem;
    hipLaunchKernelGGL(k_refit, dim3(1), dim3(RF_WG), 0, stream, steps, n_step, steps_any,
                       n_step_any, elems, n_elem, ref[0], ref[1], ref[2], eb, ub, order,
                       bounds, depths);
    return hipGetLastError();
}
