/*
 * noalloc.hip -- qrhip_render_tables_noalloc: the frame of
 * qrhip_render_tables traced without allocating device memory.
 *
 * Every device buffer the call needs beyond the caller's host inputs and
 * outputs (the staged copies of the tables, the culling programs in the
 * kernels' form, the device frame and first-hit ids) is carved from one
 * file-scope device array, "scratch", of QRHIP_NOALLOC_BYTES.  The frame is
 * traced by the one-kernel path (tracer.hip), whose only other memory is
 * each work-item's own node stack, so no queue has to be sized or grown;
 * its frames equal those of the level pipeline bit for bit.
 *
 * The array's address is looked up once per device with
 * hipGetSymbolAddress when the device is opened, and only that address is
 * handed to copies and kernels.  Every call on a device shares its array:
 * calls that overlap on one device race, and each call writes every byte it
 * reads before reading it (tables and programs are copied or converted in,
 * the kernel writes every pixel and id), so nothing needs clearing between
 * calls.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "capi_internal.h"
#include "qrhip.h"
#include "tracer_kernels.h"

static __device__ __align__(256) uint8_t scratch[QRHIP_NOALLOC_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
uint8_t *g_scratch[MAX_DEVICES];   /* per device: the array's device address */

/* the kernels' form of a culling program (DevTables::upload, dev_steps):
 * bounds grown by the box's share of the allowance; with src null, the
 * program that walks all n elements of the list */
__global__ void k_dev_steps(const qr_cull_step *src, int n, const qr_elem *elems,
                            const qr_surface *srf, int groups, float r0, float r1, float r2,
                            qr_dev_step *dst)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > n) return;
    qr_dev_step d;
    if (p == n)
    {
        qr_dev_end_step(&d);
        dst[p] = d;
        return;
    }
    for (int a = 0; a < 3; a++)
    {
        d.lo[a] = 0.0f;
        d.hi[a] = 0.0f;
    }
    const float ref[4] = {r0, r1, r2, 0.0f};
    if (src == nullptr)
    {
        qr_dev_elem_step(&d, p, elems[p].surf, elems[p].last, 0);
    }
    else
    {
        const qr_cull_step s = src[p];
        if (s.elem < 0)
        {
            qr_dev_box_step(&d, s.skip);
            refit_bounds(s.c, s.h, 1, ref, d.lo, d.hi);
        }
        else
        {
            const int surf = elems[s.elem].surf;
            const int kind = s.skip == -2 ? qr_local_kind(srf, surf, groups) : QR_LOCAL_NONE;
            qr_dev_elem_step(&d, s.elem, surf, elems[s.elem].last, kind);
            if (kind != QR_LOCAL_NONE) qr_local_bounds(srf, surf, s.c, s.h, d.lo, d.hi);
        }
    }
    dst[p] = d;
}

inline size_t up256(size_t n)
{
    return (n + 255) & ~(size_t)255;
}

} /* namespace */

namespace qrhip_impl {

hipError_t noalloc_setup(int ordinal)
{
    if (ordinal < 0 || ordinal >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[ordinal] != nullptr) return hipSuccess;
    /* with the device current: each device holds its own copy of the array */
    hipError_t e = hipSetDevice(ordinal);
    if (e != hipSuccess) return e;
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (ptr == nullptr) return hipErrorInvalidSymbol;
    g_scratch[ordinal] = (uint8_t *)ptr;
    return hipSuccess;
}

} /* namespace qrhip_impl */

using namespace qrhip_impl;

int qrhip_render_tables_noalloc(qrhip_dev *dev, const qr_scene_tables *tab,
                                const qr_frame_desc *fd, uint32_t *frame, int32_t *ids)
{
    if (!has_gpu(dev) || tab == nullptr || fd == nullptr || frame == nullptr)
    {
        return fail(QRHIP_E_ARG, "null argument");
    }
    int vrc = validate_tables(tab, fd);
    if (vrc != QRHIP_OK) return vrc;
    if (fd->x_row < fd->x_res && fd->x_row > -fd->x_res)
    {
        return fail(QRHIP_E_ARG, "frame row pitch shorter than a row");
    }

    /* the layout of the call's buffers in the array */
    const bool any = tab->n_step_any > 0 && tab->n_step > 0;
    const int n_step = tab->n_step > 0 ? tab->n_step : tab->n_elem;
    const int ns = fd->fsaa == 2 ? 4 : fd->fsaa == 1 ? 2 : 1;
    const size_t px = (size_t)fd->x_res * (size_t)fd->y_res;
    struct Part { const void *src; size_t n; size_t off; };
    Part part[] = {
        {tab->surfaces, sizeof(qr_surface) * (size_t)tab->n_surf, 0},
        {tab->materials, sizeof(qr_material) * (size_t)tab->n_mat, 0},
        {tab->lights, sizeof(qr_light) * (size_t)tab->n_lgt, 0},
        {tab->clips, sizeof(qr_clip) * (size_t)tab->n_clip, 0},
        {tab->elems, sizeof(qr_elem) * (size_t)tab->n_elem, 0},
        {tab->texels, sizeof(uint32_t) * (size_t)tab->n_texel, 0},
        {tab->steps, sizeof(qr_cull_step) * (size_t)(tab->n_step > 0 ? tab->n_step : 0), 0},
        {any ? tab->steps_any : nullptr,
         any ? sizeof(qr_cull_step) * (size_t)tab->n_step_any : 0, 0},
        {nullptr, sizeof(qr_dev_step) * ((size_t)n_step + 1), 0},            /* program */
        {nullptr, any ? sizeof(qr_dev_step) * ((size_t)tab->n_step_any + 1) : 0, 0},
        {nullptr, sizeof(uint32_t) * px, 0},                                 /* frame */
        {nullptr, ids != nullptr ? sizeof(int32_t) * px * (size_t)ns : 0, 0},
    };
    const int NPART = (int)(sizeof(part) / sizeof(part[0]));
    size_t total = 0;
    for (int i = 0; i < NPART; i++)
    {
        if (part[i].n > 0 && i < 8 && part[i].src == nullptr)
        {
            return fail(QRHIP_E_ARG, "null table");
        }
        part[i].off = total;
        total += up256(part[i].n);
    }
    if (total > (size_t)QRHIP_NOALLOC_BYTES)
    {
        return fail(QRHIP_E_ARG, "tables and frame need more device memory than the "
                                 "no-alloc scratch holds");
    }

    HIP_TRY(noalloc_setup(dev->ordinal));
    HIP_TRY(hipSetDevice(dev->ordinal));
    uint8_t *base = g_scratch[dev->ordinal];
    const hipStream_t st = dev->stream;
    for (int i = 0; i < 8; i++)
    {
        if (part[i].n == 0) continue;
        HIP_TRY(hipMemcpyAsync(base + part[i].off, part[i].src, part[i].n,
                               hipMemcpyHostToDevice, st));
    }

    /* the programs in the kernels' form, converted on the device */
    float ref[4];
    cull_reference(tab->steps, tab->n_step > 0 ? tab->n_step : 0, ref);
    qr_dev_step *dstep = (qr_dev_step *)(base + part[8].off);
    qr_dev_step *dany = any ? (qr_dev_step *)(base + part[9].off) : dstep;
    {
        /* n_step steps and the end step */
        const qr_cull_step *src = tab->n_step > 0 ? (const qr_cull_step *)(base + part[6].off)
                                                  : nullptr;
        hipLaunchKernelGGL(k_dev_steps, dim3((unsigned)((n_step + 1 + 255) / 256)), dim3(256), 0,
                           st, src, n_step, (const qr_elem *)(base + part[4].off),
                           (const qr_surface *)(base + part[0].off),
                           groups_sound(tab->steps, tab->n_step, tab->elems, tab->surfaces) ? 1 : 0,
                           ref[0], ref[1], ref[2], dstep);
    }
    if (any)
    {
        hipLaunchKernelGGL(k_dev_steps, dim3((unsigned)((tab->n_step_any + 1 + 255) / 256)),
                           dim3(256), 0, st, (const qr_cull_step *)(base + part[7].off),
                           tab->n_step_any, (const qr_elem *)(base + part[4].off),
                           (const qr_surface *)(base + part[0].off),
                           groups_sound(tab->steps_any, tab->n_step_any, tab->elems,
                                        tab->surfaces) ? 1 : 0,
                           ref[0], ref[1], ref[2], dany);
    }
    HIP_TRY(hipGetLastError());

    qr_device_tables t;
    t.surfaces = (qr_surface *)(base + part[0].off);
    t.materials = (qr_material *)(base + part[1].off);
    t.lights = (qr_light *)(base + part[2].off);
    t.clips = (qr_clip *)(base + part[3].off);
    t.elems = (qr_elem *)(base + part[4].off);
    t.texels = (uint32_t *)(base + part[5].off);
    t.steps = dstep;
    t.n_step = n_step;
    t.steps_any = dany;
    t.n_step_any = any ? tab->n_step_any : n_step;
    for (int a = 0; a < 4; a++) t.cull_ref[a] = ref[a];
    t.n_surf = tab->n_surf;
    t.n_mat = tab->n_mat;
    t.n_lgt = tab->n_lgt;
    t.n_clip = tab->n_clip;
    t.n_elem = tab->n_elem;
    t.n_texel = tab->n_texel;

    qr_frame_desc d = *fd;
    d.x_row = fd->x_res;
    uint32_t *dfr = (uint32_t *)(base + part[10].off);
    int32_t *dids = ids != nullptr ? (int32_t *)(base + part[11].off) : nullptr;
    HIP_TRY(qr_launch_whitted(&t, &d, dfr, dids, st));

    /* rows to the caller's frame at its pitch (negative: rows run backwards) */
    const size_t row = sizeof(uint32_t) * (size_t)fd->x_res;
    if (fd->x_row > 0)
    {
        HIP_TRY(hipMemcpy2DAsync(frame, sizeof(uint32_t) * (size_t)fd->x_row, dfr, row, row,
                                 (size_t)fd->y_res, hipMemcpyDeviceToHost, st));
    }
    else
    {
        for (int y = 0; y < fd->y_res; y++)
        {
            HIP_TRY(hipMemcpyAsync(frame + (ptrdiff_t)y * fd->x_row, dfr + (size_t)y * fd->x_res,
                                   row, hipMemcpyDeviceToHost, st));
        }
    }
    if (ids != nullptr)
    {
        HIP_TRY(hipMemcpyAsync(ids, dids, part[11].n, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    return QRHIP_OK;
}
