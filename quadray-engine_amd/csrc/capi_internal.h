/*
 * capi_internal.h -- state behind the C ABI's opaque handles (qrhip_dev,
 * qrhip_scene), shared by the files that implement the ABI.  Not installed.
 */
#ifndef QR_CAPI_INTERNAL_H
#define QR_CAPI_INTERNAL_H

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "engine.h"
#include "qrhip.h"
#include "tracer_kernels.h"

namespace qrhip_impl {

/* record a failure for qrhip_last_error() and return its code */
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define HIP_TRY(call)                                                        \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) return qrhip_impl::hip_fail(e_, #call);        \
    } while (0)

/* device buffer that only grows */
struct DevBuf
{
    void  *ptr = nullptr;
    size_t cap = 0;

    hipError_t reserve(size_t bytes)
    {
        if (bytes <= cap) return hipSuccess;
        if (ptr != nullptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
        size_t want = bytes < 256 ? 256 : bytes;
        hipError_t e = hipMalloc(&ptr, want);
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release()
    {
        if (ptr != nullptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
};

/* device copies of one scene's tables */
/* reference point of a program's world boxes: the middle of their centres'
 * span (the origin when there are none) */
inline void cull_reference(const qr_cull_step *s, int n, float ref[4])
{
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int p = 0; p < n; p++)
    {
        if (s[p].elem >= 0) continue;
        for (int a = 0; a < 3; a++)
        {
            if (!std::isfinite(s[p].c[a])) continue;
            lo[a] = std::min(lo[a], s[p].c[a]);
            hi[a] = std::max(hi[a], s[p].c[a]);
        }
    }
    for (int a = 0; a < 3; a++) ref[a] = lo[a] <= hi[a] ? 0.5f * (lo[a] + hi[a]) : 0.0f;
    ref[3] = 0.0f;
}

/* the kernels' form of a program: bounds with the box's share of the
 * allowance, rounded outwards (qr_dev_step) */
inline void dev_steps(const qr_cull_step *s, int n, const float ref[4],
                      std::vector<qr_dev_step> &out)
{
    out.assign((size_t)n, qr_dev_step{});
    for (int p = 0; p < n; p++)
    {
        qr_dev_step &d = out[(size_t)p];
        d.elem = s[p].elem;
        d.skip = s[p].skip;
        const bool box = s[p].elem < 0 || s[p].skip == -2;
        if (!box) continue;
        const bool world = s[p].elem < 0;
        double l1 = 0.0;
        for (int a = 0; a < 3; a++)
        {
            l1 += std::fabs((double)s[p].c[a] - (world ? (double)ref[a] : 0.0));
        }
        const double m = (double)QR_CULL_FUZZ * l1;
        for (int a = 0; a < 3; a++)
        {
            const double c = s[p].c[a], h = s[p].h[a];
            if (!std::isfinite(c) || !std::isfinite(h) || !std::isfinite(m))
            {
                d.lo[a] = -INFINITY;
                d.hi[a] = INFINITY;
                continue;
            }
            float lo = (float)(c - h - m), hi = (float)(c + h + m);
            if ((double)lo > c - h - m) lo = std::nextafter(lo, -INFINITY);
            if ((double)hi < c + h + m) hi = std::nextafter(hi, INFINITY);
            d.lo[a] = lo;
            d.hi[a] = hi;
        }
    }
}

struct DevTables
{
    DevBuf srf, mat, lgt, clp, elem, tex, step, step_any;
    qr_device_tables t;
    std::vector<qr_dev_step> dstep, dstep_any;  /* the programs as uploaded */

    hipError_t upload(const qr_scene_tables &h, hipStream_t st)
    {
        hipError_t e;
        int n_step = h.n_step;
        float ref[4];
        cull_reference(h.steps, h.n_step, ref);
        if (n_step == 0)
        {
            /* tables without a culling program walk the whole list */
            dstep.assign((size_t)h.n_elem, qr_dev_step{});
            for (int i = 0; i < h.n_elem; i++)
            {
                dstep[(size_t)i].elem = i;
                dstep[(size_t)i].skip = -1;
            }
            n_step = h.n_elem;
        }
        else
        {
            dev_steps(h.steps, h.n_step, ref, dstep);
        }
        /* any-hit walks use the nearest-hit program when they have none */
        const bool any = h.n_step_any > 0 && h.n_step > 0;
        if (any) dev_steps(h.steps_any, h.n_step_any, ref, dstep_any);
        struct Item { DevBuf *b; const void *src; size_t n; };
        Item items[8] = {
            {&step, dstep.data(), sizeof(qr_dev_step) * (size_t)n_step},
            {&step_any, any ? dstep_any.data() : nullptr,
             any ? sizeof(qr_dev_step) * (size_t)h.n_step_any : 0},
            {&srf, h.surfaces, sizeof(qr_surface) * (size_t)h.n_surf},
            {&mat, h.materials, sizeof(qr_material) * (size_t)h.n_mat},
            {&lgt, h.lights, sizeof(qr_light) * (size_t)h.n_lgt},
            {&clp, h.clips, sizeof(qr_clip) * (size_t)h.n_clip},
            {&elem, h.elems, sizeof(qr_elem) * (size_t)h.n_elem},
            {&tex, h.texels, sizeof(uint32_t) * (size_t)h.n_texel},
        };
        for (Item &it : items)
        {
            if ((e = it.b->reserve(it.n)) != hipSuccess) return e;
            if (it.n > 0)
            {
                e = hipMemcpyAsync(it.b->ptr, it.src, it.n, hipMemcpyHostToDevice, st);
                if (e != hipSuccess) return e;
            }
        }
        t.surfaces = (qr_surface *)srf.ptr;
        t.materials = (qr_material *)mat.ptr;
        t.lights = (qr_light *)lgt.ptr;
        t.clips = (qr_clip *)clp.ptr;
        t.elems = (qr_elem *)elem.ptr;
        t.texels = (uint32_t *)tex.ptr;
        t.steps = (qr_dev_step *)step.ptr;
        t.n_step = n_step;
        t.steps_any = any ? (qr_dev_step *)step_any.ptr : t.steps;
        for (int a = 0; a < 4; a++) t.cull_ref[a] = ref[a];
        t.n_step_any = any ? h.n_step_any : n_step;
        t.n_surf = h.n_surf;
        t.n_mat = h.n_mat;
        t.n_lgt = h.n_lgt;
        t.n_clip = h.n_clip;
        t.n_elem = h.n_elem;
        t.n_texel = h.n_texel;
        return hipSuccess;
    }
    void release()
    {
        srf.release(); mat.release(); lgt.release();
        clp.release(); elem.release(); tex.release(); step.release();
        step_any.release();
    }
};

} /* namespace qrhip_impl */

struct qrhip_dev
{
    int          ordinal = 0;
    hipStream_t  stream = nullptr;
    int          fsaa = QRHIP_FSAA_NO;
    int          path = QRHIP_PATH_LEVELS;
    qr_wf_state *wf = nullptr;        /* queues for qrhip_render_tables */
    qrhip_impl::DevTables scratch_tables;    /* for qrhip_render_tables */
    qrhip_impl::DevBuf    scratch_frame, scratch_ids;
    qrhip_impl::DevBuf    scratch_pt_planes, scratch_pt_seeds;
};

struct qrhip_scene
{
    qrhip_dev   *dev = nullptr;
    qr::Engine  *eng = nullptr;
    uint32_t    *frame = nullptr;      /* host frame, row 0 */
    std::vector<uint32_t> own_frame;
    int          x_res = 0, y_res = 0, x_row = 0;
    int          depth = 10;
    int          pt_on = 0;
    float        pts_c = 0.0f;
    int          fsaa = 0;
    bool         updated = false;
    int64_t      time = 0;             /* time of the last update */
    qr_frame_desc fd;
    qrhip_impl::DevTables dt;
    qrhip_impl::DevBuf dframe;         /* device frame, pitch x_res */
    qr_wf_state *wf = nullptr;         /* level pipeline queues */
    bool         pending = false;      /* launched frames not yet checked */
    /* path tracer: running-mean planes and seed plane, resident in HBM */
    qrhip_impl::DevBuf pt_planes, pt_seeds;
    size_t       pt_slots = 0;
    bool         pt_reset = true;      /* seeds and planes to (re)start */
    int          pt_fsaa = -1;         /* AA mode the planes were made with */
    uint32_t    *staging = nullptr;    /* pinned host copy of dframe */
    hipEvent_t   ev0 = nullptr, ev1 = nullptr;
};

namespace qrhip_impl {

/* a device handle opened with QRHIP_HOST_ONLY has no HIP state */
inline bool has_gpu(const qrhip_dev *d)
{
    return d != nullptr && d->ordinal >= 0;
}

/* a frame's initial path-tracer seed plane, slot (y * x_res + x) * ns + k,
 * from QuadRay's seed chain over rows of |x_row| pixels */
void pt_seed_plane(std::vector<uint32_t> &out, int x_res, int y_res, int x_row, int ns);

/* enqueue one Whitted frame on the device's chosen path */
hipError_t launch_frame(const qrhip_dev *d, qr_wf_state *wf, const qr_device_tables *t,
                        const qr_frame_desc *fd, uint32_t *frame, int32_t *ids);

} /* namespace qrhip_impl */

#endif /* QR_CAPI_INTERNAL_H */
