/*
 * capi.cpp -- the C ABI of libqrhip (include/qrhip.h): devices, scenes,
 * table upload and kernel launches.  Errors never cross the ABI as C++
 * exceptions; they become status codes plus a per-thread message.
 */
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "capi_internal.h"
#include "image.h"
#include "scenegen.hpp"

namespace qrhip_impl {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return fail(QRHIP_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

/* initial seed plane: the low 32 bits of a 48-bit LCG chain started at 1,
 * one value per framebuffer sample slot in QuadRay's order, slot
 * (y * |x_row| + x) * samples + k (rt_Scene::reset_pseed,
 * core/engine/engine.cpp:3719-3740); here gathered to (y * x_res + x) *
 * samples + k */
void pt_seed_plane(std::vector<uint32_t> &out, int x_res, int y_res, int x_row, int ns)
{
    const int ar = x_row < 0 ? -x_row : x_row;
    out.assign((size_t)x_res * y_res * ns, 0u);
    uint64_t st = 1;
    for (int y = 0; y < y_res; y++)
    {
        for (int x = 0; x < ar; x++)
        {
            for (int j = 0; j < ns; j++)
            {
                st = (st * 25214903917ull + 11ull) & 0x0000FFFFFFFFFFFFull;
                if (x < x_res) out[((size_t)y * x_res + x) * ns + j] = (uint32_t)st;
            }
        }
    }
}

/* enqueue one Whitted frame on the device's chosen path */
hipError_t launch_frame(const qrhip_dev *d, qr_wf_state *wf, const qr_device_tables *t,
                        const qr_frame_desc *fd, uint32_t *frame, int32_t *ids)
{
    if (d->path == QRHIP_PATH_LEVELS && t->n_lgt <= qr_wf_max_lights())
    {
        return qr_wf_launch(wf, t, fd, frame, ids, d->stream);
    }
    return qr_launch_whitted(t, fd, frame, ids, d->stream);
}

/* wait for the device's stream; if a level-pipeline queue overflowed in a
 * frame since the last check, the queues have grown and "relaunch" traces the
 * frame again (capacity never changes a result) */
template <class Relaunch>
static hipError_t finish_frames(const qrhip_dev *d, qr_wf_state *wf, bool pending,
                                Relaunch relaunch)
{
    hipError_t e = hipStreamSynchronize(d->stream);
    if (e != hipSuccess || !pending || wf == nullptr) return e;
    for (int attempt = 0; qr_wf_check(wf, nullptr) != 0; attempt++)
    {
        if (attempt == 16) return hipErrorOutOfMemory;
        if ((e = relaunch()) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(d->stream)) != hipSuccess) return e;
    }
    return hipSuccess;
}

/* every table index the kernels follow stays inside its table, and the
 * frame description is in range; checked on the host before a launch */
int validate_tables(const qr_scene_tables *tab, const qr_frame_desc *fd)
{
    if (fd->x_res <= 0 || fd->y_res <= 0 || fd->fsaa < 0 || fd->fsaa > 2 ||
        fd->depth < 0 || fd->depth > 14)
    {
        return fail(QRHIP_E_ARG, "frame description out of range");
    }
    /* every table index the kernels follow must stay inside the tables */
    if (tab->n_elem < 0 || tab->n_elem > QR_DEV_STEP_MAX)
        return fail(QRHIP_E_ARG, "surface list out of range");
    if (tab->n_surf < 0 || tab->n_surf > QR_DEV_SURF_MAX)
        return fail(QRHIP_E_ARG, "surface table out of range");
    for (int i = 0; i < tab->n_elem; i++)
    {
        if (tab->elems[i].surf < 0 || tab->elems[i].surf >= tab->n_surf)
            return fail(QRHIP_E_ARG, "list element outside the surface table");
    }
    for (int i = 0; i < tab->n_surf; i++)
    {
        const qr_surface &s = tab->surfaces[i];
        if (s.solver != QR_SOLVE_NONE &&
            (s.mat[0] < 0 || s.mat[0] >= tab->n_mat ||
             s.mat[1] < 0 || s.mat[1] >= tab->n_mat))
            return fail(QRHIP_E_ARG, "material index out of range");
        if (s.trm != 0 && (s.trnode < 0 || s.trnode >= tab->n_surf))
            return fail(QRHIP_E_ARG, "transform node index out of range");
        if (s.shift != 0 && s.trm == 0 && s.solver != QR_SOLVE_NONE)
            return fail(QRHIP_E_ARG, "surface in a transform node's frame without transform");
        if (s.clip_beg < 0 || s.clip_cnt < 0 || s.clip_beg + s.clip_cnt > tab->n_clip)
            return fail(QRHIP_E_ARG, "clip list out of range");
    }
    for (int i = 0; i < tab->n_clip; i++)
    {
        const qr_clip &c = tab->clips[i];
        if (c.kind < QR_CLIP_SURFACE || c.kind > QR_CLIP_LEAVE)
            return fail(QRHIP_E_ARG, "unknown clip list entry");
        if ((c.kind == QR_CLIP_SURFACE || c.kind == QR_CLIP_TRNODE) &&
            (c.surf < 0 || c.surf >= tab->n_surf))
            return fail(QRHIP_E_ARG, "clipper index out of range");
    }
    if (tab->n_step < 0 || tab->n_step > QR_DEV_STEP_MAX ||
        (tab->n_step > 0 && tab->steps == nullptr))
        return fail(QRHIP_E_ARG, "culling program out of range");
    for (int i = 0, prev = -1; i < tab->n_step; i++)
    {
        /* elements strictly in list order; box ranges inside the program */
        const qr_cull_step &c = tab->steps[i];
        if (c.elem >= 0)
        {
            if (c.elem >= tab->n_elem || c.elem <= prev)
                return fail(QRHIP_E_ARG, "culling program breaks list order");
            if (c.skip != -1 && c.skip != -2)
                return fail(QRHIP_E_ARG, "culling program out of range");
            if (c.skip == -2)
            {
                const qr_surface &s = tab->surfaces[tab->elems[c.elem].surf];
                if (s.shift == 0 || s.trm == 0 || s.tag < 0)
                    return fail(QRHIP_E_ARG, "local box on an untransformed surface or a transform node");
            }
            prev = c.elem;
        }
        else if (c.elem != -1 || c.skip <= i || c.skip > tab->n_step)
        {
            return fail(QRHIP_E_ARG, "culling program out of range");
        }
    }
    if (tab->n_step_any < 0 || tab->n_step_any > QR_DEV_STEP_MAX ||
        (tab->n_step_any > 0 && tab->steps_any == nullptr))
        return fail(QRHIP_E_ARG, "any-hit culling program out of range");
    for (int i = 0; i < tab->n_step_any; i++)
    {
        /* elements in range; box ranges inside the program */
        const qr_cull_step &c = tab->steps_any[i];
        if (c.elem >= 0)
        {
            if (c.elem >= tab->n_elem || (c.skip != -1 && c.skip != -2))
                return fail(QRHIP_E_ARG, "any-hit culling program out of range");
            if (c.skip == -2)
            {
                const qr_surface &s = tab->surfaces[tab->elems[c.elem].surf];
                if (s.shift == 0 || s.trm == 0 || s.tag < 0)
                    return fail(QRHIP_E_ARG, "local box on an untransformed surface or a transform node");
            }
        }
        else if (c.elem != -1 || c.skip <= i || c.skip > tab->n_step_any)
        {
            return fail(QRHIP_E_ARG, "any-hit culling program out of range");
        }
    }
    for (int i = 0; i < tab->n_mat; i++)
    {
        const qr_material &m = tab->materials[i];
        if (m.yshft > 24)
            return fail(QRHIP_E_ARG, "texture row shift out of range");
        uint64_t need = (uint64_t)m.tex_off + ((uint64_t)m.ymask << m.yshft) +
                        (uint64_t)m.xmask + 1;
        if (m.tex_off < 0 || need > (uint64_t)tab->n_texel)
            return fail(QRHIP_E_ARG, "texture outside the texel pool");
    }

    return QRHIP_OK;
}

} /* namespace qrhip_impl */

using namespace qrhip_impl;

extern "C" {

const char *qrhip_last_error(void)
{
    return g_err.c_str();
}

int qrhip_device_open(int ordinal, qrhip_dev **out)
{
    if (out == nullptr) return fail(QRHIP_E_ARG, "null output pointer");
    *out = nullptr;
    if (ordinal == QRHIP_HOST_ONLY)
    {
        qrhip_dev *h = new (std::nothrow) qrhip_dev();
        if (h == nullptr) return fail(QRHIP_E_DEVICE, "out of host memory");
        h->ordinal = QRHIP_HOST_ONLY;
        *out = h;
        return QRHIP_OK;
    }
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
    {
        return fail(QRHIP_E_DEVICE, "no HIP device available");
    }
    if (ordinal < 0 || ordinal >= n)
    {
        return fail(QRHIP_E_ARG, "device ordinal out of range");
    }
    HIP_TRY(hipSetDevice(ordinal));
    qrhip_dev *d = new (std::nothrow) qrhip_dev();
    if (d == nullptr) return fail(QRHIP_E_DEVICE, "out of host memory");
    d->ordinal = ordinal;
    e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
    if (e != hipSuccess)
    {
        delete d;
        return hip_fail(e, "hipStreamCreateWithFlags");
    }
    e = qr_wf_create(&d->wf);
    if (e != hipSuccess)
    {
        (void)hipStreamDestroy(d->stream);
        delete d;
        return hip_fail(e, "pipeline queues");
    }
    e = qr_pt_create(&d->pt);
    if (e != hipSuccess)
    {
        qr_wf_destroy(d->wf);
        (void)hipStreamDestroy(d->stream);
        delete d;
        return hip_fail(e, "path tracer queues");
    }
    /* the device address of the no-alloc entry point's scratch array */
    e = noalloc_setup(ordinal);
    if (e != hipSuccess)
    {
        qr_pt_destroy(d->pt);
        qr_wf_destroy(d->wf);
        (void)hipStreamDestroy(d->stream);
        delete d;
        return hip_fail(e, "hipGetSymbolAddress of the no-alloc scratch");
    }
    *out = d;
    return QRHIP_OK;
}

int qrhip_device_close(qrhip_dev *dev)
{
    if (dev == nullptr) return fail(QRHIP_E_ARG, "null device");
    if (has_gpu(dev))
    {
        (void)hipSetDevice(dev->ordinal);
        (void)hipStreamSynchronize(dev->stream);
        dev->scratch_tables.release();
        dev->scratch_frame.release();
        dev->scratch_ids.release();
        dev->scratch_pt_planes.release();
        dev->scratch_pt_seeds.release();
        qr_wf_destroy(dev->wf);
        qr_pt_destroy(dev->pt);
        (void)hipStreamDestroy(dev->stream);
    }
    delete dev;
    return QRHIP_OK;
}

int qrhip_device_set_fsaa(qrhip_dev *dev, int fsaa)
{
    if (dev == nullptr) return fail(QRHIP_E_ARG, "null device");
    /* rt_Platform::set_fsaa clamps to the supported maximum */
    if (fsaa < QRHIP_FSAA_NO) fsaa = QRHIP_FSAA_NO;
    if (fsaa > QRHIP_FSAA_4X) fsaa = QRHIP_FSAA_4X;
    dev->fsaa = fsaa;
    return fsaa;
}

int qrhip_device_get_fsaa(qrhip_dev *dev)
{
    return dev != nullptr ? dev->fsaa : fail(QRHIP_E_ARG, "null device");
}

int qrhip_device_get_fsaa_max(qrhip_dev *dev)
{
    return dev != nullptr ? QRHIP_FSAA_4X : fail(QRHIP_E_ARG, "null device");
}

int qrhip_device_set_path(qrhip_dev *dev, int path)
{
    if (dev == nullptr) return fail(QRHIP_E_ARG, "null device");
    if (path != QRHIP_PATH_LEVELS && path != QRHIP_PATH_ONE_KERNEL)
    {
        return fail(QRHIP_E_ARG, "unknown tracing path");
    }
    dev->path = path;
    return QRHIP_OK;
}

int qrhip_device_get_path(qrhip_dev *dev)
{
    return dev != nullptr ? dev->path : fail(QRHIP_E_ARG, "null device");
}

int qrhip_side_streams(void)
{
    return qr_wf_side_streams();
}

int qrhip_scene_create(qrhip_dev *dev, struct rt_SCENE *scn, int x_res,
                       int y_res, int x_row, uint32_t *frame, qrhip_scene **out)
{
    if (out == nullptr) return fail(QRHIP_E_ARG, "null output pointer");
    *out = nullptr;
    qrhip_scene *s = new (std::nothrow) qrhip_scene();
    if (s == nullptr) return fail(QRHIP_E_DEVICE, "out of host memory");
    try
    {
        s->eng = new qr::Engine(scn, x_res, y_res, x_row);
    }
    catch (const qr::Error &e)
    {
        delete s;
        return fail(QRHIP_E_SCENE, e.what());
    }
    catch (const std::bad_alloc &)
    {
        delete s;
        return fail(QRHIP_E_SCENE, "out of host memory");
    }
    s->dev = dev;
    s->x_res = x_res;
    s->y_res = y_res;
    s->x_row = x_row;
    if (frame == nullptr)
    {
        int ar = x_row < 0 ? -x_row : x_row;
        s->own_frame.assign((size_t)ar * y_res, 0u);
        s->frame = s->own_frame.data();
        if (x_row < 0) s->frame += (size_t)ar * (y_res - 1);
    }
    else
    {
        s->frame = frame;
    }
    if (has_gpu(dev))
    {
        hipError_t e = hipSetDevice(dev->ordinal);
        if (e == hipSuccess) e = s->dframe.reserve((size_t)x_res * y_res * 4);
        if (e == hipSuccess) e = hipHostMalloc((void **)&s->staging,
                                               (size_t)x_res * y_res * 4,
                                               hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreate(&s->ev0);
        if (e == hipSuccess) e = hipEventCreate(&s->ev1);
        if (e == hipSuccess) e = qr_wf_create(&s->wf);
        if (e != hipSuccess)
        {
            qrhip_scene_destroy(s);
            return hip_fail(e, "scene device set-up");
        }
    }
    *out = s;
    return QRHIP_OK;
}

int qrhip_scene_destroy(qrhip_scene *s)
{
    if (s == nullptr) return fail(QRHIP_E_ARG, "null scene");
    if (has_gpu(s->dev))
    {
        (void)hipSetDevice(s->dev->ordinal);
        (void)hipStreamSynchronize(s->dev->stream);
        s->dt.release();
        s->dframe.release();
        s->pt_planes.release();
        s->pt_seeds.release();
        if (s->staging != nullptr) (void)hipHostFree(s->staging);
        if (s->ev0 != nullptr) (void)hipEventDestroy(s->ev0);
        if (s->ev1 != nullptr) (void)hipEventDestroy(s->ev1);
        qr_wf_destroy(s->wf);
        qr_pt_destroy(s->pt);
    }
    delete s->eng;
    delete s;
    return QRHIP_OK;
}

int qrhip_scene_update(qrhip_scene *s, int64_t time)
{
    if (s == nullptr) return fail(QRHIP_E_ARG, "null scene");
    /* with the scene's tables and programs on the device, the host runs
     * the transforms alone and the device derives the records (the device
     * update stage); a changed structure goes back to the host */
    const bool on_dev = has_gpu(s->dev) && s->dt.holds_build(*s->eng);
    bool fast = false;
    try
    {
        if (on_dev) fast = s->eng->update_transforms((rt_time)time);
        else s->eng->update((rt_time)time);
    }
    catch (const qr::Error &e)
    {
        return fail(QRHIP_E_SCENE, e.what());
    }
    s->fsaa = s->dev != nullptr ? s->dev->fsaa : s->fsaa;
    s->eng->frame_desc(&s->fd, s->fsaa, s->depth);
    s->fd.x_row = s->x_res;     /* device frame pitch */
    s->updated = true;
    s->time = time;
    s->eng_time = time;
    if (!has_gpu(s->dev)) return QRHIP_OK;
    HIP_TRY(hipSetDevice(s->dev->ordinal));
    if (fast)
    {
        bool ok = false;
        HIP_TRY(s->dt.update_device(*s->eng, s->dev->stream, &ok));
        if (ok) return QRHIP_OK;
        /* an element's box flags changed: the programs are rebuilt */
        try
        {
            s->eng->materialize();
        }
        catch (const qr::Error &e)
        {
            return fail(QRHIP_E_SCENE, e.what());
        }
        s->dt.build_on_dev = 0;
    }
    const ProgramInputs in = program_inputs(*s->eng);
    HIP_TRY(s->dt.upload(s->eng->tables(), s->dev->stream, &in));
    return QRHIP_OK;
}

int qrhip_scene_program_stats(qrhip_scene *s, int *builds, int *refits, int *mismatch)
{
    if (s == nullptr) return fail(QRHIP_E_ARG, "null scene");
    if (builds != nullptr) *builds = (int)s->eng->program_build();
    if (refits != nullptr) *refits = s->dt.refits;
    if (mismatch == nullptr) return QRHIP_OK;
    *mismatch = 0;
    if (!has_gpu(s->dev) || !s->updated) return QRHIP_OK;
    try
    {
        s->eng->materialize();
    }
    catch (const qr::Error &e)
    {
        return fail(QRHIP_E_SCENE, e.what());
    }
    const qr_scene_tables &h = s->eng->tables();
    const bool any = h.n_step_any > 0 && h.n_step > 0;
    if (h.n_step == 0 || s->dt.dev_n_step != h.n_step) return QRHIP_OK;
    HIP_TRY(hipSetDevice(s->dev->ordinal));
    HIP_TRY(hipStreamSynchronize(s->dev->stream));
    for (int k = 0; k < (any ? 2 : 1); k++)
    {
        /* the steps and the end step */
        const int n = (k == 0 ? h.n_step : h.n_step_any) + 1;
        std::vector<qr_dev_step> want, got((size_t)n);
        dev_steps(k == 0 ? h.steps : h.steps_any, n - 1, s->eng->cull_ref(), h.elems,
                  h.surfaces, want);
        HIP_TRY(hipMemcpy(got.data(), k == 0 ? s->dt.step.ptr : s->dt.step_any.ptr,
                          sizeof(qr_dev_step) * (size_t)n, hipMemcpyDeviceToHost));
        for (int p = 0; p < n; p++)
        {
            if (memcmp(&got[(size_t)p], &want[(size_t)p], sizeof(qr_dev_step)) != 0) (*mismatch)++;
        }
    }
    /* the records and materials (the device update stage writes them) */
    std::vector<qr_surface> srf((size_t)h.n_surf);
    std::vector<qr_material> mat((size_t)h.n_mat);
    if (h.n_surf > 0)
    {
        HIP_TRY(hipMemcpy(srf.data(), s->dt.srf.ptr, sizeof(qr_surface) * srf.size(),
                          hipMemcpyDeviceToHost));
    }
    if (h.n_mat > 0)
    {
        HIP_TRY(hipMemcpy(mat.data(), s->dt.mat.ptr, sizeof(qr_material) * mat.size(),
                          hipMemcpyDeviceToHost));
    }
    for (int r = 0; r < h.n_surf; r++)
    {
        if (memcmp(&srf[(size_t)r], &h.surfaces[r], sizeof(qr_surface)) != 0) (*mismatch)++;
    }
    for (int m = 0; m < h.n_mat; m++)
    {
        if (memcmp(&mat[(size_t)m], &h.materials[m], sizeof(qr_material)) != 0) (*mismatch)++;
    }
    return QRHIP_OK;
}

int qrhip_scene_launch(qrhip_scene *s)
{
    if (s == nullptr) return fail(QRHIP_E_ARG, "null scene");
    if (!has_gpu(s->dev)) return fail(QRHIP_E_NODEVICE, "scene has no device");
    if (!s->updated) return fail(QRHIP_E_ARG, "scene not updated");
    if (s->pt_on != 0)
    {
        /* path tracer: pt_on passes per call (rt_Scene::render_slice,
         * core/engine/engine.cpp:3674), planes restarted when the AA mode
         * changes (3054-3057) or the path tracer was switched on */
        const int ns = s->fd.fsaa == 2 ? 4 : s->fd.fsaa == 1 ? 2 : 1;
        const size_t slots = (size_t)s->x_res * s->y_res * ns;
        if (s->pt_fsaa != s->fd.fsaa) s->pt_reset = true;
        if (s->pt_reset || s->pt_slots != slots)
        {
            std::vector<uint32_t> seeds;
            pt_seed_plane(seeds, s->x_res, s->y_res, s->x_row, ns);
            std::vector<float> zero(3 * slots, 0.0f);
            HIP_TRY(s->pt_seeds.reserve(sizeof(uint32_t) * slots));
            HIP_TRY(s->pt_planes.reserve(sizeof(float) * 3 * slots));
            HIP_TRY(hipMemcpyAsync(s->pt_seeds.ptr, seeds.data(), sizeof(uint32_t) * slots,
                                   hipMemcpyHostToDevice, s->dev->stream));
            HIP_TRY(hipMemcpyAsync(s->pt_planes.ptr, zero.data(), sizeof(float) * 3 * slots,
                                   hipMemcpyHostToDevice, s->dev->stream));
            HIP_TRY(hipStreamSynchronize(s->dev->stream));
            s->pts_c = 0.0f;
            s->pt_slots = slots;
            s->pt_reset = false;
            s->pt_fsaa = s->fd.fsaa;
        }
        if (s->pt == nullptr) HIP_TRY(qr_pt_create(&s->pt));
        /* the pt_on passes in one launch: the first counts pts_c + 1 */
        qr_frame_desc d = s->fd;
        d.pt_on = s->pt_on;
        d.pts_c = s->pts_c + 1.0f;
        HIP_TRY(qr_pt_launch(s->pt, &s->dt.t, &d, (float *)s->pt_planes.ptr,
                             (uint32_t *)s->pt_seeds.ptr, (uint32_t *)s->dframe.ptr, 0, 0,
                             s->pt_on, s->dev->stream));
        s->pts_c = s->pts_c + (float)s->pt_on;
        s->pt_pending = true;
        return QRHIP_OK;
    }
    HIP_TRY(launch_frame(s->dev, s->wf, &s->dt.t, &s->fd, (uint32_t *)s->dframe.ptr,
                         nullptr));
    s->pending = true;
    return QRHIP_OK;
}

/* wait for the scene's path-traced passes; trace the last run of passes
 * again while it overflowed a queue (it wrote nothing) */
static hipError_t pt_finish(const qrhip_dev *d, qr_pt_state *pt, bool pending)
{
    hipError_t e = hipStreamSynchronize(d->stream);
    if (e != hipSuccess || !pending) return e;
    for (int attempt = 0; qr_pt_check(pt) != 0; attempt++)
    {
        if (attempt == 16) return hipErrorOutOfMemory;
        if ((e = qr_pt_retry(pt, d->stream)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(d->stream)) != hipSuccess) return e;
    }
    return hipSuccess;
}

/* wait for the scene's frames; trace again if a queue had to grow */
static hipError_t scene_finish(qrhip_scene *s)
{
    if (s->pt_pending)
    {
        s->pt_pending = false;
        hipError_t e = pt_finish(s->dev, s->pt, true);
        if (e != hipSuccess) return e;
    }
    const bool pending = s->pending;
    s->pending = false;
    return finish_frames(s->dev, s->wf, pending, [s]() {
        return launch_frame(s->dev, s->wf, &s->dt.t, &s->fd,
                            (uint32_t *)s->dframe.ptr, nullptr);
    });
}

int qrhip_scene_sync(qrhip_scene *s)
{
    if (s == nullptr) return fail(QRHIP_E_ARG, "null scene");
    if (!has_gpu(s->dev)) return fail(QRHIP_E_NODEVICE, "scene has no device");
    HIP_TRY(scene_finish(s));
    return QRHIP_OK;
}

int qrhip_scene_download(qrhip_scene *s)
{
    if (s == nullptr) return fail(QRHIP_E_ARG, "null scene");
    if (!has_gpu(s->dev)) return fail(QRHIP_E_NODEVICE, "scene has no device");
    size_t bytes = (size_t)s->x_res * s->y_res * 4;
    HIP_TRY(scene_finish(s));
    HIP_TRY(hipMemcpyAsync(s->staging, s->dframe.ptr, bytes,
                           hipMemcpyDeviceToHost, s->dev->stream));
    HIP_TRY(hipStreamSynchronize(s->dev->stream));
    for (int y = 0; y < s->y_res; y++)
    {
        memcpy(s->frame + (ptrdiff_t)y * s->x_row,
               s->staging + (size_t)y * s->x_res, (size_t)s->x_res * 4);
    }
    return QRHIP_OK;
}

int qrhip_scene_render(qrhip_scene *s, int64_t time)
{
    int rc = qrhip_scene_update(s, time);
    if (rc != QRHIP_OK) return rc;
    if ((rc = qrhip_scene_launch(s)) != QRHIP_OK) return rc;
    return qrhip_scene_download(s);
}

int qrhip_scene_time_launches(qrhip_scene *s, int n, float *ms)
{
    if (s == nullptr || ms == nullptr || n <= 0) return fail(QRHIP_E_ARG, "bad argument");
    if (!has_gpu(s->dev)) return fail(QRHIP_E_NODEVICE, "scene has no device");
    HIP_TRY(scene_finish(s));
    /* identical frames: if the last one fitted its queues, all did; else the
     * queues have grown and the measurement is repeated */
    for (int round = 0; round < 16; round++)
    {
        HIP_TRY(hipEventRecord(s->ev0, s->dev->stream));
        for (int i = 0; i < n; i++)
        {
            int rc = qrhip_scene_launch(s);
            if (rc != QRHIP_OK) return rc;
        }
        HIP_TRY(hipEventRecord(s->ev1, s->dev->stream));
        HIP_TRY(hipEventSynchronize(s->ev1));
        s->pending = false;
        if (s->pt_pending)
        {
            /* an overflowed run of passes wrote nothing, but it was counted:
             * the planes restart at the next launch */
            s->pt_pending = false;
            if (qr_pt_check(s->pt) != 0)
            {
                s->pt_reset = true;
                continue;
            }
        }
        if (qr_wf_check(s->wf, nullptr) != 0) continue;
        float total = 0.0f;
        HIP_TRY(hipEventElapsedTime(&total, s->ev0, s->ev1));
        *ms = total / (float)n;
        return QRHIP_OK;
    }
    return fail(QRHIP_E_DEVICE, "pipeline queues kept overflowing");
}

uint32_t *qrhip_scene_get_frame(qrhip_scene *s)
{
    return s != nullptr ? s->frame : nullptr;
}

int qrhip_scene_get_x_row(qrhip_scene *s)
{
    return s != nullptr ? s->x_row : fail(QRHIP_E_ARG, "null scene");
}

int qrhip_scene_get_opts(qrhip_scene *s)
{
    return s != nullptr ? s->eng->get_opts() : fail(QRHIP_E_ARG, "null scene");
}

int qrhip_scene_set_opts(qrhip_scene *s, int opts)
{
    return s != nullptr ? s->eng->set_opts(opts) : fail(QRHIP_E_ARG, "null scene");
}

int qrhip_scene_get_pton(qrhip_scene *s)
{
    return s != nullptr ? s->pt_on : fail(QRHIP_E_ARG, "null scene");
}

int qrhip_scene_set_pton(qrhip_scene *s, int pton)
{
    if (s == nullptr) return fail(QRHIP_E_ARG, "null scene");
    /* a scene can prohibit the path tracer (rt_Scene::set_pton) */
    if ((s->eng->get_opts() & RT_OPTS_PT) == 0)
    {
        if (pton > 0 && s->pt_on == 0) s->pt_reset = true;
        s->pt_on = pton < 0 ? 0 : pton;
    }
    return s->pt_on;
}

int qrhip_scene_update_action(qrhip_scene *s, int64_t time, int action)
{
    if (s == nullptr) return fail(QRHIP_E_ARG, "null scene");
    if (action < QRHIP_CAMERA_CROUCH || action > QRHIP_CAMERA_ROTATE_RIGHT)
        return fail(QRHIP_E_ARG, "unknown camera action");
    s->eng->update_action((rt_time)time, action);
    return QRHIP_OK;
}

int qrhip_scene_next_cam(qrhip_scene *s)
{
    return s != nullptr ? s->eng->next_cam() : fail(QRHIP_E_ARG, "null scene");
}

int qrhip_scene_get_cam_idx(qrhip_scene *s)
{
    return s != nullptr ? s->eng->get_cam_idx() : fail(QRHIP_E_ARG, "null scene");
}

int qrhip_set_data_path(const char *root)
{
    if (root == nullptr) return fail(QRHIP_E_ARG, "null path");
    qr::set_data_root(root);
    return QRHIP_OK;
}

int qrhip_image_load(const char *path, uint32_t *texels, int64_t max, int *x_dim,
                     int *y_dim)
{
    if (path == nullptr || x_dim == nullptr || y_dim == nullptr)
        return fail(QRHIP_E_ARG, "null argument");
    std::vector<uint32_t> t;
    int x = 0, y = 0;
    if (!qr::bmp_load(path, t, x, y)) return fail(QRHIP_E_SCENE, "failed to load image");
    *x_dim = x;
    *y_dim = y;
    if (texels != nullptr)
    {
        if ((int64_t)t.size() > max) return fail(QRHIP_E_ARG, "image larger than the buffer");
        memcpy(texels, t.data(), sizeof(uint32_t) * t.size());
    }
    return QRHIP_OK;
}

int qrhip_image_save(const char *path, const uint32_t *pixels, int x_dim, int y_dim,
                     int64_t stride)
{
    if (path == nullptr || pixels == nullptr) return fail(QRHIP_E_ARG, "null argument");
    if (x_dim <= 0 || y_dim == 0 || x_dim > 65536 || y_dim < -65536 || y_dim > 65536)
        return fail(QRHIP_E_ARG, "image size out of range");
    if (stride == 0) stride = x_dim;
    if (stride < x_dim && -stride < x_dim) return fail(QRHIP_E_ARG, "row stride below width");
    if (!qr::bmp_save(path, pixels, x_dim, y_dim, (ptrdiff_t)stride))
        return fail(QRHIP_E_SCENE, "failed to save image");
    return QRHIP_OK;
}

int qrhip_scene_save_frame(qrhip_scene *s, int index)
{
    if (s == nullptr) return fail(QRHIP_E_ARG, "null scene");
    if (index < 0) return fail(QRHIP_E_ARG, "negative frame index");
    const std::string path = qr::data_root() + "dump/" + qr::frame_name(index);
    if (!qr::bmp_save(path, s->frame, s->x_res, -s->y_res, (ptrdiff_t)s->x_row))
        return fail(QRHIP_E_SCENE, "failed to save image");
    return QRHIP_OK;
}

int qrhip_scene_get_pose(qrhip_scene *s, float *pose)
{
    if (s == nullptr || pose == nullptr) return fail(QRHIP_E_ARG, "null argument");
    s->eng->camera_pose(pose);
    return QRHIP_OK;
}

int qrhip_scene_set_depth(qrhip_scene *s, int depth)
{
    if (s == nullptr) return fail(QRHIP_E_ARG, "null scene");
    if (depth < 0 || depth > 14) return fail(QRHIP_E_ARG, "depth out of range 0..14");
    s->depth = depth;
    return depth;
}

int qrhip_scene_get_depth(qrhip_scene *s)
{
    return s != nullptr ? s->depth : fail(QRHIP_E_ARG, "null scene");
}

int qrhip_scene_tables(qrhip_scene *s, const qr_scene_tables **tables,
                       qr_frame_desc *fd)
{
    if (s == nullptr || tables == nullptr) return fail(QRHIP_E_ARG, "null argument");
    if (!s->updated) return fail(QRHIP_E_ARG, "scene not updated");
    try
    {
        s->eng->materialize();      /* the host's copy of the device update */
    }
    catch (const qr::Error &e)
    {
        return fail(QRHIP_E_SCENE, e.what());
    }
    *tables = &s->eng->tables();
    if (fd != nullptr) *fd = s->fd;
    return QRHIP_OK;
}

int qrhip_render_tables(qrhip_dev *dev, const qr_scene_tables *tab,
                        const qr_frame_desc *fd, uint32_t *frame, int32_t *ids)
{
    if (!has_gpu(dev) || tab == nullptr || fd == nullptr || frame == nullptr)
    {
        return fail(QRHIP_E_ARG, "null argument");
    }
    int vrc = validate_tables(tab, fd);
    if (vrc != QRHIP_OK) return vrc;

    HIP_TRY(hipSetDevice(dev->ordinal));
    HIP_TRY(dev->scratch_tables.upload(*tab, dev->stream));
    const int ns = fd->fsaa == 2 ? 4 : fd->fsaa == 1 ? 2 : 1;
    size_t fbytes = (size_t)fd->x_res * fd->y_res * 4;
    size_t ibytes = fbytes * ns;
    HIP_TRY(dev->scratch_frame.reserve(fbytes));
    if (ids != nullptr) HIP_TRY(dev->scratch_ids.reserve(ibytes));
    qr_frame_desc d = *fd;
    d.x_row = fd->x_res;
    uint32_t *dfr = (uint32_t *)dev->scratch_frame.ptr;
    int32_t *dids = ids != nullptr ? (int32_t *)dev->scratch_ids.ptr : nullptr;
    HIP_TRY(launch_frame(dev, dev->wf, &dev->scratch_tables.t, &d, dfr, dids));
    HIP_TRY(finish_frames(dev, dev->wf, true, [&]() {
        return launch_frame(dev, dev->wf, &dev->scratch_tables.t, &d, dfr, dids);
    }));
    std::vector<uint32_t> tmp((size_t)fd->x_res * fd->y_res);
    HIP_TRY(hipMemcpyAsync(tmp.data(), dev->scratch_frame.ptr, fbytes,
                           hipMemcpyDeviceToHost, dev->stream));
    if (ids != nullptr)
    {
        HIP_TRY(hipMemcpyAsync(ids, dev->scratch_ids.ptr, ibytes,
                               hipMemcpyDeviceToHost, dev->stream));
    }
    HIP_TRY(hipStreamSynchronize(dev->stream));
    for (int y = 0; y < fd->y_res; y++)
    {
        memcpy(frame + (ptrdiff_t)y * fd->x_row, tmp.data() + (size_t)y * fd->x_res,
               (size_t)fd->x_res * 4);
    }
    return QRHIP_OK;
}

int qrhip_render_tables_pt(qrhip_dev *dev, const qr_scene_tables *tab,
                           const qr_frame_desc *fd, float *planes,
                           uint32_t *seeds, uint32_t *frame)
{
    if (!has_gpu(dev) || tab == nullptr || fd == nullptr || planes == nullptr ||
        seeds == nullptr || frame == nullptr)
    {
        return fail(QRHIP_E_ARG, "null argument");
    }
    int vrc = validate_tables(tab, fd);
    if (vrc != QRHIP_OK) return vrc;
    if (!(fd->pts_c >= 1.0f)) return fail(QRHIP_E_ARG, "pass count below 1");

    HIP_TRY(hipSetDevice(dev->ordinal));
    HIP_TRY(dev->scratch_tables.upload(*tab, dev->stream));
    const int ns = fd->fsaa == 2 ? 4 : fd->fsaa == 1 ? 2 : 1;
    const size_t slots = (size_t)fd->x_res * fd->y_res * ns;
    const size_t fbytes = (size_t)fd->x_res * fd->y_res * 4;
    HIP_TRY(dev->scratch_frame.reserve(fbytes));
    HIP_TRY(dev->scratch_pt_planes.reserve(sizeof(float) * 3 * slots));
    HIP_TRY(dev->scratch_pt_seeds.reserve(sizeof(uint32_t) * slots));
    HIP_TRY(hipMemcpyAsync(dev->scratch_pt_planes.ptr, planes, sizeof(float) * 3 * slots,
                           hipMemcpyHostToDevice, dev->stream));
    HIP_TRY(hipMemcpyAsync(dev->scratch_pt_seeds.ptr, seeds, sizeof(uint32_t) * slots,
                           hipMemcpyHostToDevice, dev->stream));
    qr_frame_desc d = *fd;
    d.x_row = fd->x_res;
    HIP_TRY(qr_pt_launch(dev->pt, &dev->scratch_tables.t, &d,
                         (float *)dev->scratch_pt_planes.ptr,
                         (uint32_t *)dev->scratch_pt_seeds.ptr,
                         (uint32_t *)dev->scratch_frame.ptr, 0, 0, 1, dev->stream));
    HIP_TRY(pt_finish(dev, dev->pt, true));
    std::vector<uint32_t> tmp((size_t)fd->x_res * fd->y_res);
    HIP_TRY(hipMemcpyAsync(tmp.data(), dev->scratch_frame.ptr, fbytes,
                           hipMemcpyDeviceToHost, dev->stream));
    HIP_TRY(hipMemcpyAsync(planes, dev->scratch_pt_planes.ptr, sizeof(float) * 3 * slots,
                           hipMemcpyDeviceToHost, dev->stream));
    HIP_TRY(hipMemcpyAsync(seeds, dev->scratch_pt_seeds.ptr, sizeof(uint32_t) * slots,
                           hipMemcpyDeviceToHost, dev->stream));
    HIP_TRY(hipStreamSynchronize(dev->stream));
    for (int y = 0; y < fd->y_res; y++)
    {
        memcpy(frame + (ptrdiff_t)y * fd->x_row, tmp.data() + (size_t)y * fd->x_res,
               (size_t)fd->x_res * 4);
    }
    return QRHIP_OK;
}

int qrhip_pt_seed_plane(int x_res, int y_res, int x_row, int fsaa, uint32_t *seeds)
{
    if (seeds == nullptr || x_res <= 0 || y_res <= 0 || (x_row < 0 ? -x_row : x_row) < x_res ||
        fsaa < 0 || fsaa > 2)
    {
        return fail(QRHIP_E_ARG, "bad argument");
    }
    std::vector<uint32_t> v;
    pt_seed_plane(v, x_res, y_res, x_row, fsaa == 2 ? 4 : fsaa == 1 ? 2 : 1);
    memcpy(seeds, v.data(), v.size() * sizeof(uint32_t));
    return QRHIP_OK;
}

static std::mutex g_gen_mu;
static std::map<struct rt_SCENE *, qrgen::Scene *> g_gen;

struct rt_SCENE *qrgen_create(int config, uint64_t seed, int flags, int *info)
{
    if (config < 1 || config > 64 ||
        (flags & ~(QRGEN_IDS | QRGEN_TEXTURE_FILES | QRGEN_ANIMATED)) != 0)
    {
        fail(QRHIP_E_ARG, "unknown workload");
        return nullptr;
    }
    qrgen::Scene *sc = qrgen::build(config, seed);
    if (flags & QRGEN_TEXTURE_FILES)
    {
        const std::string prefix = "qrgen_" + std::to_string(config) + "_" +
                                   std::to_string(seed) + "_";
        if (!qrgen::file_textures(*sc, prefix))
        {
            delete sc;
            fail(QRHIP_E_SCENE, "failed to save image");
            return nullptr;
        }
    }
    if (flags & QRGEN_IDS) qrgen::override_ids(*sc);
    if (flags & QRGEN_ANIMATED) qrgen::animate(*sc);
    if (info != nullptr)
    {
        info[0] = sc->x_res;
        info[1] = sc->y_res;
        info[2] = sc->fsaa;
        info[3] = sc->depth;
        info[4] = sc->pt;
    }
    std::lock_guard<std::mutex> lk(g_gen_mu);
    g_gen[&sc->root] = sc;
    return &sc->root;
}

void qrgen_destroy(struct rt_SCENE *scn)
{
    std::lock_guard<std::mutex> lk(g_gen_mu);
    auto it = g_gen.find(scn);
    if (it == g_gen.end()) return;
    delete it->second;
    g_gen.erase(it);
}

} /* extern "C" */
