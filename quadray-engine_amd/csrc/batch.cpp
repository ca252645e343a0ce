/*
 * batch.cpp -- the C ABI of frame batches (include/qrhip.h): many frames of
 * uneven resolution, scene, anti-aliasing and depth traced by one launch per
 * pipeline stage, optionally replayed from a captured HIP graph, with an
 * autotuner over the launch shape.  No counterpart in QuadRay, which renders
 * one rt_Scene frame per render() call.
 */
#include <cstring>
#include <new>
#include <vector>

#include <hip/hip_runtime.h>

#include "capi_internal.h"

using namespace qrhip_impl;

struct qrhip_batch
{
    qrhip_dev *dev = nullptr;
    std::vector<qrhip_frame_req> req;
    std::vector<qr_frame_desc>   fd;
    std::vector<qr_wf_state *>   st;        /* each frame's queues */
    std::vector<DevBuf>          dfr;       /* each frame's device frame */
    std::vector<qr_wf_item>      items;
    std::vector<qr_wf_batch *>   solo;      /* batches of one (serial mode) */
    qr_wf_batch   *wb = nullptr;
    hipGraph_t     graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int            pending = -1;            /* mode of unchecked launches */
    int            grid_per_cu = 64, tile_w = 8, rpt = 1, cadence = 1;
    hipEvent_t     ev0 = nullptr, ev1 = nullptr;
};

namespace {

void drop_graph(qrhip_batch *b)
{
    if (b->exec != nullptr) (void)hipGraphExecDestroy(b->exec);
    if (b->graph != nullptr) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

void drop_frames(qrhip_batch *b)
{
    drop_graph(b);
    for (qr_wf_batch *s : b->solo) qr_wf_batch_destroy(s);
    for (qr_wf_state *s : b->st) qr_wf_destroy(s);
    for (DevBuf &d : b->dfr) d.release();
    b->solo.clear();
    b->st.clear();
    b->dfr.clear();
    b->items.clear();
    b->req.clear();
    b->fd.clear();
    b->pending = -1;
}

/* the batched launches' shape; the serial launches (the unbatched
 * reference) keep the single-frame shape a scene's own launch has */
hipError_t apply_config(qrhip_batch *b)
{
    return qr_wf_batch_config(b->wb, b->grid_per_cu, b->tile_w, b->rpt, b->cadence);
}

hipError_t launch_direct(qrhip_batch *b, int mode)
{
    const hipStream_t st = b->dev->stream;
    const int n = (int)b->items.size();
    if (mode == QRHIP_BATCH_SERIAL)
    {
        for (int k = 0; k < n; k++)
        {
            /* a frame's batch of one is made at the first serial launch,
             * with the single-frame launch shape */
            hipError_t e = hipSuccess;
            if (b->solo[k] == nullptr)
            {
                e = qr_wf_batch_create(&b->solo[k]);
                if (e != hipSuccess) return e;
            }
            e = qr_wf_batch_launch(b->solo[k], &b->items[k], 1, st);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    return qr_wf_batch_launch(b->wb, b->items.data(), n, st);
}

/* queues overflowed in the last launches of "mode"? (they have grown) */
bool grew(qrhip_batch *b, int mode)
{
    const int n = (int)b->items.size();
    if (mode == QRHIP_BATCH_SERIAL)
    {
        bool g = false;
        for (int k = 0; k < n; k++)
        {
            g |= b->solo[k] != nullptr && qr_wf_batch_check(b->solo[k], &b->items[k], 1) != 0;
        }
        return g;
    }
    return qr_wf_batch_check(b->wb, b->items.data(), n) != 0;
}

/* wait for the stream; retrace while a queue had to grow */
hipError_t finish(qrhip_batch *b)
{
    hipError_t e = hipStreamSynchronize(b->dev->stream);
    if (e != hipSuccess || b->pending < 0) return e;
    int mode = b->pending == QRHIP_BATCH_GRAPH ? QRHIP_BATCH_DIRECT : b->pending;
    b->pending = -1;
    for (int attempt = 0; grew(b, mode); attempt++)
    {
        if (attempt == 16) return hipErrorOutOfMemory;
        if ((e = launch_direct(b, mode)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(b->dev->stream)) != hipSuccess) return e;
    }
    return hipSuccess;
}

/* capture the direct launch sequence in a graph; the frame arguments and
 * queues are settled by a direct run first, so the captured launches only
 * enqueue kernels and the counter copy */
hipError_t capture(qrhip_batch *b)
{
    hipError_t e;
    if ((e = launch_direct(b, QRHIP_BATCH_DIRECT)) != hipSuccess) return e;
    b->pending = QRHIP_BATCH_DIRECT;
    if ((e = finish(b)) != hipSuccess) return e;
    const hipStream_t st = b->dev->stream;
    if ((e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal)) != hipSuccess) return e;
    hipError_t le = qr_wf_batch_launch(b->wb, b->items.data(), (int)b->items.size(), st);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(st, &g);
    if (le != hipSuccess)
    {
        if (g != nullptr) (void)hipGraphDestroy(g);
        return le;
    }
    if (e != hipSuccess) return e;
    b->graph = g;
    return hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
}

}  /* namespace */

extern "C" {

int qrhip_batch_create(qrhip_dev *dev, qrhip_batch **out)
{
    if (out == nullptr) return fail(QRHIP_E_ARG, "null argument");
    *out = nullptr;
    if (!has_gpu(dev)) return fail(QRHIP_E_NODEVICE, "batches need a HIP device");
    qrhip_batch *b = new (std::nothrow) qrhip_batch();
    if (b == nullptr) return fail(QRHIP_E_DEVICE, "out of host memory");
    b->dev = dev;
    hipError_t e = hipSetDevice(dev->ordinal);
    if (e == hipSuccess) e = qr_wf_batch_create(&b->wb);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e != hipSuccess)
    {
        qrhip_batch_destroy(b);
        return hip_fail(e, "qrhip_batch_create");
    }
    *out = b;
    return QRHIP_OK;
}

int qrhip_batch_destroy(qrhip_batch *b)
{
    if (b == nullptr) return QRHIP_OK;
    if (b->dev != nullptr) (void)hipStreamSynchronize(b->dev->stream);
    drop_frames(b);
    qr_wf_batch_destroy(b->wb);
    if (b->ev0 != nullptr) (void)hipEventDestroy(b->ev0);
    if (b->ev1 != nullptr) (void)hipEventDestroy(b->ev1);
    delete b;
    return QRHIP_OK;
}

int qrhip_batch_set(qrhip_batch *b, const qrhip_frame_req *req, int n)
{
    if (b == nullptr || (n > 0 && req == nullptr)) return fail(QRHIP_E_ARG, "null argument");
    if (n < 1 || n > QR_WF_MAX_BATCH) return fail(QRHIP_E_ARG, "batch size out of range");
    for (int k = 0; k < n; k++)
    {
        const qrhip_frame_req &r = req[k];
        if (r.scene == nullptr || r.scene->dev != b->dev)
            return fail(QRHIP_E_ARG, "frame without a scene on the batch's device");
        if (!r.scene->updated) return fail(QRHIP_E_ARG, "scene not updated");
        if (r.x_res < 1 || r.y_res < 1 || r.x_res > 16384 || r.y_res > 16384)
            return fail(QRHIP_E_ARG, "frame size out of range");
        if (r.fsaa < QRHIP_FSAA_NO || r.fsaa > QRHIP_FSAA_4X)
            return fail(QRHIP_E_ARG, "unknown AA mode");
        if (r.depth < 0 || r.depth > 14) return fail(QRHIP_E_ARG, "depth out of range 0..14");
        if (r.scene->dt.t.n_lgt > qr_wf_max_lights())
            return fail(QRHIP_E_UNSUPPORTED, "batched frames take at most 32 lights");
    }
    HIP_TRY(hipStreamSynchronize(b->dev->stream));
    drop_frames(b);
    b->req.assign(req, req + n);
    b->fd.resize((size_t)n);
    b->st.assign((size_t)n, nullptr);
    b->dfr.resize((size_t)n);
    b->solo.assign((size_t)n, nullptr);
    for (int k = 0; k < n; k++)
    {
        const qrhip_frame_req &r = req[k];
        r.scene->eng->frame_desc(&b->fd[k], r.fsaa, r.depth, r.x_res, r.y_res);
        b->fd[k].x_row = r.x_res;
        HIP_TRY(qr_wf_create(&b->st[k]));
        HIP_TRY(b->dfr[k].reserve(sizeof(uint32_t) * (size_t)r.x_res * r.y_res));
    }
    b->items.resize((size_t)n);
    for (int k = 0; k < n; k++)
    {
        b->items[k] = qr_wf_item{b->st[k], &req[k].scene->dt.t, &b->fd[k],
                                 (uint32_t *)b->dfr[k].ptr, nullptr};
    }
    HIP_TRY(apply_config(b));
    return QRHIP_OK;
}

int qrhip_batch_config(qrhip_batch *b, int grid_per_cu, int tile_w, int rays_per_thread,
                       int cadence)
{
    if (b == nullptr) return fail(QRHIP_E_ARG, "null batch");
    const bool ok_tile = tile_w == 4 || tile_w == 8 || tile_w == 16 || tile_w == 32;
    if (grid_per_cu < 1 || grid_per_cu > 64 || !ok_tile || rays_per_thread < 1 ||
        rays_per_thread > QRHIP_BATCH_MAX_RPT || cadence < 1 ||
        cadence > QRHIP_BATCH_MAX_CADENCE)
        return fail(QRHIP_E_ARG, "launch shape out of range");
    HIP_TRY(finish(b));
    b->grid_per_cu = grid_per_cu;
    b->tile_w = tile_w;
    b->rpt = rays_per_thread;
    b->cadence = cadence;
    drop_graph(b);
    HIP_TRY(apply_config(b));
    return QRHIP_OK;
}

int qrhip_batch_launch(qrhip_batch *b, int mode)
{
    if (b == nullptr) return fail(QRHIP_E_ARG, "null batch");
    if (b->items.empty()) return fail(QRHIP_E_ARG, "empty batch");
    if (mode < QRHIP_BATCH_DIRECT || mode > QRHIP_BATCH_SERIAL)
        return fail(QRHIP_E_ARG, "unknown batch mode");
    if (b->pending >= 0 && b->pending != mode) HIP_TRY(finish(b));
    if (mode == QRHIP_BATCH_GRAPH)
    {
        if (b->exec == nullptr) HIP_TRY(capture(b));
        HIP_TRY(hipGraphLaunch(b->exec, b->dev->stream));
    }
    else
    {
        HIP_TRY(launch_direct(b, mode));
    }
    b->pending = mode;
    return QRHIP_OK;
}

int qrhip_batch_sync(qrhip_batch *b)
{
    if (b == nullptr) return fail(QRHIP_E_ARG, "null batch");
    HIP_TRY(finish(b));
    return QRHIP_OK;
}

int qrhip_batch_download(qrhip_batch *b, int k)
{
    if (b == nullptr) return fail(QRHIP_E_ARG, "null batch");
    if (k >= (int)b->req.size()) return fail(QRHIP_E_ARG, "frame index out of range");
    HIP_TRY(finish(b));
    const int k0 = k < 0 ? 0 : k, k1 = k < 0 ? (int)b->req.size() : k + 1;
    for (int j = k0; j < k1; j++)
    {
        const qrhip_frame_req &r = b->req[j];
        if (r.frame == nullptr) continue;
        HIP_TRY(hipMemcpy(r.frame, b->dfr[j].ptr, sizeof(uint32_t) * (size_t)r.x_res * r.y_res,
                          hipMemcpyDeviceToHost));
    }
    return QRHIP_OK;
}

int qrhip_batch_time(qrhip_batch *b, int n, int mode, float *ms)
{
    if (b == nullptr || ms == nullptr || n <= 0) return fail(QRHIP_E_ARG, "bad argument");
    if (b->items.empty()) return fail(QRHIP_E_ARG, "empty batch");
    if (mode < QRHIP_BATCH_DIRECT || mode > QRHIP_BATCH_SERIAL)
        return fail(QRHIP_E_ARG, "unknown batch mode");
    HIP_TRY(finish(b));
    if (mode == QRHIP_BATCH_GRAPH && b->exec == nullptr) HIP_TRY(capture(b));
    const hipStream_t st = b->dev->stream;
    /* identical batches: if the last one fitted its queues, all did; else
     * the queues have grown and the measurement is repeated */
    for (int round = 0; round < 16; round++)
    {
        HIP_TRY(hipEventRecord(b->ev0, st));
        for (int i = 0; i < n; i++)
        {
            if (mode == QRHIP_BATCH_GRAPH) HIP_TRY(hipGraphLaunch(b->exec, st));
            else HIP_TRY(launch_direct(b, mode));
        }
        HIP_TRY(hipEventRecord(b->ev1, st));
        HIP_TRY(hipEventSynchronize(b->ev1));
        if (grew(b, mode == QRHIP_BATCH_GRAPH ? QRHIP_BATCH_DIRECT : mode))
        {
            b->pending = mode;
            HIP_TRY(finish(b));
            continue;
        }
        float total = 0.0f;
        HIP_TRY(hipEventElapsedTime(&total, b->ev0, b->ev1));
        *ms = total / (float)n;
        return QRHIP_OK;
    }
    return fail(QRHIP_E_DEVICE, "pipeline queues kept overflowing");
}

int qrhip_batch_tune(qrhip_batch *b, int reps, int *best, float *ms)
{
    if (b == nullptr || best == nullptr || ms == nullptr || reps <= 0)
        return fail(QRHIP_E_ARG, "bad argument");
    static const int grids[] = {8, 16, 32, 64};
    static const int tiles[] = {4, 8, 16, 32};
    static const int rpts[] = {1, 2, 4};
    static const int cadences[] = {1, 2};
    float best_ms = 0.0f;
    int cfg[4] = {b->grid_per_cu, b->tile_w, 1, 1};
    /* one shape: timed after one untimed launch (queues grow there) */
    auto try_shape = [&](int g, int t, int r, int c) -> int
    {
        int rc = qrhip_batch_config(b, g, t, r, c);
        if (rc != QRHIP_OK) return rc;
        float m = 0.0f;
        if ((rc = qrhip_batch_time(b, 1, QRHIP_BATCH_DIRECT, &m)) != QRHIP_OK) return rc;
        if ((rc = qrhip_batch_time(b, reps, QRHIP_BATCH_DIRECT, &m)) != QRHIP_OK) return rc;
        if (best_ms == 0.0f || m < best_ms)
        {
            best_ms = m;
            cfg[0] = g;
            cfg[1] = t;
            cfg[2] = r;
            cfg[3] = c;
        }
        return QRHIP_OK;
    };
    /* the grid and the tile shape first, then rays per thread and the
     * compaction cadence at the best of them */
    for (int g : grids)
    {
        for (int t : tiles)
        {
            const int rc = try_shape(g, t, 1, 1);
            if (rc != QRHIP_OK) return rc;
        }
    }
    const int g0 = cfg[0], t0 = cfg[1];
    for (int r : rpts)
    {
        for (int c : cadences)
        {
            if (r == 1 && c == 1) continue;
            const int rc = try_shape(g0, t0, r, c);
            if (rc != QRHIP_OK) return rc;
        }
    }
    int rc = qrhip_batch_config(b, cfg[0], cfg[1], cfg[2], cfg[3]);
    if (rc != QRHIP_OK) return rc;
    for (int k = 0; k < 4; k++) best[k] = cfg[k];
    *ms = best_ms;
    return QRHIP_OK;
}

} /* extern "C" */
