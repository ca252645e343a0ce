/*
 * serve.cpp -- the serving loop of include/qrhip.h: a server thread with its
 * own HIP stream keeps its scenes resident in HBM and traces frame requests
 * as they arrive.
 *
 * Requests wait in arrival order.  Each step admits, from the head of the
 * queue, the requests whose footprint fits the budget beside those in flight
 * (and a free frame slot), traces every admitted Whitted frame as one batch of
 * the level pipeline (one launch per stage for all of them) and one pass of
 * every admitted path-traced frame, and returns the frames that finished.
 * Requests arriving during a step join the next step; path-traced frames stay
 * in flight across steps.  A request's frame does not depend on the others in
 * its steps: each frame has its own queues, planes and arguments.  No
 * counterpart in QuadRay, which renders one rt_Scene frame per render() call.
 */
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include <hip/hip_runtime.h>

#include "capi_internal.h"

using namespace qrhip_impl;

namespace {

using Clock = std::chrono::steady_clock;

struct Slot
{
    DevBuf       frame;                 /* device frame, pitch x_res */
    DevBuf       planes, seeds;         /* path tracer */
    uint32_t    *host = nullptr;        /* pinned copy of the frame */
    size_t       host_cap = 0;
    bool         busy = false;
};

/* a run of level-0 tiles of a frame, traced (Whitted) or advanced by one
 * path-tracer pass per step */
struct Chunk
{
    uint32_t tile0 = 0, ntiles = 0;
    int      passes_left = 1;
    float    pts_c = 0.0f;              /* passes done (path tracer) */
    uint64_t cost = 0;                  /* samples of one pass */
};

struct Req
{
    qrhip_request r;
    uint64_t      bytes = 0;
    int           slot = -1;
    int           steps = 0;
    int           chunks_left = 0;
    bool          update = false;       /* run the scene's update stage first */
    double        submit_ms = 0.0, admit_ms = 0.0;
    qr_frame_desc fd;
    std::vector<Chunk> chunks;
};

struct SceneState
{
    qrhip_scene *s = nullptr;
    int64_t      time = 0;              /* time of the tables on the device */
    int          in_flight = 0;         /* requests using those tables */
};

/* one chunk's work in a step */
struct Unit
{
    Req *q;
    int  c;
};

uint64_t footprint(const qrhip_request &r, int n_lgt, uint64_t step_samples)
{
    const uint64_t ns = r.fsaa == 2 ? 4 : r.fsaa == 1 ? 2 : 1;
    const uint64_t px = (uint64_t)r.x_res * (uint64_t)r.y_res;
    uint64_t b = 4 * px;                                    /* device frame */
    if (r.spp > 0) return b + 16 * px * ns;                 /* planes, seeds */
    /* queues of the largest chunk the request is traced in */
    uint64_t n0 = ((uint64_t)r.x_res + 31) * ((uint64_t)r.y_res + 31) * ns;
    if (step_samples > 0) n0 = std::min(n0, step_samples + 64 * ns);
    b += n0 * (7 * 16 + 4);                                 /* level 0 records */
    b += (uint64_t)r.depth * (n0 / 8 + 256) * (9 * 16 + 4); /* deeper levels */
    b += n0 * (uint64_t)std::max(1, n_lgt) * 4;             /* shadow jobs */
    return b;
}

}  /* namespace */

struct qrhip_server
{
    qrhip_dev                *dev = nullptr;
    std::vector<SceneState>   scenes;
    uint64_t                  budget = 0, used = 0, peak = 0;
    int                       max_active = 1;
    uint64_t                  step_samples = 0;
    hipStream_t               stream = nullptr;
    qr_wf_batch              *wb = nullptr;
    std::vector<qr_wf_state *> pool;     /* queues of the step's Whitted chunks */
    std::vector<Slot>         slots;
    Clock::time_point         t0;

    std::mutex                mu;
    std::condition_variable   cv, done_cv;
    std::deque<Req *>         waiting;
    std::deque<qrhip_result>  done;
    std::vector<Req *>        active;    /* admitted, oldest first */
    bool                      stop = false;
    uint64_t                  steps = 0, frames = 0, step_max = 0;
    double                    busy_ms = 0.0;
    int                       error = QRHIP_OK;
    std::thread               th;

    double now_ms() const
    {
        return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    }
    void run();
    hipError_t step(const std::vector<Unit> &units);
    hipError_t admit_one(Req *q);
    void pick(std::vector<Unit> &units);
    int free_slot();
};

int qrhip_server::free_slot()
{
    for (size_t k = 0; k < slots.size(); k++)
    {
        if (!slots[k].busy) return (int)k;
    }
    slots.emplace_back();
    return (int)slots.size() - 1;
}

/* claim a slot, set the request's camera, tables, chunks and path-tracer
 * start */
hipError_t qrhip_server::admit_one(Req *q)
{
    SceneState &sc = scenes[q->r.scene];
    hipError_t e;
    if (q->update)
    {
        /* the update stage for the request's time; no request in flight
         * used the scene's tables when it was admitted */
        try
        {
            sc.s->eng->update((rt_time)q->r.time);
        }
        catch (const qr::Error &)
        {
            return hipErrorInvalidValue;
        }
        if ((e = sc.s->dt.upload(sc.s->eng->tables(), stream)) != hipSuccess) return e;
    }
    const int k = free_slot();
    Slot &sl = slots[k];
    sl.busy = true;
    q->slot = k;
    sc.s->eng->frame_desc(&q->fd, q->r.fsaa, q->r.depth, q->r.x_res, q->r.y_res,
                          q->r.use_pose ? q->r.pose : nullptr);
    q->fd.x_row = q->r.x_res;
    const size_t px = (size_t)q->r.x_res * q->r.y_res;
    if ((e = sl.frame.reserve(sizeof(uint32_t) * px)) != hipSuccess) return e;
    if (sl.host_cap < px)
    {
        if (sl.host != nullptr) (void)hipHostFree(sl.host);
        sl.host = nullptr;
        sl.host_cap = 0;
        if ((e = hipHostMalloc((void **)&sl.host, sizeof(uint32_t) * px,
                               hipHostMallocDefault)) != hipSuccess)
        {
            return e;
        }
        sl.host_cap = px;
    }

    /* chunks: runs of level-0 tiles of at most the step's samples */
    const int ns = q->r.fsaa == 2 ? 4 : q->r.fsaa == 1 ? 2 : 1;
    const bool pt = q->r.spp > 0;
    const uint32_t tiles = pt ? qr_pt_tiles(q->r.x_res, q->r.y_res)
                              : qr_wf_batch_tiles(wb, q->r.x_res, q->r.y_res);
    const uint64_t per_tile = 64ull * ns * (pt ? 2 : 1);
    uint32_t run = tiles;
    if (step_samples > 0) run = (uint32_t)std::max<uint64_t>(1, step_samples / per_tile);
    q->chunks.clear();
    for (uint32_t t = 0; t < tiles; t += run)
    {
        Chunk c;
        c.tile0 = t;
        c.ntiles = std::min(run, tiles - t);
        c.passes_left = pt ? q->r.spp : 1;
        c.cost = (uint64_t)c.ntiles * per_tile;
        q->chunks.push_back(c);
    }
    q->chunks_left = (int)q->chunks.size();

    if (pt)
    {
        const size_t slots_n = px * (size_t)ns;
        std::vector<uint32_t> seeds;
        pt_seed_plane(seeds, q->r.x_res, q->r.y_res, q->r.x_res, ns);
        if ((e = sl.seeds.reserve(sizeof(uint32_t) * slots_n)) != hipSuccess) return e;
        if ((e = sl.planes.reserve(sizeof(float) * 3 * slots_n)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(sl.seeds.ptr, seeds.data(), sizeof(uint32_t) * slots_n,
                                hipMemcpyHostToDevice, stream)) != hipSuccess)
        {
            return e;
        }
        if ((e = hipMemsetAsync(sl.planes.ptr, 0, sizeof(float) * 3 * slots_n, stream)) !=
            hipSuccess)
        {
            return e;
        }
        /* the staged seeds must reach the device before "seeds" goes away */
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
    }
    return hipSuccess;
}

/* the step's chunks: one of every request in flight, oldest first, then
 * more of the oldest, within the step's samples (at least one chunk) */
void qrhip_server::pick(std::vector<Unit> &units)
{
    units.clear();
    uint64_t used_s = 0;
    const uint64_t cap = step_samples > 0 ? step_samples : ~0ull;
    std::vector<int> next(active.size(), 0);
    for (int round = 0; round < 2; round++)
    {
        for (size_t k = 0; k < active.size(); k++)
        {
            Req *q = active[k];
            int &c = next[k];
            while (c < (int)q->chunks.size())
            {
                if (q->chunks[c].passes_left == 0) { c++; continue; }
                const uint64_t cost = q->chunks[c].cost;
                if (!units.empty() && used_s + cost > cap) break;
                units.push_back(Unit{q, c});
                used_s += cost;
                c++;
                if (round == 0) break;
            }
        }
        if (used_s >= cap) break;
    }
}

/* one step: the Whitted chunks as one batch, one pass of each path-traced
 * chunk */
hipError_t qrhip_server::step(const std::vector<Unit> &units)
{
    hipError_t e;
    std::vector<qr_wf_item> items;
    for (const Unit &u : units)
    {
        if (u.q->r.spp > 0) continue;
        if (items.size() == pool.size())
        {
            qr_wf_state *st = nullptr;
            if ((e = qr_wf_create(&st)) != hipSuccess) return e;
            pool.push_back(st);
        }
        const Chunk &c = u.q->chunks[u.c];
        qr_wf_item it;
        it.s = pool[items.size()];
        it.t = &scenes[u.q->r.scene].s->dt.t;
        it.fd = &u.q->fd;
        it.frame = (uint32_t *)slots[u.q->slot].frame.ptr;
        it.ids = nullptr;
        it.tile0 = c.tile0;
        it.ntiles = c.ntiles;
        items.push_back(it);
    }
    for (int attempt = 0;; attempt++)
    {
        if (!items.empty())
        {
            e = qr_wf_batch_launch(wb, items.data(), (int)items.size(), stream);
            if (e != hipSuccess) return e;
        }
        if (attempt == 0)
        {
            for (const Unit &u : units)
            {
                if (u.q->r.spp == 0) continue;
                Chunk &c = u.q->chunks[u.c];
                Slot &sl = slots[u.q->slot];
                qr_frame_desc d = u.q->fd;
                d.pt_on = 1;
                d.pts_c = c.pts_c + 1.0f;
                e = qr_launch_pt_tiles(&scenes[u.q->r.scene].s->dt.t, &d,
                                       (float *)sl.planes.ptr, (uint32_t *)sl.seeds.ptr,
                                       (uint32_t *)sl.frame.ptr, c.tile0, c.ntiles, stream);
                if (e != hipSuccess) return e;
            }
        }
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
        if (items.empty() || qr_wf_batch_check(wb, items.data(), (int)items.size()) == 0) break;
        if (attempt == 16) return hipErrorOutOfMemory;
        /* a queue grew: trace the Whitted chunks again */
    }
    for (const Unit &u : units)
    {
        Chunk &c = u.q->chunks[u.c];
        c.pts_c = c.pts_c + 1.0f;
        if (--c.passes_left == 0) u.q->chunks_left--;
    }
    /* frames that finished */
    bool any = false;
    for (Req *q : active)
    {
        if (q->chunks_left > 0) continue;
        Slot &sl = slots[q->slot];
        const size_t bytes = sizeof(uint32_t) * (size_t)q->r.x_res * q->r.y_res;
        if ((e = hipMemcpyAsync(sl.host, sl.frame.ptr, bytes, hipMemcpyDeviceToHost,
                                stream)) != hipSuccess)
        {
            return e;
        }
        any = true;
    }
    if (any && (e = hipStreamSynchronize(stream)) != hipSuccess) return e;
    for (Req *q : active)
    {
        if (q->chunks_left == 0 && q->r.frame != nullptr)
        {
            memcpy(q->r.frame, slots[q->slot].host,
                   sizeof(uint32_t) * (size_t)q->r.x_res * q->r.y_res);
        }
    }
    return hipSuccess;
}

void qrhip_server::run()
{
    (void)hipSetDevice(dev->ordinal);
    std::vector<Unit> units;
    for (;;)
    {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [this] { return stop || !waiting.empty() || !active.empty(); });
        if (waiting.empty() && active.empty()) break;      /* stop, drained */

        /* admission from the head of the queue */
        std::vector<Req *> admitted;
        while (!waiting.empty() && (int)(active.size() + admitted.size()) < max_active)
        {
            Req *q = waiting.front();
            SceneState &sc = scenes[q->r.scene];
            if (used + q->bytes > budget) break;
            /* one time per scene in flight */
            if (sc.in_flight > 0 && q->r.time != sc.time) break;
            waiting.pop_front();
            used += q->bytes;
            peak = std::max(peak, used);
            q->update = q->r.time != sc.time;
            sc.time = q->r.time;
            sc.in_flight++;
            q->admit_ms = now_ms();
            admitted.push_back(q);
        }
        if (admitted.empty() && active.empty())
        {
            /* nothing in flight and the head cannot start: cannot happen
             * with the submit checks, but never spin on it */
            Req *q = waiting.front();
            waiting.pop_front();
            qrhip_result res;
            memset(&res, 0, sizeof(res));
            res.tag = q->r.tag;
            res.status = QRHIP_E_ARG;
            res.submit_ms = q->submit_ms;
            res.admit_ms = res.done_ms = now_ms();
            done.push_back(res);
            frames++;
            delete q;
            lk.unlock();
            done_cv.notify_all();
            continue;
        }
        lk.unlock();

        const double t_step = now_ms();
        hipError_t e = hipSuccess;
        for (Req *q : admitted)
        {
            if (e == hipSuccess) e = admit_one(q);
            active.push_back(q);
        }
        if (e == hipSuccess)
        {
            pick(units);
            e = step(units);
        }
        const double t_end = now_ms();

        std::vector<Req *> fin;
        std::vector<Req *> keep;
        for (Req *q : active)
        {
            bool served = false;
            for (const Unit &u : units) served |= u.q == q;
            if (served) q->steps++;
            if (e != hipSuccess || q->chunks_left == 0) fin.push_back(q);
            else keep.push_back(q);
        }
        uint64_t reqs_in_step = 0;
        {
            std::vector<const Req *> seen;
            for (const Unit &u : units)
            {
                if (std::find(seen.begin(), seen.end(), u.q) == seen.end()) seen.push_back(u.q);
            }
            reqs_in_step = seen.size();
        }
        active.swap(keep);

        lk.lock();
        steps++;
        step_max = std::max<uint64_t>(step_max, reqs_in_step);
        busy_ms += t_end - t_step;
        if (e != hipSuccess) error = QRHIP_E_DEVICE;
        for (Req *q : fin)
        {
            qrhip_result res;
            res.tag = q->r.tag;
            res.status = e == hipSuccess ? QRHIP_OK : QRHIP_E_DEVICE;
            res.steps = q->steps;
            res.submit_ms = q->submit_ms;
            res.admit_ms = q->admit_ms;
            res.done_ms = t_end;
            done.push_back(res);
            used -= q->bytes;
            if (q->slot >= 0) slots[q->slot].busy = false;
            scenes[q->r.scene].in_flight--;
            frames++;
            delete q;
        }
        lk.unlock();
        done_cv.notify_all();
    }
}

extern "C" {

int qrhip_server_create(qrhip_dev *dev, qrhip_scene *const *scenes, int n_scenes,
                        uint64_t hbm_budget, int max_active, uint64_t step_samples,
                        qrhip_server **out)
{
    if (out == nullptr || (n_scenes > 0 && scenes == nullptr)) return fail(QRHIP_E_ARG, "null argument");
    *out = nullptr;
    if (!has_gpu(dev)) return fail(QRHIP_E_NODEVICE, "a server needs a HIP device");
    if (n_scenes < 1 || max_active < 1 || hbm_budget == 0)
        return fail(QRHIP_E_ARG, "server needs scenes, a budget and room for a request");
    for (int k = 0; k < n_scenes; k++)
    {
        if (scenes[k] == nullptr || scenes[k]->dev != dev || !scenes[k]->updated)
            return fail(QRHIP_E_ARG, "scene not updated on the server's device");
        if (scenes[k]->dt.t.n_lgt > qr_wf_max_lights())
            return fail(QRHIP_E_UNSUPPORTED, "served scenes take at most 32 lights");
    }
    qrhip_server *s = new (std::nothrow) qrhip_server();
    if (s == nullptr) return fail(QRHIP_E_DEVICE, "out of host memory");
    s->dev = dev;
    s->budget = hbm_budget;
    s->max_active = max_active;
    s->step_samples = step_samples;
    for (int k = 0; k < n_scenes; k++)
    {
        SceneState st;
        st.s = scenes[k];
        st.time = scenes[k]->time;
        s->scenes.push_back(st);
    }
    hipError_t e = hipSetDevice(dev->ordinal);
    if (e == hipSuccess) e = hipStreamCreate(&s->stream);
    if (e == hipSuccess) e = qr_wf_batch_create(&s->wb);
    if (e != hipSuccess)
    {
        if (s->stream != nullptr) (void)hipStreamDestroy(s->stream);
        qr_wf_batch_destroy(s->wb);
        delete s;
        return hip_fail(e, "qrhip_server_create");
    }
    /* the scenes' uploads were enqueued on the device stream */
    e = hipStreamSynchronize(dev->stream);
    if (e != hipSuccess)
    {
        (void)hipStreamDestroy(s->stream);
        qr_wf_batch_destroy(s->wb);
        delete s;
        return hip_fail(e, "qrhip_server_create");
    }
    s->t0 = Clock::now();
    try
    {
        s->th = std::thread([s] { s->run(); });
    }
    catch (...)
    {
        (void)hipStreamDestroy(s->stream);
        qr_wf_batch_destroy(s->wb);
        delete s;
        return fail(QRHIP_E_DEVICE, "cannot start the server thread");
    }
    *out = s;
    return QRHIP_OK;
}

int qrhip_server_footprint(qrhip_server *s, const qrhip_request *req, uint64_t *bytes)
{
    if (s == nullptr || req == nullptr || bytes == nullptr) return fail(QRHIP_E_ARG, "null argument");
    if (req->scene < 0 || req->scene >= (int)s->scenes.size())
        return fail(QRHIP_E_ARG, "scene index out of range");
    *bytes = footprint(*req, s->scenes[req->scene].s->dt.t.n_lgt, s->step_samples);
    return QRHIP_OK;
}

int qrhip_server_submit(qrhip_server *s, const qrhip_request *req)
{
    if (s == nullptr || req == nullptr) return fail(QRHIP_E_ARG, "null argument");
    const qrhip_request &r = *req;
    if (r.scene < 0 || r.scene >= (int)s->scenes.size())
        return fail(QRHIP_E_ARG, "scene index out of range");
    if (r.x_res < 1 || r.y_res < 1 || r.x_res > 16384 || r.y_res > 16384)
        return fail(QRHIP_E_ARG, "frame size out of range");
    if (r.fsaa < QRHIP_FSAA_NO || r.fsaa > QRHIP_FSAA_4X) return fail(QRHIP_E_ARG, "unknown AA mode");
    if (r.depth < 0 || r.depth > 14) return fail(QRHIP_E_ARG, "depth out of range 0..14");
    if (r.spp < 0 || r.spp > 1 << 20) return fail(QRHIP_E_ARG, "samples out of range");
    Req *q = new (std::nothrow) Req();
    if (q == nullptr) return fail(QRHIP_E_DEVICE, "out of host memory");
    q->r = r;
    q->bytes = footprint(r, s->scenes[r.scene].s->dt.t.n_lgt, s->step_samples);
    if (q->bytes > s->budget)
    {
        delete q;
        return fail(QRHIP_E_ARG, "request needs more device memory than the server's budget");
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop)
        {
            delete q;
            return fail(QRHIP_E_ARG, "server is stopping");
        }
        q->submit_ms = s->now_ms();
        s->waiting.push_back(q);
    }
    s->cv.notify_one();
    return QRHIP_OK;
}

int qrhip_server_poll(qrhip_server *s, qrhip_result *out, int max, int wait_ms)
{
    if (s == nullptr || (max > 0 && out == nullptr)) return fail(QRHIP_E_ARG, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && wait_ms > 0)
    {
        s->done_cv.wait_for(lk, std::chrono::milliseconds(wait_ms),
                            [s] { return !s->done.empty(); });
    }
    int n = 0;
    while (n < max && !s->done.empty())
    {
        out[n++] = s->done.front();
        s->done.pop_front();
    }
    return n;
}

int qrhip_server_get_stats(qrhip_server *s, qrhip_server_stats *st)
{
    if (s == nullptr || st == nullptr) return fail(QRHIP_E_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    st->steps = s->steps;
    st->frames = s->frames;
    st->step_frames_max = s->step_max;
    st->bytes_in_use = s->used;
    st->bytes_peak = s->peak;
    st->busy_ms = s->busy_ms;
    return s->error;
}

int qrhip_server_destroy(qrhip_server *s)
{
    if (s == nullptr) return QRHIP_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->th.joinable()) s->th.join();
    (void)hipStreamSynchronize(s->stream);
    for (qr_wf_state *st : s->pool) qr_wf_destroy(st);
    for (Slot &sl : s->slots)
    {
        sl.frame.release();
        sl.planes.release();
        sl.seeds.release();
        if (sl.host != nullptr) (void)hipHostFree(sl.host);
    }
    qr_wf_batch_destroy(s->wb);
    (void)hipStreamDestroy(s->stream);
    const int err = s->error;
    delete s;
    return err;
}

} /* extern "C" */
