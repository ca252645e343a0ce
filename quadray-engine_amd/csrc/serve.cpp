/*
 * serve.cpp -- the serving loop of include/qrhip.h: a server thread keeps its
 * scenes resident in HBM and traces frame requests as they arrive.
 *
 * Requests wait in arrival order and are admitted while their footprint fits
 * the budget beside the requests in flight.  An admitted frame is split into
 * chunks (runs of level-0 tiles).  Each step takes, within a budget of camera
 * samples, a chunk of the oldest request in flight and then chunks of the
 * requests with the least work left; it traces its Whitted chunks as one
 * batch of the level pipeline (one launch per stage for all of them) and one
 * pass of each path-traced chunk.  Steps are pipelined: step N+1 is enqueued before step N
 * is retired, so the device never waits for the host's bookkeeping, and the
 * frames that finish in step N are copied back on a second stream while step
 * N+1 runs.  Two step contexts (queues, frame arguments) alternate.  A
 * request's frame does not depend on the others in its steps: each chunk has
 * its own queues and each frame its own planes and arguments.  No counterpart
 * in QuadRay, which renders one rt_Scene frame per render() call.
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
    DevBuf    frame;                    /* device frame, pitch x_res */
    DevBuf    planes, seeds;            /* path tracer */
    uint32_t *host = nullptr;           /* pinned: frame copy back */
    size_t    host_cap = 0;
    uint32_t *staged = nullptr;         /* pinned: seed plane upload */
    size_t    staged_cap = 0;
    bool      busy = false;
};

/* a run of level-0 tiles of a frame: traced once (Whitted) or advanced by one
 * path-tracer pass per step */
struct Chunk
{
    uint32_t tile0 = 0, ntiles = 0;
    int      passes = 1;                /* passes the chunk needs */
    int      enqueued = 0;              /* passes launched */
    int      finished = 0;              /* passes retired */
    uint64_t cost = 0;                  /* samples of one pass */
};

struct Req
{
    qrhip_request r;
    uint64_t      bytes = 0;
    int           slot = -1;
    int           steps = 0;
    int           chunks_left = 0;      /* chunks with passes not retired */
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

/* one chunk pass of a step */
struct Unit
{
    Req  *q;
    int   c;
    float pts_c;                        /* pass number (path tracer) */
};

/* a step in flight: its chunk passes, queues and launch context */
struct StepCtx
{
    std::vector<Unit>          units;
    std::vector<qr_wf_item>    items;
    std::vector<qr_wf_state *> pool;
    std::vector<qr_pt_state *> pt_pool;     /* one per path-traced unit */
    qr_wf_batch               *wb = nullptr;
    hipEvent_t                 ev = nullptr;
    bool                       pending = false;
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
    b += n0 * (uint64_t)std::max(1, n_lgt) * 12;            /* shadow jobs, 3 sets */
    return b;
}

hipError_t pinned_reserve(uint32_t *&p, size_t &cap, size_t n)
{
    if (cap >= n) return hipSuccess;
    if (p != nullptr) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    hipError_t e = hipHostMalloc((void **)&p, sizeof(uint32_t) * n, hipHostMallocDefault);
    if (e == hipSuccess) cap = n;
    return e;
}

}  /* namespace */

struct qrhip_server
{
    qrhip_dev                *dev = nullptr;
    std::vector<SceneState>   scenes;
    uint64_t                  budget = 0, used = 0, peak = 0;
    int                       max_active = 1;
    uint64_t                  step_samples = 0;
    hipStream_t               stream = nullptr;     /* kernels */
    hipStream_t               copies = nullptr;     /* frames back to the host */
    hipEvent_t                copied = nullptr;
    StepCtx                   ctx[2];
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
    hipError_t admit_one(Req *q);
    void pick(std::vector<Unit> &units);
    hipError_t launch(StepCtx &c);
    hipError_t retire(StepCtx &c);
    hipError_t finish(const std::vector<Req *> &fin);
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

/* claim a slot; set the request's tables, camera, chunks and path-tracer
 * start (enqueued on the kernel stream, ahead of the request's first pass) */
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
        const ProgramInputs in = program_inputs(*sc.s->eng);
        if ((e = sc.s->dt.upload(sc.s->eng->tables(), stream, &in)) != hipSuccess) return e;
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
    if ((e = pinned_reserve(sl.host, sl.host_cap, px)) != hipSuccess) return e;

    /* chunks: runs of level-0 tiles of at most the step's samples */
    const int ns = q->r.fsaa == 2 ? 4 : q->r.fsaa == 1 ? 2 : 1;
    const bool pt = q->r.spp > 0;
    const uint32_t tiles = pt ? qr_pt_tiles(q->r.x_res, q->r.y_res)
                              : qr_wf_batch_tiles(ctx[0].wb, q->r.x_res, q->r.y_res);
    const uint64_t per_tile = 64ull * ns * (pt ? 2 : 1);
    uint32_t run = tiles;
    if (step_samples > 0) run = (uint32_t)std::max<uint64_t>(1, step_samples / per_tile);
    q->chunks.clear();
    for (uint32_t t = 0; t < tiles; t += run)
    {
        Chunk c;
        c.tile0 = t;
        c.ntiles = std::min(run, tiles - t);
        c.passes = pt ? q->r.spp : 1;
        c.cost = (uint64_t)c.ntiles * per_tile;
        q->chunks.push_back(c);
    }
    q->chunks_left = (int)q->chunks.size();

    if (pt)
    {
        /* the seed plane and cleared colour planes; the pinned staging copy
         * stays untouched until the slot is reassigned */
        const size_t slots_n = px * (size_t)ns;
        std::vector<uint32_t> seeds;
        pt_seed_plane(seeds, q->r.x_res, q->r.y_res, q->r.x_res, ns);
        if ((e = pinned_reserve(sl.staged, sl.staged_cap, slots_n)) != hipSuccess) return e;
        memcpy(sl.staged, seeds.data(), sizeof(uint32_t) * slots_n);
        if ((e = sl.seeds.reserve(sizeof(uint32_t) * slots_n)) != hipSuccess) return e;
        if ((e = sl.planes.reserve(sizeof(float) * 3 * slots_n)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(sl.seeds.ptr, sl.staged, sizeof(uint32_t) * slots_n,
                                hipMemcpyHostToDevice, stream)) != hipSuccess)
        {
            return e;
        }
        if ((e = hipMemsetAsync(sl.planes.ptr, 0, sizeof(float) * 3 * slots_n, stream)) !=
            hipSuccess)
        {
            return e;
        }
    }
    return hipSuccess;
}

/* the step's chunk passes within the step's samples: first a chunk of the
 * oldest request in flight (nothing starves), then chunks of the requests
 * with the least work left (short frames finish first); a chunk gives at
 * most one pass per step, and a path-traced chunk none while its last pass
 * is in flight */
void qrhip_server::pick(std::vector<Unit> &units)
{
    units.clear();
    if (active.empty()) return;
    uint64_t used_s = 0;
    const uint64_t cap = step_samples > 0 ? step_samples : ~0ull;
    std::vector<std::pair<uint64_t, size_t>> order;
    for (size_t k = 0; k < active.size(); k++)
    {
        uint64_t left = 0;
        for (const Chunk &ch : active[k]->chunks)
        {
            left += (uint64_t)(ch.passes - ch.enqueued) * ch.cost;
        }
        order.push_back({left, k});
    }
    std::stable_sort(order.begin() + 1, order.end());
    for (const auto &o : order)
    {
        Req *q = active[o.second];
        for (int c = 0; c < (int)q->chunks.size(); c++)
        {
            Chunk &ch = q->chunks[c];
            if (ch.enqueued == ch.passes) continue;
            /* a path-traced chunk's next pass reads the planes its last pass
             * writes, and a pass that overflowed its queues is traced again
             * when its step retires: one pass of a chunk in flight at a time */
            if (q->r.spp > 0 && ch.enqueued != ch.finished) continue;
            if (!units.empty() && used_s + ch.cost > cap) break;
            ch.enqueued++;
            units.push_back(Unit{q, c, (float)ch.enqueued});
            used_s += ch.cost;
        }
        if (used_s >= cap) break;
    }
}

/* enqueue a step: its Whitted chunks as one batch, then one pass of each
 * path-traced chunk, then the step's event */
hipError_t qrhip_server::launch(StepCtx &c)
{
    hipError_t e;
    c.items.clear();
    for (const Unit &u : c.units)
    {
        if (u.q->r.spp > 0) continue;
        if (c.items.size() == c.pool.size())
        {
            qr_wf_state *st = nullptr;
            if ((e = qr_wf_create(&st)) != hipSuccess) return e;
            c.pool.push_back(st);
        }
        const Chunk &ch = u.q->chunks[u.c];
        qr_wf_item it;
        it.s = c.pool[c.items.size()];
        it.t = &scenes[u.q->r.scene].s->dt.t;
        it.fd = &u.q->fd;
        it.frame = (uint32_t *)slots[u.q->slot].frame.ptr;
        it.ids = nullptr;
        it.tile0 = ch.tile0;
        it.ntiles = ch.ntiles;
        c.items.push_back(it);
    }
    if (!c.items.empty())
    {
        e = qr_wf_batch_launch(c.wb, c.items.data(), (int)c.items.size(), stream);
        if (e != hipSuccess) return e;
    }
    size_t npt = 0;
    for (const Unit &u : c.units)
    {
        if (u.q->r.spp == 0) continue;
        if (npt == c.pt_pool.size())
        {
            qr_pt_state *st = nullptr;
            if ((e = qr_pt_create(&st)) != hipSuccess) return e;
            c.pt_pool.push_back(st);
        }
        const Chunk &ch = u.q->chunks[u.c];
        Slot &sl = slots[u.q->slot];
        qr_frame_desc d = u.q->fd;
        d.pt_on = 1;
        d.pts_c = u.pts_c;
        e = qr_pt_launch(c.pt_pool[npt++], &scenes[u.q->r.scene].s->dt.t, &d,
                         (float *)sl.planes.ptr, (uint32_t *)sl.seeds.ptr,
                         (uint32_t *)sl.frame.ptr, ch.tile0, ch.ntiles, 1, stream);
        if (e != hipSuccess) return e;
    }
    if ((e = hipEventRecord(c.ev, stream)) != hipSuccess) return e;
    c.pending = true;
    return hipSuccess;
}

/* wait for a step; if one of its Whitted chunks overflowed a queue (the
 * queues have grown), trace its Whitted chunks again; then count its passes */
hipError_t qrhip_server::retire(StepCtx &c)
{
    hipError_t e;
    for (int attempt = 0;; attempt++)
    {
        if ((e = hipEventSynchronize(c.ev)) != hipSuccess) return e;
        if (c.items.empty() || qr_wf_batch_check(c.wb, c.items.data(), (int)c.items.size()) == 0)
        {
            break;
        }
        if (attempt == 16) return hipErrorOutOfMemory;
        e = qr_wf_batch_launch(c.wb, c.items.data(), (int)c.items.size(), stream);
        if (e != hipSuccess) return e;
        if ((e = hipEventRecord(c.ev, stream)) != hipSuccess) return e;
    }
    /* a path-traced pass that overflowed its queues wrote nothing: trace it
     * again with the grown queues */
    size_t npt = 0;
    for (const Unit &u : c.units)
    {
        if (u.q->r.spp == 0) continue;
        qr_pt_state *st = c.pt_pool[npt++];
        for (int attempt = 0; qr_pt_check(st) != 0; attempt++)
        {
            if (attempt == 16) return hipErrorOutOfMemory;
            if ((e = qr_pt_retry(st, stream)) != hipSuccess) return e;
            if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
        }
    }
    for (const Unit &u : c.units)
    {
        Chunk &ch = u.q->chunks[u.c];
        if (++ch.finished == ch.passes) u.q->chunks_left--;
    }
    c.pending = false;
    return hipSuccess;
}

/* copy the finished frames back on the copy stream (their last writes were
 * in the retired step; the next step keeps the device busy meanwhile) into
 * the callers' buffers */
hipError_t qrhip_server::finish(const std::vector<Req *> &fin)
{
    hipError_t e;
    if (fin.empty()) return hipSuccess;
    for (Req *q : fin)
    {
        Slot &sl = slots[q->slot];
        const size_t bytes = sizeof(uint32_t) * (size_t)q->r.x_res * q->r.y_res;
        if ((e = hipMemcpyAsync(sl.host, sl.frame.ptr, bytes, hipMemcpyDeviceToHost,
                                copies)) != hipSuccess)
        {
            return e;
        }
    }
    if ((e = hipEventRecord(copied, copies)) != hipSuccess) return e;
    if ((e = hipEventSynchronize(copied)) != hipSuccess) return e;
    for (Req *q : fin)
    {
        if (q->r.frame != nullptr)
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
    int cur = 0;
    StepCtx *prev = nullptr;
    for (;;)
    {
        std::unique_lock<std::mutex> lk(mu);
        if (prev == nullptr)
        {
            cv.wait(lk, [this] { return stop || !waiting.empty() || !active.empty(); });
            if (waiting.empty() && active.empty()) break;      /* stop, drained */
        }
        const double t_begin = now_ms();

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
        lk.unlock();

        hipError_t e = hipSuccess;
        for (Req *q : admitted)
        {
            if (e == hipSuccess) e = admit_one(q);
            active.push_back(q);
        }

        /* the next step goes in before the previous one is retired */
        StepCtx &c = ctx[cur];
        bool launched = false;
        uint64_t reqs_in_step = 0;
        if (e == hipSuccess)
        {
            pick(c.units);
            if (!c.units.empty())
            {
                e = launch(c);
                launched = e == hipSuccess;
                for (Req *q : active)
                {
                    for (const Unit &u : c.units)
                    {
                        if (u.q == q)
                        {
                            q->steps++;
                            reqs_in_step++;
                            break;
                        }
                    }
                }
            }
        }
        if (e == hipSuccess && prev != nullptr) e = retire(*prev);

        std::vector<Req *> fin, keep;
        for (Req *q : active)
        {
            if (e != hipSuccess || q->chunks_left == 0) fin.push_back(q);
            else keep.push_back(q);
        }
        if (e == hipSuccess) e = finish(fin);
        if (e != hipSuccess)
        {
            /* a device error ends every request in flight */
            (void)hipStreamSynchronize(stream);
            fin = active;
            keep.clear();
            launched = false;
            c.pending = false;
            if (prev != nullptr) prev->pending = false;
        }
        active.swap(keep);
        const double t_end = now_ms();

        lk.lock();
        if (launched) steps++;
        step_max = std::max<uint64_t>(step_max, reqs_in_step);
        busy_ms += t_end - t_begin;
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

        prev = launched ? &c : nullptr;
        if (launched) cur ^= 1;
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
    auto release = [s]()
    {
        for (StepCtx &c : s->ctx)
        {
            qr_wf_batch_destroy(c.wb);
            if (c.ev != nullptr) (void)hipEventDestroy(c.ev);
        }
        if (s->copied != nullptr) (void)hipEventDestroy(s->copied);
        if (s->stream != nullptr) (void)hipStreamDestroy(s->stream);
        if (s->copies != nullptr) (void)hipStreamDestroy(s->copies);
        delete s;
    };
    hipError_t e = hipSetDevice(dev->ordinal);
    if (e == hipSuccess) e = hipStreamCreate(&s->stream);
    if (e == hipSuccess) e = hipStreamCreate(&s->copies);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->copied, hipEventDisableTiming);
    for (StepCtx &c : s->ctx)
    {
        if (e == hipSuccess) e = qr_wf_batch_create(&c.wb);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c.ev, hipEventDisableTiming);
    }
    /* the scenes' uploads were enqueued on the device stream */
    if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
    if (e != hipSuccess)
    {
        release();
        return hip_fail(e, "qrhip_server_create");
    }
    s->t0 = Clock::now();
    try
    {
        s->th = std::thread([s] { s->run(); });
    }
    catch (...)
    {
        release();
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
    (void)hipStreamSynchronize(s->copies);
    for (StepCtx &c : s->ctx)
    {
        for (qr_wf_state *st : c.pool) qr_wf_destroy(st);
        for (qr_pt_state *st : c.pt_pool) qr_pt_destroy(st);
        qr_wf_batch_destroy(c.wb);
        if (c.ev != nullptr) (void)hipEventDestroy(c.ev);
    }
    for (Slot &sl : s->slots)
    {
        sl.frame.release();
        sl.planes.release();
        sl.seeds.release();
        if (sl.host != nullptr) (void)hipHostFree(sl.host);
        if (sl.staged != nullptr) (void)hipHostFree(sl.staged);
    }
    if (s->copied != nullptr) (void)hipEventDestroy(s->copied);
    (void)hipStreamDestroy(s->stream);
    (void)hipStreamDestroy(s->copies);
    const int err = s->error;
    delete s;
    return err;
}

} /* extern "C" */
