/*
 * serve.cpp -- the serving loop of include/qrhip.h: a server thread keeps its
 * scenes resident in HBM and traces frame requests as they arrive.
 *
 * Requests wait in arrival order and are admitted while their footprint fits
 * the budget beside the requests in flight.  Each admitted request gets its
 * own copy of its scene's tables on the device, made by the scene's update
 * stage for the request's time, so requests for one scene at different
 * times are in flight together.  An admitted frame is split into chunks
 * (runs of level-0 tiles).
 *
 * Each step takes chunks of the requests in flight in order of their
 * deadlines (submit time plus 16 ms for Whitted frames up to 720 rows, 48 ms
 * for larger ones, 16 ms a pass for path-traced ones), within a budget of
 * estimated device time and a cap of camera samples; it traces its Whitted
 * chunks as one batch of the level pipeline (one launch per stage for all
 * of them) and one run of passes of each path-traced chunk (as many passes
 * as the budget leaves room for).  The estimate is each scene's measured
 * device time per sample, Whitted and path-traced apart: every retired step
 * times its Whitted batch and its path-traced runs with events and scales
 * the rates of the scenes in it by measured over predicted.
 *
 * Steps are pipelined: step N+1 is enqueued before step N is retired, so
 * the device never waits for the host's bookkeeping, and the frames that
 * finish in step N are copied back on a second stream while step N+1 runs.
 * Two step contexts (queues, frame arguments) alternate.  A request's frame
 * does not depend on the others in its steps: each chunk has its own queues
 * and each frame its own tables, planes and arguments.  No counterpart in
 * QuadRay, which renders one rt_Scene frame per render() call.
 */
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include <hip/hip_runtime.h>

#include "capi_internal.h"

using namespace qrhip_impl;

namespace {

using Clock = std::chrono::steady_clock;

struct Slot
{
    DevTables dt;                       /* the request's scene tables */
    int       dt_scene = -1;            /* scene and time they hold */
    int64_t   dt_time = 0;
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
    double        submit_ms = 0.0, admit_ms = 0.0;
    double        deadline_ms = 0.0;    /* order of service */
    double        device_ms = 0.0;      /* measured device time, attributed */
    qr_frame_desc fd;
    std::vector<Chunk> chunks;
};

/* a scene's engine and its measured device time per sample */
struct SceneState
{
    qrhip_scene *s = nullptr;
    int64_t      time = 0;              /* time of the engine's last update */
    double       wt_ns = 0.5;           /* Whitted: ns per camera sample */
    double       pt_ns = 2.0;           /* path tracer: ns per sample and pass */
};

/* fixed device time of a step's Whitted batch (about 22 launches) and of a
 * path-traced run (reset, bounces, fold), in ms */
constexpr double WHITTED_STEP_MS = 0.15;
constexpr double PT_RUN_MS = 0.10;

/* one chunk of a step: a Whitted chunk, or npass passes of a path-traced
 * chunk from pass pts_c on */
struct Unit
{
    Req  *q;
    int   c;
    float pts_c;                        /* first pass number (path tracer) */
    int   npass;
    double est_ms;                      /* predicted device time */
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
    hipEvent_t                 t0 = nullptr, t1 = nullptr, t2 = nullptr;  /* timing */
    bool                       pending = false;
};

/* bytes of a scene's tables on the device */
uint64_t tables_bytes(const qr_device_tables &t)
{
    return sizeof(qr_surface) * (uint64_t)t.n_surf + sizeof(qr_material) * (uint64_t)t.n_mat +
           sizeof(qr_light) * (uint64_t)t.n_lgt + sizeof(qr_clip) * (uint64_t)t.n_clip +
           sizeof(qr_elem) * (uint64_t)t.n_elem + 4 * (uint64_t)t.n_texel +
           sizeof(qr_dev_step) * (uint64_t)(t.n_step + t.n_step_any + 2) +
           (sizeof(qr_refit_elem) + 32) * (uint64_t)t.n_elem;
}

uint64_t footprint(const qrhip_request &r, const qr_device_tables &t, uint64_t step_samples)
{
    const int n_lgt = t.n_lgt;
    const uint64_t ns = r.fsaa == 2 ? 4 : r.fsaa == 1 ? 2 : 1;
    const uint64_t px = (uint64_t)r.x_res * (uint64_t)r.y_res;
    uint64_t b = 4 * px + tables_bytes(t);                  /* device frame, tables */
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
    double                    step_ms = 0.0;       /* device-time budget, 0: none */
    double                    dev_whitted = 0.0, dev_pt = 0.0;
    hipStream_t               stream = nullptr;     /* kernels */
    hipStream_t               copies = nullptr;     /* frames back to the host */
    hipEvent_t                copied = nullptr;
    StepCtx                   ctx[2];
    std::deque<Slot>          slots;     /* stable addresses: steps in flight
                                            point at their slots' tables */
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
    std::string               error_what;          /* the first device error */
    std::thread               th;

    double now_ms() const
    {
        return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    }
    void run();
    hipError_t admit_one(Req *q);
    void pick(std::vector<Unit> &units, double budget_ms);
    void account(StepCtx &c);
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
    const int k = free_slot();
    Slot &sl = slots[k];
    sl.busy = true;
    q->slot = k;
    /* the engine at the request's time (its camera gives the frame's), and
     * the tables for that time in the request's slot; the tables of
     * requests in flight are theirs and stay untouched */
    if (sc.time != q->r.time)
    {
        try
        {
            sc.s->eng->update((rt_time)q->r.time);
        }
        catch (const qr::Error &)
        {
            return hipErrorInvalidValue;
        }
        sc.time = q->r.time;
        sc.s->eng_time = q->r.time;
    }
    if (sl.dt_scene != q->r.scene || sl.dt_time != q->r.time)
    {
        /* the host tables of the engine's time (a scene updated on the
         * device leaves them behind until asked) */
        try
        {
            sc.s->eng->materialize();
        }
        catch (const qr::Error &)
        {
            return hipErrorInvalidValue;
        }
        /* programs built by another scene's engine are never refitted */
        if (sl.dt_scene != q->r.scene) sl.dt.build_on_dev = 0;
        const ProgramInputs in = program_inputs(*sc.s->eng);
        if ((e = sl.dt.upload(sc.s->eng->tables(), stream, &in)) != hipSuccess) return e;
        sl.dt_scene = q->r.scene;
        sl.dt_time = q->r.time;
    }
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
    const uint64_t per_tile = 64ull * ns;
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
    /* service order: Whitted frames of up to 720 rows within 16 ms, larger
     * ones within 48 ms, path-traced ones within 16 ms a pass */
    q->deadline_ms = q->submit_ms + (pt ? 16.0 * q->r.spp : q->r.y_res <= 720 ? 16.0 : 48.0);

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

/* the step's chunks: requests in order of their deadlines, each chunk
 * taken while the step's predicted device time stays within budget_ms (0:
 * no time budget) and its samples within the server's cap; a step takes at
 * least one chunk.  A path-traced chunk takes as many passes as the budget
 * leaves room for, and none while its last run is in flight. */
void qrhip_server::pick(std::vector<Unit> &units, double budget_ms)
{
    units.clear();
    if (active.empty()) return;
    uint64_t used_s = 0;
    double used_ms = 0.0;
    bool any_whitted = false;
    const uint64_t cap = step_samples > 0 ? step_samples : ~0ull;
    std::vector<Req *> order(active.begin(), active.end());
    std::stable_sort(order.begin(), order.end(), [](const Req *x, const Req *y) {
        return x->deadline_ms < y->deadline_ms;
    });
    for (Req *q : order)
    {
        const bool pt = q->r.spp > 0;
        const SceneState &sc = scenes[q->r.scene];
        const double ns_per = pt ? sc.pt_ns : sc.wt_ns;
        for (int c = 0; c < (int)q->chunks.size(); c++)
        {
            Chunk &ch = q->chunks[c];
            if (ch.enqueued == ch.passes) continue;
            /* a path-traced chunk's next pass reads the planes its last pass
             * writes, and a run that overflowed its queues is traced again
             * when its step retires: one run of a chunk in flight at a time */
            if (pt && ch.enqueued != ch.finished) continue;
            const double pass_ms = (double)ch.cost * ns_per * 1e-6;
            const double fixed = pt ? PT_RUN_MS : any_whitted ? 0.0 : WHITTED_STEP_MS;
            if (!units.empty() &&
                ((budget_ms > 0.0 && used_ms + fixed + pass_ms > budget_ms) ||
                 used_s + ch.cost > cap))
            {
                break;      /* the rest of this request waits; others may fit */
            }
            int npass = 1;
            if (pt)
            {
                const int left = ch.passes - ch.enqueued;
                int by_time = left, by_cap = left;
                if (budget_ms > 0.0 && pass_ms > 0.0)
                {
                    by_time = (int)std::max(1.0, (budget_ms - used_ms - fixed) / pass_ms);
                }
                else
                {
                    by_time = 1;
                }
                if (cap != ~0ull)
                {
                    by_cap = (int)std::max<uint64_t>(1, (cap - std::min(cap, used_s)) / ch.cost);
                }
                npass = std::max(1, std::min(left, std::min(by_time, by_cap)));
            }
            const double est = fixed + pass_ms * npass;
            units.push_back(Unit{q, c, (float)(ch.enqueued + 1), npass, est});
            ch.enqueued += npass;
            used_ms += est;
            used_s += ch.cost * (uint64_t)npass;
            any_whitted = any_whitted || !pt;
        }
        if ((budget_ms > 0.0 && used_ms >= budget_ms) || used_s >= cap) break;
    }
}

/* the retired step's measured device time, Whitted batch and path-traced
 * runs apart: each unit is charged its share of its part by prediction,
 * and the rates of the step's scenes are scaled by measured over predicted
 * (bounded, and averaged with their old values) */
void qrhip_server::account(StepCtx &c)
{
    float w_ms = 0.0f, p_ms = 0.0f;
    if (hipEventElapsedTime(&w_ms, c.t0, c.t1) != hipSuccess) w_ms = 0.0f;
    if (hipEventElapsedTime(&p_ms, c.t1, c.t2) != hipSuccess) p_ms = 0.0f;
    for (int part = 0; part < 2; part++)
    {
        const bool pt = part == 1;
        double pred = 0.0, fixed = 0.0, est = 0.0;
        int n = 0;
        for (const Unit &u : c.units)
        {
            if ((u.q->r.spp > 0) != pt) continue;
            n++;
            est += u.est_ms;
            const double ns_per = pt ? scenes[u.q->r.scene].pt_ns : scenes[u.q->r.scene].wt_ns;
            pred += (double)u.q->chunks[u.c].cost * u.npass * ns_per * 1e-6;
            fixed += pt ? PT_RUN_MS : 0.0;
        }
        if (n == 0) continue;
        if (!pt) fixed = WHITTED_STEP_MS;
        const double meas = pt ? p_ms : w_ms;
        {
            std::lock_guard<std::mutex> lk(mu);
            if (pt) dev_pt += meas;
            else dev_whitted += meas;
        }
        for (const Unit &u : c.units)
        {
            if ((u.q->r.spp > 0) != pt) continue;
            u.q->device_ms += est > 0.0 ? meas * (u.est_ms / est) : meas / n;
        }
        if (pred <= 0.0 || meas <= 0.0) continue;
        const double ratio = std::min(4.0, std::max(0.25, std::max(meas - fixed, 0.05 * meas) / pred));
        std::vector<int> seen;
        for (const Unit &u : c.units)
        {
            if ((u.q->r.spp > 0) != pt) continue;
            const int sid = u.q->r.scene;
            if (std::find(seen.begin(), seen.end(), sid) != seen.end()) continue;
            seen.push_back(sid);
            double &rate = pt ? scenes[sid].pt_ns : scenes[sid].wt_ns;
            rate = rate * (0.5 + 0.5 * ratio);
        }
    }
}

/* enqueue a step: its Whitted chunks as one batch, then one pass of each
 * path-traced chunk, then the step's event */
hipError_t qrhip_server::launch(StepCtx &c)
{
    hipError_t e;
    c.items.clear();
    if ((e = hipEventRecord(c.t0, stream)) != hipSuccess) return e;
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
        it.t = &slots[u.q->slot].dt.t;
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
    if ((e = hipEventRecord(c.t1, stream)) != hipSuccess) return e;
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
        e = qr_pt_launch(c.pt_pool[npt++], &sl.dt.t, &d, (float *)sl.planes.ptr,
                         (uint32_t *)sl.seeds.ptr, (uint32_t *)sl.frame.ptr, ch.tile0,
                         ch.ntiles, u.npass, stream);
        if (e != hipSuccess) return e;
    }
    if ((e = hipEventRecord(c.t2, stream)) != hipSuccess) return e;
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
        if (attempt == 0) account(c);
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
        ch.finished += u.npass;
        if (ch.finished == ch.passes) u.q->chunks_left--;
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
            if (used + q->bytes > budget) break;
            waiting.pop_front();
            used += q->bytes;
            peak = std::max(peak, used);
            q->admit_ms = now_ms();
            admitted.push_back(q);
        }
        const double budget_ms = step_ms;
        lk.unlock();

        hipError_t e = hipSuccess;
        const char *stage = "admission";
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
            pick(c.units, budget_ms);
            if (!c.units.empty())
            {
                stage = "step launch";
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
        if (e == hipSuccess && prev != nullptr)
        {
            stage = "step retirement";
            e = retire(*prev);
        }

        std::vector<Req *> fin, keep;
        for (Req *q : active)
        {
            if (e != hipSuccess || q->chunks_left == 0) fin.push_back(q);
            else keep.push_back(q);
        }
        if (e == hipSuccess)
        {
            stage = "frame copy";
            e = finish(fin);
        }
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
        if (e != hipSuccess && error == QRHIP_OK)
        {
            error = QRHIP_E_DEVICE;
            error_what = std::string(stage) + ": " + hipGetErrorString(e);
        }
        for (Req *q : fin)
        {
            qrhip_result res;
            res.tag = q->r.tag;
            res.status = e == hipSuccess ? QRHIP_OK : QRHIP_E_DEVICE;
            res.steps = q->steps;
            res.submit_ms = q->submit_ms;
            res.admit_ms = q->admit_ms;
            res.done_ms = t_end;
            res.device_ms = q->device_ms;
            done.push_back(res);
            used -= q->bytes;
            if (q->slot >= 0) slots[q->slot].busy = false;
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
        st.time = scenes[k]->eng_time;
        s->scenes.push_back(st);
    }
    auto release = [s]()
    {
        for (StepCtx &c : s->ctx)
        {
            qr_wf_batch_destroy(c.wb);
            for (hipEvent_t ev : {c.ev, c.t0, c.t1, c.t2})
            {
                if (ev != nullptr) (void)hipEventDestroy(ev);
            }
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
        if (e == hipSuccess) e = hipEventCreate(&c.t0);
        if (e == hipSuccess) e = hipEventCreate(&c.t1);
        if (e == hipSuccess) e = hipEventCreate(&c.t2);
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
    *bytes = footprint(*req, s->scenes[req->scene].s->dt.t, s->step_samples);
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
    q->bytes = footprint(r, s->scenes[r.scene].s->dt.t, s->step_samples);
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
    st->device_ms_whitted = s->dev_whitted;
    st->device_ms_pt = s->dev_pt;
    if (s->error != QRHIP_OK) return fail(s->error, "server: " + s->error_what);
    return QRHIP_OK;
}

int qrhip_server_set_step_ms(qrhip_server *s, double ms)
{
    if (s == nullptr || !(ms >= 0.0)) return fail(QRHIP_E_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(s->mu);
    s->step_ms = ms;
    return QRHIP_OK;
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
    /* the scenes go back to their callers with their engines at the time of
     * their own last update again (their device tables and camera are that
     * update's; the server moved the engines per request) */
    for (SceneState &sc : s->scenes)
    {
        qrhip_scene *scn = sc.s;
        if (scn->eng_time == scn->time) continue;
        try
        {
            scn->eng->update((rt_time)scn->time);
            scn->eng_time = scn->time;
        }
        catch (const qr::Error &)
        {
            if (s->error == QRHIP_OK)
            {
                s->error = QRHIP_E_SCENE;
                s->error_what = "restoring a scene's update";
            }
        }
    }
    (void)hipStreamSynchronize(s->stream);
    (void)hipStreamSynchronize(s->copies);
    for (StepCtx &c : s->ctx)
    {
        for (qr_wf_state *st : c.pool) qr_wf_destroy(st);
        for (qr_pt_state *st : c.pt_pool) qr_pt_destroy(st);
        qr_wf_batch_destroy(c.wb);
        for (hipEvent_t ev : {c.ev, c.t0, c.t1, c.t2})
        {
            if (ev != nullptr) (void)hipEventDestroy(ev);
        }
    }
    for (Slot &sl : s->slots)
    {
        sl.dt.release();
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
    const std::string what = s->error_what;
    delete s;
    return err == QRHIP_OK ? QRHIP_OK : fail(err, "server: " + what);
}

} /* extern "C" */
