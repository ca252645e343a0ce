/*
 * capi_internal.h -- state behind the C ABI's opaque handles (qrhip_dev,
 * qrhip_scene), shared by the files that implement the ABI.  Not installed.
 */
#ifndef QR_CAPI_INTERNAL_H
#define QR_CAPI_INTERNAL_H

#include <algorithm>
#include <cmath>
#include <cstring>
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

/* Pinned staging of host-to-device copies: the bytes of one upload are
 * gathered into a page-locked buffer and copied from there, so the copies
 * are truly asynchronous; the buffer is reused only after the event
 * recorded behind the previous upload's copies. */
struct Stager
{
    uint8_t   *host = nullptr;
    size_t     cap = 0, used = 0;
    hipEvent_t ev = nullptr;

    hipError_t begin(size_t total)
    {
        hipError_t e;
        if (ev == nullptr &&
            (e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess)
        {
            return e;
        }
        if ((e = hipEventSynchronize(ev)) != hipSuccess) return e;
        used = 0;
        if (total <= cap) return hipSuccess;
        if (host != nullptr) (void)hipHostFree(host);
        host = nullptr;
        cap = 0;
        if ((e = hipHostMalloc((void **)&host, total, hipHostMallocDefault)) != hipSuccess)
        {
            return e;
        }
        cap = total;
        return hipSuccess;
    }
    /* copy n bytes of src to dst (device) through the buffer */
    hipError_t copy(void *dst, const void *src, size_t n, hipStream_t st)
    {
        uint8_t *p = host + used;
        memcpy(p, src, n);
        used += (n + 255) & ~(size_t)255;
        return hipMemcpyAsync(dst, p, n, hipMemcpyHostToDevice, st);
    }
    hipError_t end(hipStream_t st) { return hipEventRecord(ev, st); }
    void release()
    {
        if (ev != nullptr) (void)hipEventSynchronize(ev);
        if (host != nullptr) (void)hipHostFree(host);
        if (ev != nullptr) (void)hipEventDestroy(ev);
        host = nullptr;
        cap = used = 0;
        ev = nullptr;
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

/* whether every element step of a program that has a member's local box
 * (a surface in a transform node's frame) is reached only through its
 * group's transform-node step, the nearest such step before it: the walk
 * tests a member's box against the ray that step put in the node's frame.
 * Nodes and members follow each other in list order and the programs keep
 * a group whole, so engine-built programs always are; a box step whose
 * range would enter a group after its node's step is not. */
inline bool groups_sound(const qr_cull_step *s, int n, const qr_elem *elems,
                         const qr_surface *srf)
{
    std::vector<int> node_at((size_t)n + 1, -1);  /* nearest node step at or before */
    std::vector<int> last_member((size_t)n, -1);  /* per node step: its last member step */
    int node = -1;
    for (int q = 0; q < n; q++)
    {
        if (s[q].elem >= 0)
        {
            const int surf = elems[s[q].elem].surf;
            const qr_surface &r = srf[surf];
            if (r.tag < 0 && r.trm != 0)
            {
                node = q;
            }
            else if (s[q].skip == -2 && r.trnode != surf)
            {
                /* a member: its node is the nearest node step, and the
                 * member lies in that node's group */
                if (node < 0) return false;
                const int ne = s[node].elem;
                if (r.trnode != elems[ne].surf || s[q].elem <= ne ||
                    s[q].elem > elems[ne].last)
                {
                    return false;
                }
                last_member[(size_t)node] = q;
            }
        }
        node_at[(size_t)q] = node;
    }
    for (int b = 0; b < n; b++)
    {
        if (s[b].elem >= 0) continue;
        const int k = s[b].skip;
        if (k >= n) continue;
        const int t = node_at[(size_t)k];
        if (t >= 0 && b < t && t < k && k <= last_member[(size_t)t]) return false;
    }
    return true;
}

/* the kernels' form of a program: bounds with the box's share of the
 * allowance, rounded outwards (qr_dev_step, refit_bounds), and the end step */
inline void dev_steps(const qr_cull_step *s, int n, const float ref[4], const qr_elem *elems,
                      const qr_surface *srf, std::vector<qr_dev_step> &out)
{
    out.assign((size_t)n + 1, qr_dev_step{});
    qr_dev_end_step(&out[(size_t)n]);
    const int groups = groups_sound(s, n, elems, srf) ? 1 : 0;
    for (int p = 0; p < n; p++)
    {
        qr_dev_step &d = out[(size_t)p];
        if (s[p].elem < 0)
        {
            qr_dev_box_step(&d, s[p].skip);
            refit_bounds(s[p].c, s[p].h, 1, ref, d.lo, d.hi);
            continue;
        }
        const qr_elem &el = elems[s[p].elem];
        const int kind = s[p].skip == -2 ? qr_local_kind(srf, el.surf, groups) : QR_LOCAL_NONE;
        qr_dev_elem_step(&d, s[p].elem, el.surf, el.last, kind);
        if (kind != QR_LOCAL_NONE) qr_local_bounds(srf, el.surf, s[p].c, s[p].h, d.lo, d.hi);
    }
}


/* what the engine knows about the programs it hands over: the element
 * inputs of their boxes, the build they come from, whether its update only
 * refitted that build's boxes, and the allowance's reference point */
struct ProgramInputs
{
    const qr_refit_elem *elems = nullptr;
    int                  n_elem = 0;
    uint64_t             build = 0;
    bool                 refit = false;
    const float         *ref = nullptr;
};

/* the engine's side of an upload */
inline ProgramInputs program_inputs(const qr::Engine &eng)
{
    ProgramInputs in;
    in.elems = eng.refit_inputs().data();
    in.n_elem = (int)eng.refit_inputs().size();
    in.build = eng.program_build();
    in.refit = eng.refitted();
    in.ref = eng.cull_ref();
    return in;
}

struct DevTables
{
    DevBuf srf, mat, lgt, clp, elem, tex, step, step_any, refit_in, refit_box;
    DevBuf geom, expect, flag;      /* the device update stage's inputs */
    uint32_t *flag_host = nullptr;  /* pinned copy of flag */
    uint64_t expect_build = 0;      /* the build "expect" holds the flags of */
    Stager stage;
    std::vector<uint8_t> sent[8];   /* what each table's buffer holds now */
    qr_device_tables t;
    std::vector<qr_dev_step> dstep, dstep_any;  /* the programs as uploaded */
    uint64_t build_on_dev = 0;  /* the engine build whose programs the device holds */
    int      dev_n_step = 0, dev_n_any = 0;
    int      refits = 0;        /* programs refitted on the device */

    /* the tables to the device.  When "in" says the engine only refitted the
     * boxes of the programs the device already holds, the device refits
     * them itself from the element inputs (qr_refit_programs) instead of
     * receiving them. */
    hipError_t upload(const qr_scene_tables &h, hipStream_t st,
                      const ProgramInputs *in = nullptr)
    {
        hipError_t e;
        int n_step = h.n_step;
        float ref[4];
        if (in != nullptr && in->ref != nullptr)
        {
            for (int a = 0; a < 4; a++) ref[a] = in->ref[a];
        }
        else
        {
            cull_reference(h.steps, h.n_step, ref);
        }
        const bool any = h.n_step_any > 0 && h.n_step > 0;
        const bool on_dev = in != nullptr && in->refit && in->build != 0 &&
                            in->build == build_on_dev && h.n_step > 0 &&
                            h.n_step == dev_n_step && (any ? h.n_step_any : 0) == dev_n_any &&
                            in->n_elem == h.n_elem;
        if (on_dev)
        {
            const size_t bytes = sizeof(qr_refit_elem) * (size_t)in->n_elem;
            if ((e = refit_in.reserve(bytes)) != hipSuccess) return e;
            if ((e = refit_box.reserve(sizeof(float) * 8 * (size_t)in->n_elem)) != hipSuccess)
            {
                return e;
            }
            if ((e = upload_rest(h, st, n_step, any, ref, false, in)) != hipSuccess) return e;
            if ((e = qr_refit_programs((qr_dev_step *)step.ptr, h.n_step,
                                       any ? (qr_dev_step *)step_any.ptr : nullptr,
                                       any ? h.n_step_any : 0,
                                       (const qr_refit_elem *)refit_in.ptr, in->n_elem,
                                       (const qr_surface *)srf.ptr, ref,
                                       (float *)refit_box.ptr, st)) != hipSuccess)
            {
                return e;
            }
            /* the device now holds boxes no host copy has */
            sent[0].clear();
            sent[1].clear();
            refits++;
            return hipSuccess;
        }
        if (n_step == 0)
        {
            /* tables without a culling program walk the whole list */
            dstep.assign((size_t)h.n_elem + 1, qr_dev_step{});
            for (int i = 0; i < h.n_elem; i++)
            {
                qr_dev_elem_step(&dstep[(size_t)i], i, h.elems[i].surf, h.elems[i].last, 0);
            }
            qr_dev_end_step(&dstep[(size_t)h.n_elem]);
            n_step = h.n_elem;
        }
        else
        {
            dev_steps(h.steps, h.n_step, ref, h.elems, h.surfaces, dstep);
        }
        /* any-hit walks use the nearest-hit program when they have none */
        if (any) dev_steps(h.steps_any, h.n_step_any, ref, h.elems, h.surfaces, dstep_any);
        build_on_dev = in != nullptr ? in->build : 0;
        dev_n_step = n_step;
        dev_n_any = any ? h.n_step_any : 0;
        return upload_rest(h, st, n_step, any, ref, true, nullptr);
    }

    /* the tables, and the programs when "programs" (else they are in place);
     * the element inputs of a device refit when "refit" is given.  A table
     * whose bytes equal what its buffer holds is not sent again. */
    hipError_t upload_rest(const qr_scene_tables &h, hipStream_t st, int n_step, bool any,
                           const float ref[4], bool programs, const ProgramInputs *refit)
    {
        hipError_t e;
        struct Item { DevBuf *b; const void *src; size_t n; };
        Item items[8] = {
            {&step, dstep.data(), programs ? sizeof(qr_dev_step) * ((size_t)n_step + 1) : 0},
            {&step_any, any ? dstep_any.data() : nullptr,
             programs && any ? sizeof(qr_dev_step) * ((size_t)h.n_step_any + 1) : 0},
            {&srf, h.surfaces, sizeof(qr_surface) * (size_t)h.n_surf},
            {&mat, h.materials, sizeof(qr_material) * (size_t)h.n_mat},
            {&lgt, h.lights, sizeof(qr_light) * (size_t)h.n_lgt},
            {&clp, h.clips, sizeof(qr_clip) * (size_t)h.n_clip},
            {&elem, h.elems, sizeof(qr_elem) * (size_t)h.n_elem},
            {&tex, h.texels, sizeof(uint32_t) * (size_t)h.n_texel},
        };
        bool send[8];
        size_t total = 0;
        for (int i = 0; i < 8; i++)
        {
            Item &it = items[i];
            void *was = it.b->ptr;
            if ((e = it.b->reserve(it.n)) != hipSuccess) return e;
            send[i] = it.n > 0 &&
                      (it.b->ptr != was || sent[i].size() != it.n ||
                       memcmp(sent[i].data(), it.src, it.n) != 0);
            if (send[i]) total += (it.n + 255) & ~(size_t)255;
        }
        const size_t rbytes = refit != nullptr ? sizeof(qr_refit_elem) * (size_t)refit->n_elem
                                               : 0;
        total += (rbytes + 255) & ~(size_t)255;
        if (total > 0)
        {
            if ((e = stage.begin(total)) != hipSuccess) return e;
            for (int i = 0; i < 8; i++)
            {
                if (!send[i]) continue;
                const uint8_t *src = (const uint8_t *)items[i].src;
                sent[i].assign(src, src + items[i].n);
                if ((e = stage.copy(items[i].b->ptr, src, items[i].n, st)) != hipSuccess)
                {
                    sent[i].clear();
                    return e;
                }
            }
            if (rbytes > 0 && (e = stage.copy(refit_in.ptr, refit->elems, rbytes, st)) != hipSuccess)
            {
                return e;
            }
            if ((e = stage.end(st)) != hipSuccess) return e;
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
        step_any.release(); refit_in.release(); refit_box.release();
        geom.release(); expect.release(); flag.release();
        if (flag_host != nullptr) (void)hipHostFree(flag_host);
        flag_host = nullptr;
        stage.release();
        for (std::vector<uint8_t> &v : sent) v.clear();
        build_on_dev = 0;
        expect_build = 0;
    }

    /* whether this holds the tables and programs of eng's current build
     * (the device update stage then derives the records itself) */
    bool holds_build(const qr::Engine &eng) const
    {
        const qr_scene_tables &h = eng.tables();
        return build_on_dev != 0 && build_on_dev == eng.program_build() && dev_n_step > 0 &&
               h.n_step == dev_n_step && t.n_surf == h.n_surf && t.n_elem == h.n_elem &&
               t.n_lgt == h.n_lgt && (int)eng.geoms().size() == h.n_surf &&
               eng.program_key().size() == 4 * (size_t)h.n_elem + 1;
    }

    /* The device update stage (update.hip): every record, the planes'
     * texture transforms and the element inputs from the engine's record
     * inputs (after update_transforms), then the programs refitted, and the
     * lights from the host.  The device checks every element's box flags
     * against the programs' build; *ok is false if one differs (the caller
     * then rebuilds on the host).  Waits for the device. */
    hipError_t update_device(const qr::Engine &eng, hipStream_t st, bool *ok)
    {
        hipError_t e;
        *ok = false;
        const qr_scene_tables &h = eng.tables();
        const std::vector<qr_geom> &gs = eng.geoms();
        const int n_elem = h.n_elem;
        const bool any = dev_n_any > 0;
        const size_t gbytes = sizeof(qr_geom) * gs.size();
        const size_t ebytes = sizeof(int32_t) * 2 * (size_t)n_elem;
        if ((e = geom.reserve(gbytes)) != hipSuccess) return e;
        if ((e = expect.reserve(ebytes)) != hipSuccess) return e;
        if ((e = flag.reserve(sizeof(uint32_t))) != hipSuccess) return e;
        if ((e = refit_in.reserve(sizeof(qr_refit_elem) * (size_t)n_elem)) != hipSuccess) return e;
        if ((e = refit_box.reserve(sizeof(float) * 8 * (size_t)n_elem)) != hipSuccess) return e;
        if (flag_host == nullptr &&
            (e = hipHostMalloc((void **)&flag_host, sizeof(uint32_t), hipHostMallocDefault)) !=
            hipSuccess)
        {
            return e;
        }
        std::vector<int32_t> want;
        const bool send_expect = expect_build != build_on_dev;
        if (send_expect)
        {
            const std::vector<int32_t> &key = eng.program_key();
            want.resize(2 * (size_t)n_elem);
            for (int k = 0; k < n_elem; k++)
            {
                want[2 * (size_t)k] = key[4 * (size_t)k + 2];
                want[2 * (size_t)k + 1] = key[4 * (size_t)k + 3];
            }
        }
        const size_t lbytes = sizeof(qr_light) * (size_t)h.n_lgt;
        const bool send_lgt = lbytes > 0 && (sent[4].size() != lbytes ||
                                             memcmp(sent[4].data(), h.lights, lbytes) != 0);
        const uint32_t zero = 0u;
        size_t total = ((gbytes + 255) & ~(size_t)255) + 256;
        if (send_expect) total += (ebytes + 255) & ~(size_t)255;
        if (send_lgt) total += (lbytes + 255) & ~(size_t)255;
        if ((e = stage.begin(total)) != hipSuccess) return e;
        if ((e = stage.copy(geom.ptr, gs.data(), gbytes, st)) != hipSuccess) return e;
        if ((e = stage.copy(flag.ptr, &zero, sizeof(zero), st)) != hipSuccess) return e;
        if (send_expect && (e = stage.copy(expect.ptr, want.data(), ebytes, st)) != hipSuccess)
        {
            return e;
        }
        if (send_lgt)
        {
            const uint8_t *src = (const uint8_t *)h.lights;
            sent[4].assign(src, src + lbytes);
            if ((e = stage.copy(lgt.ptr, src, lbytes, st)) != hipSuccess) return e;
        }
        if ((e = stage.end(st)) != hipSuccess) return e;
        if (send_expect) expect_build = build_on_dev;
        e = qr_update_records((const qr_geom *)geom.ptr, h.n_surf, (const qr_clip *)clp.ptr,
                              (qr_surface *)srf.ptr, (qr_material *)mat.ptr,
                              (qr_refit_elem *)refit_in.ptr, (const int32_t *)expect.ptr,
                              (uint32_t *)flag.ptr, st);
        if (e != hipSuccess) return e;
        e = qr_refit_programs((qr_dev_step *)step.ptr, dev_n_step,
                              any ? (qr_dev_step *)step_any.ptr : nullptr, any ? dev_n_any : 0,
                              (const qr_refit_elem *)refit_in.ptr, n_elem,
                              (const qr_surface *)srf.ptr, t.cull_ref,
                              (float *)refit_box.ptr, st);
        if (e != hipSuccess) return e;
        if ((e = hipMemcpyAsync(flag_host, flag.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                st)) != hipSuccess)
        {
            return e;
        }
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
        /* the device now holds records, materials and boxes no host copy has */
        for (int i : {0, 1, 2, 3}) sent[i].clear();
        refits++;
        *ok = *flag_host == 0u;
        return hipSuccess;
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
    qr_pt_state *pt = nullptr;        /* queues for qrhip_render_tables_pt */
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
    int64_t      eng_time = 0;         /* time the engine's state is for (a
                                          server moves it per request) */
    qr_frame_desc fd;
    qrhip_impl::DevTables dt;
    qrhip_impl::DevBuf dframe;         /* device frame, pitch x_res */
    qr_wf_state *wf = nullptr;         /* level pipeline queues */
    bool         pending = false;      /* launched frames not yet checked */
    /* path tracer: running-mean planes and seed plane, resident in HBM;
     * its bounce queues, and whether launched passes are still unchecked */
    qrhip_impl::DevBuf pt_planes, pt_seeds;
    qr_pt_state *pt = nullptr;
    bool         pt_pending = false;
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

/* every table index the kernels follow stays inside its table, and the
 * frame description is in range (QRHIP_OK, or the failure reported) */
int validate_tables(const qr_scene_tables *tab, const qr_frame_desc *fd);

/* the device address of the no-alloc scratch array on device "ordinal",
 * looked up once (noalloc.hip) */
hipError_t noalloc_setup(int ordinal);

/* enqueue one Whitted frame on the device's chosen path */
hipError_t launch_frame(const qrhip_dev *d, qr_wf_state *wf, const qr_device_tables *t,
                        const qr_frame_desc *fd, uint32_t *frame, int32_t *ids);

} /* namespace qrhip_impl */

#endif /* QR_CAPI_INTERNAL_H */
