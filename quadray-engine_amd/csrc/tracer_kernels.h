/*
 * tracer_kernels.h -- launch interface of the HIP kernels (internal to
 * libqrhip; the public boundary is include/qrhip.h).
 */
#ifndef QR_TRACER_KERNELS_H
#define QR_TRACER_KERNELS_H

#include <hip/hip_runtime.h>

#include "qr_soa.h"
#include "refit.h"

/* a culling step as the kernels read it: the box of a qr_cull_step as
 * bounds grown by the box's share of the allowance, QR_CULL_FUZZ times the
 * L1 distance of its centre from the reference point (world boxes:
 * cull_ref; local boxes: the origin of the frame they are in).  A ray adds its own
 * share, QR_CULL_FUZZ times the L1 distance of its origin from the same
 * point; the sum bounds QR_CULL_FUZZ times the distance from the origin to
 * the centre, so every box is at least the one the allowance defines. */
typedef struct qr_dev_step
{
    float    lo[3], hi[3];
    int32_t  elem;
    int32_t  skip;
} qr_dev_step;

/* A box step has elem -1 and skip = the byte offset of the step after its
 * range (QR_DEV_STEP_BYTES times its index), which the walk adds to the
 * program's address as it is.  An element step carries what the walk needs
 * before it can load the surface record, so one scalar load of the step
 * starts it: skip is the record's byte offset in the surface table with the
 * kind of its local box in the low two bits (records are 4-byte multiples),
 * and an element without a local box keeps its "last" (a transform node's
 * group end, else -1) in lo[0]'s bits.  Local boxes:
 *
 *   QR_LOCAL_GROUP  a surface in a transform node's frame (a member of the
 *                   node's group): the box in that frame, about the node's
 *                   origin, tested against the ray as the node's step left
 *                   it in that frame, before the surface's own set-up;
 *   QR_LOCAL_OWN    a surface that is its own transform node: the box in its
 *                   frame about its origin, tested after its set-up.
 *
 * A device program of n steps is followed by an end step (elem
 * QR_DEV_STEP_END), where every walk stops, so a walk never compares a
 * step's index with the program's length. */
#define QR_DEV_STEP_BYTES 32
#define QR_DEV_STEP_END   0x7fffffff
/* most steps a device program holds (its byte offsets fit an int32) */
#define QR_DEV_STEP_MAX   ((1 << 26) - 1)
#define QR_LOCAL_NONE  0
#define QR_LOCAL_GROUP 1
#define QR_LOCAL_OWN   2
/* most records a surface table of the kernels holds (byte offsets fit an
 * int32) */
#define QR_DEV_SURF_MAX ((int)(0x7fffffff / sizeof(qr_surface)))
static inline QR_HD void qr_dev_box_step(qr_dev_step *d, int skip_index)
{
    d->elem = -1;
    d->skip = QR_DEV_STEP_BYTES * skip_index;
}
static inline QR_HD int qr_dev_box_end(const qr_dev_step *d)
{
    return d->skip / QR_DEV_STEP_BYTES;
}
static inline QR_HD void qr_dev_end_step(qr_dev_step *d)
{
    for (int a = 0; a < 3; a++) d->lo[a] = d->hi[a] = 0.0f;
    d->elem = QR_DEV_STEP_END;
    d->skip = -1;
}
static inline QR_HD void qr_dev_elem_step(qr_dev_step *d, int elem, int surf, int last,
                                          int local)
{
    d->elem = elem;
    d->skip = (int32_t)(surf * (int)sizeof(qr_surface)) | local;
    if (local == QR_LOCAL_NONE)
    {
        for (int a = 0; a < 3; a++) d->lo[a] = d->hi[a] = 0.0f;
        union { int32_t i; float f; } u;
        u.i = last;
        d->lo[0] = u.f;
    }
}
/* the kind of an element step's local box (QR_LOCAL_*) */
static inline QR_HD int qr_dev_step_local(const qr_dev_step *d)
{
    return d->skip & 3;
}
static inline QR_HD int qr_dev_step_surf(const qr_dev_step *d)
{
    return (d->skip & ~3) / (int)sizeof(qr_surface);
}
/* the kind of local box the kernels give a list element whose program step
 * has one (a qr_cull_step with skip -2); groups 0: a program whose group
 * members are not always reached through their node's step (qr_groups_sound)
 * tests no member box */
static inline QR_HD int qr_local_kind(const qr_surface *srf, int surf, int groups)
{
    return srf[surf].trnode == surf ? QR_LOCAL_OWN : groups ? QR_LOCAL_GROUP : QR_LOCAL_NONE;
}
/* the kernels' bounds of an element's local box (centre c, half extents
 * h, about the surface's origin in its tracing frame) */
static inline QR_HD void qr_local_bounds(const qr_surface *srf, int surf, const float *c,
                                         const float *h, float *lo, float *hi)
{
    if (srf[surf].trnode == surf)
    {
        const float zero[3] = {0.0f, 0.0f, 0.0f};
        refit_bounds(c, h, 0, zero, lo, hi);
    }
    else
    {
        refit_bounds_about(c, h, srf[surf].pos, lo, hi);
    }
}

/* device copies of the scene tables (device pointers) */
typedef struct qr_device_tables
{
    qr_surface  *surfaces;
    qr_material *materials;
    qr_light    *lights;
    qr_clip     *clips;
    qr_elem     *elems;
    uint32_t    *texels;
    qr_dev_step *steps;         /* nearest-hit culling program (n_step + 1 steps:
                                   the last is the end step) */
    qr_dev_step *steps_any;     /* any-hit culling program (n_step_any + 1) */
    int32_t      n_surf, n_mat, n_lgt, n_clip, n_elem, n_texel, n_step, n_step_any;
    float        cull_ref[4];   /* reference point of the world boxes */
} qr_device_tables;

/* the device update stage's records (update.hip): every record of the
 * surface table, the planes' texture transforms in mat, and the surface
 * list's element inputs of the culling programs, from the records' inputs
 * (fields.h); clips is the table's clip list; expect holds per element the
 * world and local flags the programs were built for, and an element that
 * differs sets *mismatch */
struct qr_geom;
hipError_t qr_update_records(const struct qr_geom *geoms, int n_rec, const qr_clip *clips,
                             qr_surface *srf, qr_material *mat, struct qr_refit_elem *refit_in,
                             const int32_t *expect, uint32_t *mismatch, hipStream_t stream);

/* the device update stage's refit (update.hip): both programs' boxes (lo/hi
 * of every box step and of every element step with a local box) from the
 * element inputs and the surface records (srf: a member's position in its
 * group's frame), ranges and elements unchanged (n_step and n_step_any do
 * not count the end steps, which stay as they are); scratch holds 8 floats
 * per element; steps_any may be null when n_step_any is 0 */
hipError_t qr_refit_programs(qr_dev_step *steps, int n_step, qr_dev_step *steps_any,
                             int n_step_any, const qr_refit_elem *elems, int n_elem,
                             const qr_surface *srf, const float ref[4], float *scratch,
                             hipStream_t stream);

/* Whitted frame: ARGB8 into frame (row y at frame + y * fd->x_row), and the
 * first-hit id of every sample into ids when ids is not null */
hipError_t qr_launch_whitted(const qr_device_tables *t, const qr_frame_desc *fd,
                             uint32_t *frame, int32_t *ids, hipStream_t stream);

/* the path tracer (tracer_pt.hip), traced as a wavefront: queues of one
 * user (a scene, a server step, a device), reused by every launch */
typedef struct qr_pt_state qr_pt_state;
hipError_t qr_pt_create(qr_pt_state **out);
void       qr_pt_destroy(qr_pt_state *s);
/* "passes" passes over the chunk of 8x8-pixel tiles tile0 .. tile0 +
 * ntiles - 1 (ntiles 0: all), fd->pts_c counting the first: planes (red,
 * green, blue running means, 3 x slots floats) and seeds (one per slot),
 * slot = (y * x_res + x) * samples + k, are updated, and the chunk's pixels
 * resolved into frame (pitch fd->x_res) */
hipError_t qr_pt_launch(qr_pt_state *s, const qr_device_tables *t, const qr_frame_desc *fd,
                        float *planes, uint32_t *seeds, uint32_t *frame, uint32_t tile0,
                        uint32_t ntiles, int passes, hipStream_t stream);
/* after the stream finished a launch: 0 if it is complete, 1 if its last
 * run of passes overflowed a queue (it wrote nothing; the queues are grown:
 * call qr_pt_retry, then check again) */
int        qr_pt_check(qr_pt_state *s);
hipError_t qr_pt_retry(qr_pt_state *s, hipStream_t stream);
uint32_t   qr_pt_tiles(int x_res, int y_res);

/* level-by-level pipeline (tracer_wf.hip): queues and counters of one
 * device, reused by every frame traced with it */
typedef struct qr_wf_state qr_wf_state;

typedef struct qr_wf_stats
{
    uint32_t level_rays[16];    /* rays queued per tree level (level 0: 0) */
} qr_wf_stats;

/* most lights the pipeline handles (visibility mask bits) */
int        qr_wf_max_lights(void);
hipError_t qr_wf_create(qr_wf_state **out);
void       qr_wf_destroy(qr_wf_state *s);
/* enqueue one Whitted frame on "stream" (frame pitch fd->x_res) */
hipError_t qr_wf_launch(qr_wf_state *s, const qr_device_tables *t,
                        const qr_frame_desc *fd, uint32_t *frame, int32_t *ids,
                        hipStream_t stream);
/* after the stream finished the frame: 0 if it is complete, 1 if a queue
 * overflowed (queues are grown; trace the frame again) */
int        qr_wf_check(qr_wf_state *s, qr_wf_stats *stats);

/* Batches: many frames, each with its own queues (state), tables, camera,
 * resolution, anti-aliasing and depth, traced by one launch per stage.  The
 * launches depend only on the number of frames and the deepest depth, so a
 * batch's launch sequence can be captured in a HIP graph once its frame
 * arguments are on the device (they are uploaded when they change). */
#define QR_WF_MAX_BATCH 65536
typedef struct qr_wf_batch qr_wf_batch;
typedef struct qr_wf_item
{
    qr_wf_state              *s;        /* the frame's queues */
    const qr_device_tables   *t;
    const qr_frame_desc      *fd;
    uint32_t                 *frame;    /* device, pitch fd->x_res */
    int32_t                  *ids;      /* device first-hit ids or null */
    uint32_t                  tile0;    /* a chunk of the frame: level-0 tiles */
    uint32_t                  ntiles;   /* tile0 .. tile0 + ntiles - 1; 0: all */
} qr_wf_item;

hipError_t qr_wf_batch_create(qr_wf_batch **out);
/* second streams made so far (one per device at most, shared by batches) */
int        qr_wf_side_streams(void);
void       qr_wf_batch_destroy(qr_wf_batch *b);
/* launch shape: persistent workgroups per compute unit (1..64), level-0
 * tile width in pixels (4, 8, 16 or 32; tiles hold 64 pixels), rays per
 * thread (1..QR_WF_MAX_RPT: blocks of 256 queue entries a virtual workgroup
 * takes) and compaction cadence (1..QR_WF_MAX_CADENCE: children are
 * compacted into the queue of level L only when L is a multiple of it) */
#define QR_WF_MAX_RPT 8
#define QR_WF_MAX_CADENCE 4
hipError_t qr_wf_batch_config(qr_wf_batch *b, int grid_per_cu, int tile_w, int rpt,
                              int cadence);
/* level-0 tiles of an x_res x y_res frame under the batch's tile shape */
uint32_t   qr_wf_batch_tiles(const qr_wf_batch *b, int x_res, int y_res);
hipError_t qr_wf_batch_launch(qr_wf_batch *b, const qr_wf_item *items, int n,
                              hipStream_t stream);
/* after the stream finished the batch: 1 if a frame's queue overflowed (its
 * queues are grown; trace the batch again), else 0 */
int        qr_wf_batch_check(qr_wf_batch *b, const qr_wf_item *items, int n);
void       qr_wf_batch_stats(const qr_wf_batch *b, int k, qr_wf_stats *stats);

#endif /* QR_TRACER_KERNELS_H */
