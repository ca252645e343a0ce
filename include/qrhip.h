/*
 * qrhip.h -- C ABI of the MI355X renderer (libqrhip.so).
 *
 * The drop-in boundary for QuadRay's rendering core.  Each entry point names
 * the QuadRay interface it replaces; all of them return 0 on success and a
 * negative QRHIP_E_* code on failure, with qrhip_last_error() describing the
 * failure of the calling thread's last call (QuadRay throws rt_Exception for
 * the same conditions, core/engine/engine.cpp:2842-2958).  No C++ or torch
 * types cross this boundary.  The C++ shim include/qr_scene.hpp rebuilds
 * QuadRay's rt_Platform / rt_Scene classes on top of it and re-throws.
 *
 * Threading: a device handle may be shared; a scene must not be used from two
 * threads at once (as QuadRay's rt_Scene).  Frames produced by render calls
 * are complete when the call returns.
 */
#ifndef QRHIP_H
#define QRHIP_H

#include <stdint.h>

#include "qr_soa.h"

#ifdef __cplusplus
extern "C" {
#endif

#define QRHIP_OK            0
#define QRHIP_E_ARG        -1   /* invalid argument */
#define QRHIP_E_SCENE      -2   /* scene data rejected (QuadRay rt_Exception) */
#define QRHIP_E_DEVICE     -3   /* HIP runtime error */
#define QRHIP_E_NODEVICE   -4   /* scene was created without a device */
#define QRHIP_E_UNSUPPORTED -5

#define QRHIP_FSAA_NO       0   /* QuadRay RT_FSAA_NO (core/engine/engine.h:52-55) */
#define QRHIP_FSAA_2X       1
#define QRHIP_FSAA_4X       2

typedef struct qrhip_dev   qrhip_dev;
typedef struct qrhip_scene qrhip_scene;
struct rt_SCENE;

/* message of the calling thread's last failed call ("" if none) */
const char *qrhip_last_error(void);

/* ---- platform (replaces rt_Platform, core/engine/engine.h:133-154) ---- */

/* Open HIP device "ordinal" (its own non-blocking stream).  Replaces the
 * rt_Platform constructor's thread-pool and SIMD-target set-up.
 * QRHIP_HOST_ONLY opens a platform without HIP: scenes created on it run the
 * update stage and give their tables, but do not render. */
#define QRHIP_HOST_ONLY    -1
int  qrhip_device_open(int ordinal, qrhip_dev **out);
int  qrhip_device_close(qrhip_dev *dev);
/* Antialiasing mode for scenes created afterwards and rendered on this
 * device: rt_Platform::set_fsaa / get_fsaa / get_fsaa_max (engine.h:146-149) */
int  qrhip_device_set_fsaa(qrhip_dev *dev, int fsaa);
int  qrhip_device_get_fsaa(qrhip_dev *dev);
int  qrhip_device_get_fsaa_max(qrhip_dev *dev);

/* Tracing path for frames rendered on this device (no counterpart in
 * QuadRay, which has one CPU path): QRHIP_PATH_LEVELS, the default, traces
 * the Whitted tree level by level over compacted ray queues;
 * QRHIP_PATH_ONE_KERNEL traces each sample's whole tree in one kernel (also
 * used for scenes with more than 32 lights).  Both give identical frames. */
#define QRHIP_PATH_LEVELS      0
#define QRHIP_PATH_ONE_KERNEL  1
int  qrhip_device_set_path(qrhip_dev *dev, int path);
int  qrhip_device_get_path(qrhip_dev *dev);
/* HIP streams the library holds for overlapped stages beside the devices'
 * own streams (no counterpart in QuadRay): at most one per device, shared by
 * every scene, batch and server on it */
int  qrhip_side_streams(void);

/* ---- scene (replaces rt_Scene, core/engine/engine.h:333-364) ---- */

/* rt_Scene::rt_Scene(rt_SCENE*, x_res, y_res, x_row, frame, pfm):
 * instantiate the object hierarchy from the scene structs and lock them.
 * "frame" is the caller's host framebuffer (row y at frame + y * x_row) or
 * NULL, in which case the scene owns one; dev may be NULL for a host-only
 * scene (update stage and tables, no rendering). */
int  qrhip_scene_create(qrhip_dev *dev, struct rt_SCENE *scn, int x_res,
                        int y_res, int x_row, uint32_t *frame,
                        qrhip_scene **out);
/* rt_Scene::~rt_Scene: unlock and restore the scene structs */
int  qrhip_scene_destroy(qrhip_scene *scn);

/* rt_Scene::render(time): update stage for "time", upload, trace, and
 * copy the frame to the host framebuffer. */
int  qrhip_scene_render(qrhip_scene *scn, int64_t time);

/* split form of render for callers that keep frames in HBM: update stage and
 * table upload (also sets up the device frame), then launches that leave the
 * frame in device memory, then the copy back */
int  qrhip_scene_update(qrhip_scene *scn, int64_t time);
int  qrhip_scene_launch(qrhip_scene *scn);
int  qrhip_scene_sync(qrhip_scene *scn);
int  qrhip_scene_download(qrhip_scene *scn);
/* time n back-to-back launches of the current frame with HIP events on the
 * scene's stream; *ms_per_launch receives the mean kernel time */
int  qrhip_scene_time_launches(qrhip_scene *scn, int n, float *ms_per_launch);

/* rt_Scene::get_frame / get_x_row (engine.h:363, 356) */
uint32_t *qrhip_scene_get_frame(qrhip_scene *scn);
int  qrhip_scene_get_x_row(qrhip_scene *scn);

/* rt_Scene::get_opts / set_opts (engine.h:356-357): option bits minus those
 * the scene data turned off */
int  qrhip_scene_get_opts(qrhip_scene *scn);
int  qrhip_scene_set_opts(qrhip_scene *scn, int opts);
/* rt_Scene::get_pton / set_pton (engine.h:358-359): path-tracer passes per
 * render call, 0 off; switching it on restarts the colour planes and the
 * seed plane, which stay in device memory between calls */
int  qrhip_scene_get_pton(qrhip_scene *scn);
int  qrhip_scene_set_pton(qrhip_scene *scn, int pton);
/* rt_Scene::update(time, action) (core/engine/engine.h:340,
 * engine.cpp:3006-3009; rt_Camera::update_action, object.cpp:495-573): move
 * or turn the current camera by its speeds (rt_CAMERA dps per 50 ms: right,
 * forward, up; drt: turn, tilt) times the time since the camera's last
 * update, along the heading of that update.  Like QuadRay it writes the
 * caller's scene data (the camera's transform, restored when the scene is
 * destroyed) and shows at the next render.  Actions are QuadRay's
 * RT_CAMERA_* values (core/engine/object.h:47-63); crouch, jump and lean
 * move nothing there either. */
#define QRHIP_CAMERA_CROUCH          0
#define QRHIP_CAMERA_JUMP            1
#define QRHIP_CAMERA_MOVE_DOWN       2
#define QRHIP_CAMERA_MOVE_UP         3
#define QRHIP_CAMERA_MOVE_LEFT       4
#define QRHIP_CAMERA_MOVE_RIGHT      5
#define QRHIP_CAMERA_MOVE_BACK       6
#define QRHIP_CAMERA_MOVE_FORWARD    7
#define QRHIP_CAMERA_LEAN_LEFT       8
#define QRHIP_CAMERA_LEAN_RIGHT      9
#define QRHIP_CAMERA_ROTATE_DOWN     10
#define QRHIP_CAMERA_ROTATE_UP       11
#define QRHIP_CAMERA_ROTATE_LEFT     12
#define QRHIP_CAMERA_ROTATE_RIGHT    13
int  qrhip_scene_update_action(qrhip_scene *scn, int64_t time, int action);
/* rt_Scene::next_cam / get_cam_idx (engine.h:361-362) */
int  qrhip_scene_next_cam(qrhip_scene *scn);
int  qrhip_scene_get_cam_idx(qrhip_scene *scn);
/* secondary ray depth: QuadRay's compile-time RT_STACK_DEPTH
 * (core/tracer/tracer.h:51) as a runtime setting, 0..14, default 10 */
int  qrhip_scene_set_depth(qrhip_scene *scn, int depth);
int  qrhip_scene_get_depth(qrhip_scene *scn);

/* ---- lower boundary: the structure-of-arrays tables ---- */

/* tables and frame description produced by the last update stage (host
 * memory owned by the scene; valid until the next update or destroy) */
int  qrhip_scene_tables(qrhip_scene *scn, const qr_scene_tables **tables,
                        qr_frame_desc *fd);
/* render caller-built tables on a device: upload, trace one Whitted frame,
 * copy back "frame" (x_row stride from fd) and, if "ids" is not NULL, the
 * first-hit id of every sample ([y][x][sample], (id << 1 | side) or -1) */
int  qrhip_render_tables(qrhip_dev *dev, const qr_scene_tables *tables,
                         const qr_frame_desc *fd, uint32_t *frame,
                         int32_t *ids);

/* qrhip_render_tables for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, hosts that forbid hipMalloc);
 * no counterpart in QuadRay.  Same arguments and results, traced by the
 * one-kernel path; every device buffer it needs beyond the caller's host
 * inputs and outputs (staged tables, culling programs, device frame and ids)
 * comes from one static device array of QRHIP_NOALLOC_BYTES per GPU, held
 * until the process ends.  It calls no device allocator; tables plus frame
 * (4 bytes a pixel) plus ids (4 bytes a sample) beyond that size are
 * rejected with QRHIP_E_ARG, and so is a row pitch shorter than a row.
 * Calls that overlap on one GPU share the array and race: serialise them. */
#define QRHIP_NOALLOC_BYTES (64u << 20)
int  qrhip_render_tables_noalloc(qrhip_dev *dev, const qr_scene_tables *tables,
                                 const qr_frame_desc *fd, uint32_t *frame,
                                 int32_t *ids);

/* one path-tracer pass on caller-built tables: planes (red, green and blue
 * running means, 3 x slots floats) and seeds (slots values) are read and
 * updated in place, slot = (y * x_res + x) * samples + k; fd->pts_c counts
 * this pass (1 for the first); the frame is resolved from the planes.
 * The pass of rt_Scene::render with set_pton (core/tracer/tracer.cpp PT
 * domain, 2113-2342, 3420-3920, 4686-4752, 6955-6990). */
int  qrhip_render_tables_pt(qrhip_dev *dev, const qr_scene_tables *tables,
                            const qr_frame_desc *fd, float *planes,
                            uint32_t *seeds, uint32_t *frame);
/* a scene's initial seed plane for a frame of x_res x y_res (row pitch
 * x_row) and AA mode fsaa, in the slot order above (rt_Scene::reset_pseed,
 * core/engine/engine.cpp:3719-3740) */
int  qrhip_pt_seed_plane(int x_res, int y_res, int x_row, int fsaa,
                         uint32_t *seeds);

/* ---- files (QuadRay core/engine/rtimag.cpp) ---- */

/* root of the data folders: texture files named by a material (rt_TEX with a
 * file name, RT_TEX_LOAD) load from root + "data/textures/", frame dumps go to
 * root + "dump/"; QuadRay's RT_PATH, "../" by default */
int  qrhip_set_data_path(const char *root);
/* a 24-bit uncompressed BMP as texels 0x00RRGGBB, rows as stored
 * (load_image, rtimag.cpp:47-186); texels may be NULL to ask for the size;
 * fails if the image holds more than max texels */
int  qrhip_image_load(const char *path, uint32_t *texels, int64_t max, int *x_dim,
                      int *y_dim);
/* write a 24-bit BMP (save_image, rtimag.cpp:204-361): x_dim x |y_dim|
 * pixels, y_dim < 0 for top-down rows, row y at pixels + y * stride (stride 0:
 * x_dim; a negative stride walks rows backwards, |stride| >= x_dim) */
int  qrhip_image_save(const char *path, const uint32_t *pixels, int x_dim, int y_dim,
                      int64_t stride);
/* rt_Scene::save_frame (engine.cpp:3863-3899): the scene's frame to
 * root + "dump/scrXXX.bmp" (scrXXX-Y.bmp from index 1000), top-down rows;
 * rows follow the scene's x_row also when it is negative (QuadRay steps a
 * negative x_row as x_res) */
int  qrhip_scene_save_frame(qrhip_scene *scn, int index);

/* ---- batches of frames (no counterpart in QuadRay, which renders one
 *      rt_Scene frame per render() call) ----
 *
 * Many frames of uneven resolution, scene, anti-aliasing and depth traced
 * together: every pipeline stage is one launch over all of them, its
 * workgroups drawn from a global queue of (frame, tile range) entries.  The
 * launch sequence can be captured in a HIP graph and replayed.  Every frame
 * equals the frame qrhip_scene_render gives for the same scene, camera,
 * resolution, AA mode and depth. */
typedef struct qrhip_batch qrhip_batch;

typedef struct qrhip_frame_req
{
    qrhip_scene *scene;     /* tables and camera of the scene's last update */
    int32_t      x_res;     /* frame size, 1..16384 each */
    int32_t      y_res;
    int32_t      fsaa;      /* QRHIP_FSAA_NO, _2X or _4X */
    int32_t      depth;     /* secondary-ray depth 0..14 */
    uint32_t    *frame;     /* host frame (pitch x_res) filled by
                               qrhip_batch_download, or NULL */
} qrhip_frame_req;

#define QRHIP_BATCH_DIRECT   0   /* one launch per stage for the batch */
#define QRHIP_BATCH_GRAPH    1   /* the same launches replayed from a HIP graph */
#define QRHIP_BATCH_SERIAL   2   /* each frame on its own, one after another,
                                    with the single-frame launch shape of a
                                    scene's own launch (the unbatched
                                    reference) */

int  qrhip_batch_create(qrhip_dev *dev, qrhip_batch **out);
int  qrhip_batch_destroy(qrhip_batch *b);
/* the batch's frames (requests are copied; scenes must stay alive and
 * updated); allocates each frame's queues and device frame */
int  qrhip_batch_set(qrhip_batch *b, const qrhip_frame_req *req, int n);
/* launch shape of the batched launches (serial launches keep the
 * single-frame one): persistent workgroups per compute unit (1..64, default 64),
 * level-0 tile width in pixels (4, 8, 16 or 32, default 8; tiles hold 64
 * pixels), rays per thread (1..QRHIP_BATCH_MAX_RPT, default 1: blocks of 256
 * queue entries each virtual workgroup takes) and compaction cadence
 * (1..QRHIP_BATCH_MAX_CADENCE, default 1: the children of level L are
 * compacted into dense queues only when L + 1 is a multiple of it; in
 * between every ray gets two child slots, live or not).  No shape changes a
 * frame. */
#define QRHIP_BATCH_MAX_RPT      8
#define QRHIP_BATCH_MAX_CADENCE  4
int  qrhip_batch_config(qrhip_batch *b, int grid_per_cu, int tile_w, int rays_per_thread,
                        int cadence);
/* enqueue the batch (mode QRHIP_BATCH_*) on the device's stream */
int  qrhip_batch_launch(qrhip_batch *b, int mode);
/* wait; a frame whose queue had to grow is traced again */
int  qrhip_batch_sync(qrhip_batch *b);
/* after sync: copy frame k (or every frame, k < 0) to its request's buffer */
int  qrhip_batch_download(qrhip_batch *b, int k);
/* mean device time of n launches in a mode, in ms per batch */
int  qrhip_batch_time(qrhip_batch *b, int n, int mode, float *ms);
/* autotuner: times the batch under the launch shapes (workgroups per CU x
 * tile width {4, 8, 16, 32}^2 first, then rays per thread {1, 2, 4} x
 * compaction cadence {1, 2} at the best of those), keeps the fastest and
 * returns it in best[0..3] (workgroups per CU, tile width, rays per thread,
 * cadence) with its time in *ms; reps launches per shape */
int  qrhip_batch_tune(qrhip_batch *b, int reps, int *best, float *ms);

/* ---- serving loop (no counterpart in QuadRay) ----
 *
 * A server thread keeps its scenes resident in HBM and traces frame requests
 * as they arrive.  Requests wait in arrival order and are admitted while
 * their memory footprint fits the budget beside the requests in flight.  An
 * admitted frame is split into chunks (runs of level-0 tiles); each step
 * takes, within a budget of camera samples, a chunk of the oldest request in
 * flight and then chunks of the requests with the least work left, traces
 * the Whitted chunks as one batch (one launch per pipeline stage) and one
 * pass of each path-traced chunk, and returns the frames that finished.
 * Steps are pipelined (the next is launched before the last is retired).
 * Requests that arrive during a step join the next one, so a small frame
 * does not wait for a large one to finish.  A frame is the same whatever
 * else shares its steps. */
typedef struct qrhip_server qrhip_server;

typedef struct qrhip_request
{
    uint64_t  tag;          /* returned with the result */
    int32_t   scene;        /* index into the server's scenes */
    int32_t   x_res, y_res; /* 1..16384 each */
    int32_t   fsaa;         /* QRHIP_FSAA_* */
    int32_t   depth;        /* Whitted secondary-ray depth 0..14 */
    int32_t   spp;          /* 0: Whitted frame; n > 0: path tracer, n passes */
    int32_t   use_pose;     /* 1: camera from pose[], else the scene's camera */
    int64_t   time;         /* animation time: the scene's update stage runs
                               for it before the frame, into tables of the
                               request's own */
    float     pose[12];     /* right, down and view axes, then position */
    uint32_t *frame;        /* host frame x_res * y_res (pitch x_res); stays
                               the caller's, written before the result */
} qrhip_request;

typedef struct qrhip_result
{
    uint64_t  tag;
    int32_t   status;       /* QRHIP_OK or a QRHIP_E_* code */
    int32_t   steps;        /* steps the request took part in */
    double    submit_ms;    /* times since the server started */
    double    admit_ms;
    double    done_ms;
    double    device_ms;    /* device time of its steps charged to it: each
                               step's measured Whitted batch and path-traced
                               runs shared by predicted cost */
} qrhip_result;

typedef struct qrhip_server_stats
{
    uint64_t  steps;            /* steps run */
    uint64_t  frames;           /* requests completed */
    uint64_t  step_frames_max;  /* most requests served by one step */
    uint64_t  bytes_in_use;     /* footprint of requests in flight now */
    uint64_t  bytes_peak;       /* largest footprint in flight */
    double    busy_ms;          /* time spent in steps */
    double    device_ms_whitted;    /* measured device time of Whitted batches */
    double    device_ms_pt;         /* and of path-traced runs */
} qrhip_server_stats;

/* scenes must be updated, on dev, and left to the server until it is
 * destroyed; hbm_budget bounds the summed footprint of requests in flight;
 * max_active bounds the requests in flight; step_samples is the camera
 * samples (path-traced samples count twice) a step takes, 0 for whole frames
 * (max_active 1 and step_samples 0: one frame at a time, unbatched) */
int  qrhip_server_create(qrhip_dev *dev, qrhip_scene *const *scenes, int n_scenes,
                         uint64_t hbm_budget, int max_active, uint64_t step_samples,
                         qrhip_server **out);
/* queue a request (thread-safe); fails if it cannot ever fit the budget */
int  qrhip_server_submit(qrhip_server *s, const qrhip_request *req);
/* completed requests, up to max, waiting up to wait_ms for the first one;
 * returns how many were written */
int  qrhip_server_poll(qrhip_server *s, qrhip_result *out, int max, int wait_ms);
/* device memory a request needs while in flight, in bytes */
int  qrhip_server_footprint(qrhip_server *s, const qrhip_request *req, uint64_t *bytes);
int  qrhip_server_get_stats(qrhip_server *s, qrhip_server_stats *st);
/* a step's budget of predicted device time in ms (0, the default: only the
 * step_samples cap); requests are served in order of their deadlines (16 ms
 * after submission for Whitted frames of up to 720 rows, 48 ms for larger
 * ones, 16 ms a pass for path-traced ones), and the prediction is each
 * scene's device time per sample as measured in the steps so far */
int  qrhip_server_set_step_ms(qrhip_server *s, double ms);
/* finishes the requests already submitted, then stops */
int  qrhip_server_destroy(qrhip_server *s);

/* the scene's culling programs (no counterpart in QuadRay, whose tile
 * lists are rebuilt every frame): how many times the engine built them
 * (an update builds them when the surface list or a surface's boxedness
 * changed, and otherwise only refits their boxes), how many updates the
 * device refitted them itself (the update stage's device side), and, for
 * checks, in how many steps, records and materials the device's tables
 * differ from the host's (0 expected; the host derives its copy on demand
 * after device updates; mismatch may be NULL, which skips the comparison) */
int  qrhip_scene_program_stats(qrhip_scene *s, int *builds, int *refits, int *mismatch);

/* the camera basis of a scene's current camera: right, down and view axes
 * (QuadRay's camera matrix rows 0-2) and position, for qrhip_request.pose */
int  qrhip_scene_get_pose(qrhip_scene *s, float *pose);

/* ---- synthetic workloads (no counterpart in QuadRay) ---- */

/* seeded rt_SCENE trees: config 1 plain 800x480, 2 Whitted 1080p 64
 * surfaces, 3 path-traced box, 4 and 5 Whitted with 256 / 384 surfaces,
 * 6 the 64-surface Whitted scene with glass everywhere, 7 and 8 Whitted
 * scenes of 8 to 512 / 16 to 1024 quadrics (log-uniform, drawn from the
 * seed).  Flags: QRGEN_IDS replaces every material by an id-encoding LIGHT
 * material; QRGEN_TEXTURE_FILES writes every texel array as a BMP under the
 * data path's data/textures/ (qrgen_<config>_<seed>_<k>.bmp) and names that
 * file in its material instead; QRGEN_ANIMATED gives the root's child arrays
 * animators (turning about z, or moving up and down, as functions of the
 * time).  Resolution, AA, depth and samples the
 * workload is specified with are returned in info[0..4] (x_res, y_res,
 * fsaa, depth, spp). */
#define QRGEN_IDS           1
#define QRGEN_TEXTURE_FILES 2
#define QRGEN_ANIMATED      4
struct rt_SCENE *qrgen_create(int config, uint64_t seed, int flags, int *info);
void qrgen_destroy(struct rt_SCENE *scn);

#ifdef __cplusplus
}
#endif

#endif /* QRHIP_H */
