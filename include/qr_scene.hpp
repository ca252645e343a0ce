/*
 * qr_scene.hpp -- QuadRay's rt_Platform / rt_Scene API over libqrhip.so.
 *
 * Header-only C++ layer on the C ABI of qrhip.h, for callers written against
 * QuadRay's engine (core/engine/engine.h:133-154 rt_Platform, 333-364
 * rt_Scene).  Methods keep QuadRay's names, argument meaning and blocking
 * behaviour; failures throw qr::Exception, whose "err" member carries the
 * message as rt_Exception's does (core/engine/engine.cpp:2842-2958).
 *
 * Differences a caller sees: the platform is a HIP device instead of a
 * thread pool (no allocator or thread callbacks); SIMD-target selection has
 * no meaning here; the secondary-ray depth, a compile-time constant in
 * QuadRay (RT_STACK_DEPTH, core/tracer/tracer.h:51), is a runtime setting.
 */
#ifndef QR_SCENE_HPP
#define QR_SCENE_HPP

#include <stdint.h>

#include <string>

#include "qrhip.h"

/* QuadRay's camera actions under their engine names
 * (core/engine/object.h:47-63), for rt_Scene::update(time, action) */
#ifndef RT_CAMERA_CROUCH
#define RT_CAMERA_CROUCH        QRHIP_CAMERA_CROUCH
#define RT_CAMERA_JUMP          QRHIP_CAMERA_JUMP
#define RT_CAMERA_MOVE_DOWN     QRHIP_CAMERA_MOVE_DOWN
#define RT_CAMERA_MOVE_UP       QRHIP_CAMERA_MOVE_UP
#define RT_CAMERA_MOVE_LEFT     QRHIP_CAMERA_MOVE_LEFT
#define RT_CAMERA_MOVE_RIGHT    QRHIP_CAMERA_MOVE_RIGHT
#define RT_CAMERA_MOVE_BACK     QRHIP_CAMERA_MOVE_BACK
#define RT_CAMERA_MOVE_FORWARD  QRHIP_CAMERA_MOVE_FORWARD
#define RT_CAMERA_LEAN_LEFT     QRHIP_CAMERA_LEAN_LEFT
#define RT_CAMERA_LEAN_RIGHT    QRHIP_CAMERA_LEAN_RIGHT
#define RT_CAMERA_ROTATE_DOWN   QRHIP_CAMERA_ROTATE_DOWN
#define RT_CAMERA_ROTATE_UP     QRHIP_CAMERA_ROTATE_UP
#define RT_CAMERA_ROTATE_LEFT   QRHIP_CAMERA_ROTATE_LEFT
#define RT_CAMERA_ROTATE_RIGHT  QRHIP_CAMERA_ROTATE_RIGHT
#endif

namespace qr {

/* thrown by every method that fails; err points at what() */
class Exception
{
  public:
    explicit Exception(const char *msg) : msg_(msg != nullptr ? msg : ""),
                                          err(msg_.c_str()) {}
    Exception(const Exception &o) : msg_(o.msg_), err(msg_.c_str()) {}
    Exception &operator=(const Exception &o)
    {
        msg_ = o.msg_;
        err = msg_.c_str();
        return *this;
    }
    const char *what() const { return err; }

  private:
    std::string msg_;

  public:
    const char *err;
};

inline int check(int rc)
{
    if (rc < 0) throw Exception(qrhip_last_error());
    return rc;
}

/* rt_Platform: the device frames are traced on */
class Platform
{
  public:
    /* ordinal of the HIP device, or QRHIP_HOST_ONLY for a platform that runs
     * the update stage only */
    explicit Platform(int ordinal = 0)
    {
        check(qrhip_device_open(ordinal, &dev_));
    }
    ~Platform()
    {
        if (dev_ != nullptr) qrhip_device_close(dev_);
    }
    Platform(const Platform &) = delete;
    Platform &operator=(const Platform &) = delete;

    /* rt_Platform::set_fsaa / get_fsaa / get_fsaa_max (engine.h:146-149):
     * RT_FSAA_NO 0, RT_FSAA_2X 1, RT_FSAA_4X 2; set_fsaa clamps and returns
     * the mode in effect */
    int set_fsaa(int fsaa) { return check(qrhip_device_set_fsaa(dev_, fsaa)); }
    int get_fsaa() { return check(qrhip_device_get_fsaa(dev_)); }
    int get_fsaa_max() { return check(qrhip_device_get_fsaa_max(dev_)); }

    /* tracing path (QRHIP_PATH_*), no counterpart in QuadRay */
    int set_path(int path) { return check(qrhip_device_set_path(dev_, path)); }
    int get_path() { return check(qrhip_device_get_path(dev_)); }

    qrhip_dev *handle() const { return dev_; }

  private:
    qrhip_dev *dev_ = nullptr;
};

/* rt_Scene: a scene instance bound to a framebuffer and a platform */
class Scene
{
  public:
    /* rt_Scene(rt_SCENE *scn, x_res, y_res, x_row, frame, pfm): frame may be
     * null (the scene allocates one); row y of the frame starts at
     * frame + y * x_row, x_row may be negative */
    Scene(struct rt_SCENE *scn, int x_res, int y_res, int x_row, uint32_t *frame,
          Platform *pfm)
    {
        check(qrhip_scene_create(pfm != nullptr ? pfm->handle() : nullptr, scn,
                                 x_res, y_res, x_row, frame, &scn_));
    }
    ~Scene()
    {
        if (scn_ != nullptr) qrhip_scene_destroy(scn_);
    }
    Scene(const Scene &) = delete;
    Scene &operator=(const Scene &) = delete;

    /* rt_Scene::render(time): update stage and a full frame, complete in
     * the framebuffer on return */
    void render(int64_t time) { check(qrhip_scene_render(scn_, time)); }

    uint32_t *get_frame() { return qrhip_scene_get_frame(scn_); }
    /* may be negative (bottom-up frame) */
    int get_x_row() { return qrhip_scene_get_x_row(scn_); }

    /* option bits (RT_OPTS_*; bit 31 may be set) */
    int get_opts() { return qrhip_scene_get_opts(scn_); }
    int set_opts(int opts) { return qrhip_scene_set_opts(scn_, opts); }
    int get_pton() { return check(qrhip_scene_get_pton(scn_)); }
    int set_pton(int pton) { return check(qrhip_scene_set_pton(scn_, pton)); }
    /* rt_Scene::update(time, action): camera actions (RT_CAMERA_*) */
    void update(int64_t time, int action)
    {
        check(qrhip_scene_update_action(scn_, time, action));
    }
    int next_cam() { return check(qrhip_scene_next_cam(scn_)); }
    int get_cam_idx() { return check(qrhip_scene_get_cam_idx(scn_)); }

    /* rt_Scene::save_frame (engine.cpp:3863-3899): the frame to
     * <data path>dump/scrXXX.bmp, throws "failed to save image" */
    void save_frame(int index) { check(qrhip_scene_save_frame(scn_, index)); }

    /* secondary-ray depth 0..14 (QuadRay: RT_STACK_DEPTH, default 10) */
    int set_depth(int depth) { return check(qrhip_scene_set_depth(scn_, depth)); }
    int get_depth() { return check(qrhip_scene_get_depth(scn_)); }

    qrhip_scene *handle() const { return scn_; }

  private:
    qrhip_scene *scn_ = nullptr;
};

} /* namespace qr */

#endif /* QR_SCENE_HPP */
