"""ctypes binding of libqrhip.so (include/qrhip.h) for tests and bench.py.

The shared library is the product; this module only marshals arguments.
Structures mirror include/qr_soa.h field for field.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libqrhip.so")


class QrSurface(C.Structure):
    _fields_ = [
        ("pos", C.c_float * 3), ("cmin", C.c_float * 3), ("cmax", C.c_float * 3),
        ("tc", (C.c_float * 3) * 3), ("sci", C.c_float * 4), ("scj", C.c_float * 3),
        ("d_eps", C.c_float), ("t_eps", C.c_float),
        ("amap", C.c_int32 * 3), ("asgn", C.c_int32 * 3), ("shift", C.c_int32),
        ("trm", C.c_int32), ("tag", C.c_int32), ("solver", C.c_int32),
        ("nrm_kind", C.c_int32), ("clp_kind", C.c_int32), ("apex", C.c_int32),
        ("trnode", C.c_int32), ("mat", C.c_int32 * 2), ("props", C.c_uint32 * 2),
        ("clip_beg", C.c_int32), ("clip_cnt", C.c_int32), ("cmin_on", C.c_int32),
        ("cmax_on", C.c_int32), ("id", C.c_int32), ("pad", C.c_int32 * 3),
    ]


class QrClip(C.Structure):
    _fields_ = [("kind", C.c_int32), ("surf", C.c_int32), ("rel", C.c_int32),
                ("pad", C.c_int32)]


class QrMaterial(C.Structure):
    _fields_ = [
        ("xscal", C.c_float), ("yscal", C.c_float), ("xoffs", C.c_float),
        ("yoffs", C.c_float), ("xmask", C.c_uint32), ("ymask", C.c_uint32),
        ("yshft", C.c_uint32), ("tex_off", C.c_int32), ("tmap", C.c_int32 * 2),
        ("l_dff", C.c_float), ("l_spc", C.c_float), ("l_pow", C.c_uint32),
        ("c_rfl", C.c_float), ("c_trn", C.c_float), ("c_rfr", C.c_float),
        ("rfr_2", C.c_float), ("c_rcp", C.c_float), ("ext_2", C.c_float),
        ("col", C.c_float * 3), ("e_src", C.c_float), ("pad", C.c_int32 * 2),
    ]


class QrLight(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("col", C.c_float * 3), ("l_src", C.c_float),
                ("a_qdr", C.c_float), ("a_lnr", C.c_float), ("a_cnt", C.c_float),
                ("a_rng", C.c_float), ("t_max", C.c_float)]


class QrElem(C.Structure):
    _fields_ = [("surf", C.c_int32), ("last", C.c_int32)]


class QrCullStep(C.Structure):
    _fields_ = [("c", C.c_float * 3), ("h", C.c_float * 3),
                ("elem", C.c_int32), ("skip", C.c_int32)]


class QrSceneTables(C.Structure):
    _fields_ = [
        ("n_surf", C.c_int32), ("n_mat", C.c_int32), ("n_lgt", C.c_int32),
        ("n_clip", C.c_int32), ("n_elem", C.c_int32), ("n_texel", C.c_int32),
        ("n_user", C.c_int32), ("pad", C.c_int32),
        ("surfaces", C.POINTER(QrSurface)), ("materials", C.c_void_p),
        ("lights", C.c_void_p), ("clips", C.c_void_p), ("elems", C.c_void_p),
        ("texels", C.c_void_p), ("n_step", C.c_int32), ("pad2", C.c_int32),
        ("steps", C.POINTER(QrCullStep)), ("n_step_any", C.c_int32),
        ("pad3", C.c_int32), ("steps_any", C.POINTER(QrCullStep)),
    ]


class QrFrameDesc(C.Structure):
    _fields_ = [
        ("org", C.c_float * 3), ("dir", C.c_float * 3), ("hor", C.c_float * 3),
        ("ver", C.c_float * 3), ("amb", C.c_float * 4), ("t_min", C.c_float),
        ("x_res", C.c_int32), ("y_res", C.c_int32), ("x_row", C.c_int32),
        ("fsaa", C.c_int32), ("depth", C.c_int32), ("gamma", C.c_uint32),
        ("pt_on", C.c_int32), ("pts_c", C.c_float), ("pt_split", C.c_int32),
        ("pt_roul", C.c_int32), ("pad", C.c_int32),
        ("aa_h", C.c_float * 4), ("aa_v", C.c_float * 4),
    ]


class QrFrameReq(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("x_res", C.c_int32), ("y_res", C.c_int32),
                ("fsaa", C.c_int32), ("depth", C.c_int32),
                ("frame", C.POINTER(C.c_uint32))]


class QrRequest(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("scene", C.c_int32), ("x_res", C.c_int32),
                ("y_res", C.c_int32), ("fsaa", C.c_int32), ("depth", C.c_int32),
                ("spp", C.c_int32), ("use_pose", C.c_int32), ("time", C.c_int64),
                ("pose", C.c_float * 12), ("frame", C.POINTER(C.c_uint32))]


class QrResult(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("status", C.c_int32), ("steps", C.c_int32),
                ("submit_ms", C.c_double), ("admit_ms", C.c_double),
                ("done_ms", C.c_double), ("device_ms", C.c_double)]


class QrServerStats(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("frames", C.c_uint64),
                ("step_frames_max", C.c_uint64), ("bytes_in_use", C.c_uint64),
                ("bytes_peak", C.c_uint64), ("busy_ms", C.c_double),
                ("device_ms_whitted", C.c_double), ("device_ms_pt", C.c_double)]


# the arrays of a qr_scene_tables: (pointer field, count field, record type)
TABLE_ARRAYS = (
    ("surfaces", "n_surf", QrSurface), ("materials", "n_mat", QrMaterial),
    ("lights", "n_lgt", QrLight), ("clips", "n_clip", QrClip),
    ("elems", "n_elem", QrElem), ("texels", "n_texel", C.c_uint32),
    ("steps", "n_step", QrCullStep), ("steps_any", "n_step_any", QrCullStep),
)
TABLE_COUNTS = ("n_surf", "n_mat", "n_lgt", "n_clip", "n_elem", "n_texel", "n_user",
                "n_step", "n_step_any")


def tables_snapshot(tables_ptr, fd):
    """Copy a scene's tables and frame description into plain numpy arrays
    (records as raw bytes), e.g. to store them as a file."""
    t = tables_ptr.contents
    out = {"counts": np.array([getattr(t, k) for k in TABLE_COUNTS], dtype=np.int64),
           "fd": np.frombuffer(C.string_at(C.addressof(fd), C.sizeof(fd)),
                               dtype=np.uint8).copy()}
    for field, count, rec in TABLE_ARRAYS:
        n = getattr(t, count) * C.sizeof(rec)
        ptr = C.cast(getattr(t, field), C.c_void_p).value
        out[field] = (np.frombuffer(C.string_at(ptr, n), dtype=np.uint8).copy()
                      if n > 0 and ptr else np.zeros(0, dtype=np.uint8))
    return out


class TablesFromSnapshot:
    """qr_scene_tables and qr_frame_desc rebuilt over the arrays of
    tables_snapshot(); .ptr and .fd stay valid while this object lives."""

    def __init__(self, snap):
        self.arrays = {}
        t = QrSceneTables()
        for k, v in zip(TABLE_COUNTS, snap["counts"].tolist()):
            setattr(t, k, int(v))
        for field, count, rec in TABLE_ARRAYS:
            a = np.ascontiguousarray(snap[field], dtype=np.uint8)
            if a.size != getattr(t, count) * C.sizeof(rec):
                raise ValueError(f"snapshot array {field} has the wrong size")
            self.arrays[field] = a
            ptr = a.ctypes.data if a.size else None
            setattr(t, field, C.cast(ptr, _TABLE_FIELD_TYPES[field]))
        self.tables = t
        self.ptr = C.pointer(t)
        self.fd = QrFrameDesc.from_buffer_copy(np.ascontiguousarray(snap["fd"]).tobytes())


_TABLE_FIELD_TYPES = dict(QrSceneTables._fields_)

_lib = None


def lib():
    """Load libqrhip.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libqrhip.so is not built: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64 = C.c_void_p, C.c_int, C.c_int64
        sig = {
            "qrhip_last_error": (C.c_char_p, []),
            "qrhip_device_open": (i32, [i32, C.POINTER(vp)]),
            "qrhip_device_close": (i32, [vp]),
            "qrhip_device_set_fsaa": (i32, [vp, i32]),
            "qrhip_device_get_fsaa": (i32, [vp]),
            "qrhip_device_get_fsaa_max": (i32, [vp]),
            "qrhip_device_set_path": (i32, [vp, i32]),
            "qrhip_device_get_path": (i32, [vp]),
            "qrhip_side_streams": (i32, []),
            "qrhip_scene_create": (i32, [vp, vp, i32, i32, i32, vp, C.POINTER(vp)]),
            "qrhip_scene_destroy": (i32, [vp]),
            "qrhip_scene_render": (i32, [vp, i64]),
            "qrhip_scene_update": (i32, [vp, i64]),
            "qrhip_scene_launch": (i32, [vp]),
            "qrhip_scene_sync": (i32, [vp]),
            "qrhip_scene_download": (i32, [vp]),
            "qrhip_scene_time_launches": (i32, [vp, i32, C.POINTER(C.c_float)]),
            "qrhip_scene_get_frame": (C.POINTER(C.c_uint32), [vp]),
            "qrhip_scene_get_x_row": (i32, [vp]),
            "qrhip_scene_get_opts": (i32, [vp]),
            "qrhip_scene_set_opts": (i32, [vp, i32]),
            "qrhip_scene_get_pton": (i32, [vp]),
            "qrhip_scene_set_pton": (i32, [vp, i32]),
            "qrhip_scene_next_cam": (i32, [vp]),
            "qrhip_scene_update_action": (i32, [vp, i64, i32]),
            "qrhip_scene_get_cam_idx": (i32, [vp]),
            "qrhip_scene_set_depth": (i32, [vp, i32]),
            "qrhip_scene_get_depth": (i32, [vp]),
            "qrhip_scene_tables": (i32, [vp, C.POINTER(C.POINTER(QrSceneTables)),
                                         C.POINTER(QrFrameDesc)]),
            "qrhip_render_tables": (i32, [vp, C.POINTER(QrSceneTables),
                                          C.POINTER(QrFrameDesc), vp, vp]),
            "qrhip_render_tables_noalloc": (i32, [vp, C.POINTER(QrSceneTables),
                                                  C.POINTER(QrFrameDesc), vp, vp]),
            "qrhip_render_tables_pt": (i32, [vp, C.POINTER(QrSceneTables),
                                             C.POINTER(QrFrameDesc), vp, vp, vp]),
            "qrhip_pt_seed_plane": (i32, [i32, i32, i32, i32, vp]),
            "qrhip_batch_create": (i32, [vp, C.POINTER(vp)]),
            "qrhip_batch_destroy": (i32, [vp]),
            "qrhip_batch_set": (i32, [vp, C.POINTER(QrFrameReq), i32]),
            "qrhip_batch_config": (i32, [vp, i32, i32, i32, i32]),
            "qrhip_batch_launch": (i32, [vp, i32]),
            "qrhip_batch_sync": (i32, [vp]),
            "qrhip_batch_download": (i32, [vp, i32]),
            "qrhip_batch_time": (i32, [vp, i32, i32, C.POINTER(C.c_float)]),
            "qrhip_batch_tune": (i32, [vp, i32, C.POINTER(C.c_int), C.POINTER(C.c_float)]),
            "qrhip_server_create": (i32, [vp, C.POINTER(vp), i32, C.c_uint64, i32,
                                          C.c_uint64, C.POINTER(vp)]),
            "qrhip_server_submit": (i32, [vp, C.POINTER(QrRequest)]),
            "qrhip_server_poll": (i32, [vp, C.POINTER(QrResult), i32, i32]),
            "qrhip_server_footprint": (i32, [vp, C.POINTER(QrRequest),
                                             C.POINTER(C.c_uint64)]),
            "qrhip_server_get_stats": (i32, [vp, C.POINTER(QrServerStats)]),
            "qrhip_server_set_step_ms": (i32, [vp, C.c_double]),
            "qrhip_server_destroy": (i32, [vp]),
            "qrhip_scene_get_pose": (i32, [vp, C.POINTER(C.c_float)]),
            "qrhip_scene_program_stats": (i32, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                C.POINTER(C.c_int)]),
            "qrhip_set_data_path": (i32, [C.c_char_p]),
            "qrhip_image_load": (i32, [C.c_char_p, vp, i64, C.POINTER(C.c_int),
                                       C.POINTER(C.c_int)]),
            "qrhip_image_save": (i32, [C.c_char_p, vp, i32, i32, i64]),
            "qrhip_scene_save_frame": (i32, [vp, i32]),
            "qrgen_create": (vp, [i32, C.c_uint64, i32, C.POINTER(C.c_int)]),
            "qrgen_destroy": (None, [vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class QrError(RuntimeError):
    pass


def check(rc):
    if rc < 0:
        raise QrError(lib().qrhip_last_error().decode())
    return rc


HOST_ONLY = -1
# camera actions (QuadRay's RT_CAMERA_*, core/engine/object.h:47-63)
CAMERA_MOVE_DOWN, CAMERA_MOVE_UP, CAMERA_MOVE_LEFT, CAMERA_MOVE_RIGHT = 2, 3, 4, 5
CAMERA_MOVE_BACK, CAMERA_MOVE_FORWARD = 6, 7
CAMERA_ROTATE_DOWN, CAMERA_ROTATE_UP, CAMERA_ROTATE_LEFT, CAMERA_ROTATE_RIGHT = 10, 11, 12, 13
PATH_LEVELS = 0
PATH_ONE_KERNEL = 1
BATCH_DIRECT = 0
BATCH_GRAPH = 1
BATCH_SERIAL = 2


class Device:
    """rt_Platform over the C ABI; ordinal HOST_ONLY opens one without HIP."""

    def __init__(self, ordinal=0):
        self.h = C.c_void_p()
        check(lib().qrhip_device_open(ordinal, C.byref(self.h)))

    def set_fsaa(self, fsaa):
        return check(lib().qrhip_device_set_fsaa(self.h, fsaa))

    def set_path(self, path):
        return check(lib().qrhip_device_set_path(self.h, path))

    def close(self):
        if self.h:
            lib().qrhip_device_close(self.h)
            self.h = C.c_void_p()

    def render_tables_pt(self, tables_ptr, fd, planes, seeds):
        """One path-tracer pass on caller tables; planes (float32 [3][slots])
        and seeds (uint32 [slots]) are updated in place; returns the frame."""
        assert planes.dtype == np.float32 and planes.flags.c_contiguous
        assert seeds.dtype == np.uint32 and seeds.flags.c_contiguous
        frame = np.zeros((fd.y_res, fd.x_row), dtype=np.uint32)
        check(lib().qrhip_render_tables_pt(self.h, tables_ptr, C.byref(fd),
                                           planes.ctypes.data, seeds.ctypes.data,
                                           frame.ctypes.data))
        return frame[:, :fd.x_res]

    def render_tables_noalloc(self, tables_ptr, fd, want_ids=False):
        """render_tables through the entry point that allocates no device
        memory; returns (frame, ids or None)."""
        ns = {0: 1, 1: 2, 2: 4}[fd.fsaa]
        frame = np.zeros((fd.y_res, fd.x_row), dtype=np.uint32)
        ids = np.zeros((fd.y_res, fd.x_res, ns), dtype=np.int32) if want_ids else None
        check(lib().qrhip_render_tables_noalloc(
            self.h, tables_ptr, C.byref(fd), frame.ctypes.data,
            ids.ctypes.data if want_ids else None))
        return frame[:, :fd.x_res], ids

    def render_tables(self, tables_ptr, fd, want_ids=False):
        """Trace caller-provided tables; returns (frame, ids or None)."""
        ns = {0: 1, 1: 2, 2: 4}[fd.fsaa]
        frame = np.zeros((fd.y_res, fd.x_row), dtype=np.uint32)
        ids = np.zeros((fd.y_res, fd.x_res, ns), dtype=np.int32) if want_ids else None
        check(lib().qrhip_render_tables(
            self.h, tables_ptr, C.byref(fd), frame.ctypes.data,
            ids.ctypes.data if want_ids else None))
        return frame[:, :fd.x_res], ids


def pt_seed_plane(x_res, y_res, x_row, fsaa):
    """Initial seed plane of a path-traced frame (uint32 per sample slot)."""
    ns = {0: 1, 1: 2, 2: 4}[fsaa]
    seeds = np.zeros(x_res * y_res * ns, dtype=np.uint32)
    check(lib().qrhip_pt_seed_plane(x_res, y_res, x_row, fsaa, seeds.ctypes.data))
    return seeds


def set_data_path(root):
    """Root of data/textures/ and dump/ (QuadRay's RT_PATH, "../" at start)."""
    check(lib().qrhip_set_data_path(os.fsencode(root)))


def image_load(path):
    """A 24-bit BMP as uint32 texels [y_dim][x_dim], rows as stored."""
    x, y = C.c_int(), C.c_int()
    check(lib().qrhip_image_load(os.fsencode(path), None, 0, C.byref(x), C.byref(y)))
    tex = np.zeros((y.value, x.value), dtype=np.uint32)
    check(lib().qrhip_image_load(os.fsencode(path), tex.ctypes.data, tex.size,
                                 C.byref(x), C.byref(y)))
    return tex


def image_save(path, pixels, top_down=False):
    """Write uint32 pixels [rows][cols] as a 24-bit BMP (height stored
    negative when top_down)."""
    assert pixels.dtype == np.uint32 and pixels.ndim == 2 and pixels.strides[1] == 4
    rows, cols = pixels.shape
    check(lib().qrhip_image_save(os.fsencode(path), pixels.ctypes.data, cols,
                                 -rows if top_down else rows, pixels.strides[0] // 4))


class Workload:
    """Seeded synthetic rt_SCENE (config 1 plain, 2 Whitted 64 quadrics, ...)."""

    def __init__(self, config, seed, ids=False, texture_files=False, animated=False):
        info = (C.c_int * 5)()
        flags = (1 if ids else 0) | (2 if texture_files else 0) | (4 if animated else 0)
        self.ptr = lib().qrgen_create(config, seed, flags, info)
        if not self.ptr:
            raise QrError(lib().qrhip_last_error().decode())
        self.x_res, self.y_res, self.fsaa, self.depth, self.spp = list(info)

    def close(self):
        if self.ptr:
            lib().qrgen_destroy(self.ptr)
            self.ptr = None


class Scene:
    """rt_Scene over the C ABI; dev=None gives a host-only scene."""

    def __init__(self, dev, scn_ptr, x_res, y_res, x_row=None):
        self.h = C.c_void_p()
        self.x_res, self.y_res = x_res, y_res
        self.x_row = x_res if x_row is None else x_row
        check(lib().qrhip_scene_create(dev.h if dev is not None else None, scn_ptr,
                                       x_res, y_res, self.x_row, None,
                                       C.byref(self.h)))

    def destroy(self):
        if self.h:
            lib().qrhip_scene_destroy(self.h)
            self.h = C.c_void_p()

    def set_depth(self, d):
        return check(lib().qrhip_scene_set_depth(self.h, d))

    def set_pton(self, n):
        return check(lib().qrhip_scene_set_pton(self.h, n))

    def update(self, time=0):
        check(lib().qrhip_scene_update(self.h, time))

    def update_action(self, time, action):
        """rt_Scene::update(time, action): a camera action (CAMERA_*)."""
        check(lib().qrhip_scene_update_action(self.h, time, action))

    def render(self, time=0):
        check(lib().qrhip_scene_render(self.h, time))
        return self.frame()

    def frame(self):
        """Copy of the host framebuffer as [y][x] (row 0 first, also for a
        negative x_row, whose rows run backwards in memory)."""
        p = lib().qrhip_scene_get_frame(self.h)
        ar = abs(self.x_row)
        base = C.addressof(p.contents)
        if self.x_row < 0:
            base -= (self.y_res - 1) * ar * 4
        buf = (C.c_uint32 * (self.y_res * ar)).from_address(base)
        a = np.ctypeslib.as_array(buf).reshape(self.y_res, ar)
        if self.x_row < 0:
            a = a[::-1]
        return a[:, :self.x_res].copy()

    def tables(self):
        t = C.POINTER(QrSceneTables)()
        fd = QrFrameDesc()
        check(lib().qrhip_scene_tables(self.h, C.byref(t), C.byref(fd)))
        return t, fd

    def pose(self):
        """Current camera: right, down and view axes, then position."""
        p = (C.c_float * 12)()
        check(lib().qrhip_scene_get_pose(self.h, p))
        return np.array(list(p), dtype=np.float32)

    def program_stats(self, compare=False):
        """(builds, device refits, steps differing from the host tables or
        None when not compared) of the scene's culling programs."""
        b, r, m = C.c_int(), C.c_int(), C.c_int()
        check(lib().qrhip_scene_program_stats(self.h, C.byref(b), C.byref(r),
                                              C.byref(m) if compare else None))
        return b.value, r.value, (m.value if compare else None)

    def save_frame(self, index):
        """rt_Scene::save_frame: the frame to <data path>dump/scrXXX.bmp."""
        check(lib().qrhip_scene_save_frame(self.h, index))

    def time_launches(self, n):
        ms = C.c_float()
        check(lib().qrhip_scene_time_launches(self.h, n, C.byref(ms)))
        return ms.value


class Batch:
    """Frames traced together (qrhip_batch_*): set(requests) with requests
    (scene, x_res, y_res, fsaa, depth); frames come back as numpy arrays."""

    def __init__(self, dev):
        self.h = C.c_void_p()
        check(lib().qrhip_batch_create(dev.h, C.byref(self.h)))
        self.reqs, self.bufs = None, []

    def set(self, requests):
        n = len(requests)
        self.reqs = (QrFrameReq * n)()
        self.bufs = []
        for k, (scene, x, y, fsaa, depth) in enumerate(requests):
            buf = np.zeros((y, x), dtype=np.uint32)
            self.bufs.append(buf)
            self.reqs[k] = QrFrameReq(scene.h, x, y, fsaa, depth,
                                      buf.ctypes.data_as(C.POINTER(C.c_uint32)))
        check(lib().qrhip_batch_set(self.h, self.reqs, n))

    def config(self, grid_per_cu, tile_w, rays_per_thread=1, cadence=1):
        check(lib().qrhip_batch_config(self.h, grid_per_cu, tile_w, rays_per_thread, cadence))

    def launch(self, mode=BATCH_DIRECT):
        check(lib().qrhip_batch_launch(self.h, mode))

    def sync(self):
        check(lib().qrhip_batch_sync(self.h))

    def frames(self):
        """Sync, download every frame and return copies."""
        check(lib().qrhip_batch_download(self.h, -1))
        return [b.copy() for b in self.bufs]

    def render(self, mode=BATCH_DIRECT):
        self.launch(mode)
        return self.frames()

    def time(self, n, mode=BATCH_DIRECT):
        ms = C.c_float()
        check(lib().qrhip_batch_time(self.h, n, mode, C.byref(ms)))
        return ms.value

    def tune(self, reps=2):
        """(workgroups per CU, tile width, rays per thread, cadence), ms"""
        best = (C.c_int * 4)()
        ms = C.c_float()
        check(lib().qrhip_batch_tune(self.h, reps, best, C.byref(ms)))
        return tuple(best), ms.value

    def destroy(self):
        if self.h:
            lib().qrhip_batch_destroy(self.h)
            self.h = C.c_void_p()


class Server:
    """The serving loop (qrhip_server_*): scenes stay resident; submit()
    queues a request and returns its host frame buffer, poll() returns
    finished results as dicts."""

    def __init__(self, dev, scenes, budget_bytes, max_active, step_samples=0, step_ms=0.0):
        self.h = C.c_void_p()
        arr = (C.c_void_p * len(scenes))(*[s.h.value for s in scenes])
        check(lib().qrhip_server_create(dev.h, arr, len(scenes), budget_bytes,
                                        max_active, step_samples, C.byref(self.h)))
        if step_ms > 0.0:
            check(lib().qrhip_server_set_step_ms(self.h, step_ms))
        self.frames = {}

    @staticmethod
    def request(tag, scene, x, y, fsaa=0, depth=4, spp=0, time=0, pose=None):
        r = QrRequest()
        r.tag, r.scene, r.x_res, r.y_res = tag, scene, x, y
        r.fsaa, r.depth, r.spp, r.time = fsaa, depth, spp, time
        if pose is not None:
            r.use_pose = 1
            for k in range(12):
                r.pose[k] = float(pose[k])
        return r

    def footprint(self, req):
        b = C.c_uint64()
        check(lib().qrhip_server_footprint(self.h, C.byref(req), C.byref(b)))
        return b.value

    def submit(self, req):
        buf = np.zeros((req.y_res, req.x_res), dtype=np.uint32)
        req.frame = buf.ctypes.data_as(C.POINTER(C.c_uint32))
        check(lib().qrhip_server_submit(self.h, C.byref(req)))
        self.frames[req.tag] = buf
        return buf

    def poll(self, max_n=256, wait_ms=100):
        out = (QrResult * max_n)()
        n = check(lib().qrhip_server_poll(self.h, out, max_n, wait_ms))
        return [{"tag": r.tag, "status": r.status, "steps": r.steps,
                 "submit_ms": r.submit_ms, "admit_ms": r.admit_ms,
                 "done_ms": r.done_ms, "device_ms": r.device_ms} for r in out[:n]]

    def stats(self):
        st = QrServerStats()
        check(lib().qrhip_server_get_stats(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in QrServerStats._fields_}

    def destroy(self):
        if self.h:
            rc = lib().qrhip_server_destroy(self.h)
            self.h = C.c_void_p()
            check(rc)
