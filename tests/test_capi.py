"""The C ABI of libqrhip.so: exported symbols, argument checks and the
rt_Scene error behaviour, on the host (no HIP device calls)."""
import ctypes as C
import os
import re

import pytest

import qrhip
from helpers import HostScene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(inc, fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(qrhip_\w+|qrgen_\w+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_function():
    lib = qrhip.lib()
    names = declared_functions()
    assert len(names) >= 28
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_function():
    """The ctypes layer sets a signature for every C entry point."""
    lib = qrhip.lib()
    for n in declared_functions():
        assert getattr(lib, n).argtypes is not None, n


def test_struct_sizes_match_header():
    # qr_surface: 22 floats + 21 ints/uints (incl. pad) = 172 bytes + ...
    assert C.sizeof(qrhip.QrSurface) % 4 == 0
    assert C.sizeof(qrhip.QrFrameDesc) == 4 * (12 + 4 + 1 + 11 + 8)


def test_host_only_platform_fsaa():
    d = qrhip.Device(qrhip.HOST_ONLY)
    try:
        assert qrhip.lib().qrhip_device_get_fsaa(d.h) == 0
        assert d.set_fsaa(7) == 2          # clamped like rt_Platform::set_fsaa
        assert d.set_fsaa(-3) == 0
        assert qrhip.lib().qrhip_device_get_fsaa_max(d.h) == 2
    finally:
        d.close()


def test_scene_locking_and_errors():
    w = qrhip.Workload(1, 1)
    d = qrhip.Device(qrhip.HOST_ONLY)
    try:
        s = qrhip.Scene(d, w.ptr, 64, 32)
        with pytest.raises(qrhip.QrError, match="locked by another instance"):
            qrhip.Scene(d, w.ptr, 64, 32)
        s.destroy()
        s2 = qrhip.Scene(d, w.ptr, 64, 32)   # unlocked again by destroy
        s2.destroy()
        with pytest.raises(qrhip.QrError, match="dimensions are not valid"):
            qrhip.Scene(d, w.ptr, 0, 32)
        with pytest.raises(qrhip.QrError, match="dimensions are not valid"):
            qrhip.Scene(d, w.ptr, 64, 32, x_row=32)
    finally:
        d.close()
        w.close()


def test_null_arguments_fail_cleanly():
    L = qrhip.lib()
    out = C.c_void_p()
    assert L.qrhip_scene_create(None, None, 64, 32, 64, None, C.byref(out)) < 0
    assert L.qrhip_scene_render(None, 0) == qrhip_err("E_ARG")
    assert L.qrhip_device_close(None) == qrhip_err("E_ARG")
    assert L.qrhip_scene_get_frame(None) is None or not L.qrhip_scene_get_frame(None)


def qrhip_err(name):
    return {"E_ARG": -1, "E_SCENE": -2, "E_DEVICE": -3, "E_NODEVICE": -4,
            "E_UNSUPPORTED": -5}[name]


def test_scene_settings_round_trip():
    hs = HostScene(2, 1, 64, 36, 0, 4)
    try:
        L, h = qrhip.lib(), hs.scene.h
        assert L.qrhip_scene_get_depth(h) == 4
        assert L.qrhip_scene_set_depth(h, 15) == qrhip_err("E_ARG")
        assert L.qrhip_scene_set_depth(h, -1) == qrhip_err("E_ARG")
        assert L.qrhip_scene_get_depth(h) == 4
        opts = L.qrhip_scene_get_opts(h)
        assert opts > 0
        L.qrhip_scene_set_opts(h, opts & ~0x1)
        assert L.qrhip_scene_get_opts(h) == opts & ~0x1
        assert L.qrhip_scene_get_pton(h) == 0
        assert L.qrhip_scene_get_x_row(h) == 64
        # a host-only scene updates but does not render
        assert L.qrhip_scene_render(h, 0) == qrhip_err("E_NODEVICE")
        assert L.qrhip_scene_get_cam_idx(h) == 0
    finally:
        hs.close()


def test_tables_are_consistent():
    """Every index in the tables stays inside its table (what the GPU entry
    points check before a launch)."""
    for cfg in (1, 2, 3, 4):
        hs = HostScene(cfg, 3, 64, 36, 0, 4)
        try:
            t = hs.tables.contents
            n = t.n_surf
            assert n > 0 and t.n_elem > 0 and t.n_lgt > 0
            for i in range(n):
                s = t.surfaces[i]
                if s.trm:
                    assert 0 <= s.trnode < n
                assert 0 <= s.clip_beg and s.clip_beg + s.clip_cnt <= t.n_clip
                if s.solver:
                    assert 0 <= s.mat[0] < t.n_mat and 0 <= s.mat[1] < t.n_mat
            assert hs.fd.x_res == 64 and hs.fd.y_res == 36 and hs.fd.depth == 4
        finally:
            hs.close()


def test_no_device_here_is_reported():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    out = C.c_void_p()
    rc = qrhip.lib().qrhip_device_open(0, C.byref(out))
    assert rc == qrhip_err("E_DEVICE")
    assert b"no HIP device" in qrhip.lib().qrhip_last_error()


def test_tables_snapshot_round_trip():
    """qrhip.tables_snapshot copies every table as bytes of the record sizes
    the header declares (the golden test-scene fixtures are such snapshots):
    the oracle renders the rebuilt tables as it renders the scene's own."""
    import qr_oracle
    hs = HostScene(9, 3, 96, 54, 2, 6)
    try:
        want = hs.oracle()[0]
        snap = qrhip.tables_snapshot(hs.tables, hs.fd)
    finally:
        hs.close()
    t = qrhip.TablesFromSnapshot(snap)
    got = qr_oracle.render(t.ptr, t.fd)[0]
    assert (got == want).all()
    assert t.tables.n_surf > 20 and t.tables.n_step > 0 and t.tables.n_texel > 0
