"""BMP texture files and frame dumps (QuadRay core/engine/rtimag.cpp
load_image 47-186, save_image 204-361; rt_Scene::save_frame,
core/engine/engine.cpp:3863-3899), on the host.

The loader is pinned by BMP files built byte by byte here from the format's
published layout (a 54-byte header, 24-bit pixels stored B, G, R, rows padded
to 4 bytes) with known texels, and by files that break one of QuadRay's
checks; the writer by the exact header QuadRay writes and by a round trip;
texture files in a scene by the tables and the oracle's frame of the same
scene with texel arrays."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

import qrhip
import qr_oracle
from helpers import HostScene


def bmp_bytes(texels, height_sign=1, depth=24, planes=1, comp=0, sig=b"BM",
              offset=54, pad_last=True, gap=b""):
    """A BMP of texels [rows][cols] (0x00RRGGBB) in file row order."""
    rows, cols = texels.shape
    pad = (4 - (3 * cols) % 4) % 4
    body = bytearray()
    for y in range(rows):
        for x in range(cols):
            v = int(texels[y, x])
            body += bytes([v & 0xFF, (v >> 8) & 0xFF, (v >> 16) & 0xFF])
        if y < rows - 1 or pad_last:
            body += b"\0" * pad
    head = sig + struct.pack("<IHHI", 54 + len(body), 0, 0, offset + len(gap))
    head += struct.pack("<IiiHHIIiiII", 40, cols, height_sign * rows, planes, depth,
                        comp, len(body), 2835, 2835, 0, 0)
    assert len(head) == 54
    return head + (b"\0" * (offset - 54)) + gap + bytes(body)


@pytest.fixture
def data_dir(tmp_path):
    (tmp_path / "data" / "textures").mkdir(parents=True)
    (tmp_path / "dump").mkdir()
    qrhip.set_data_path(str(tmp_path) + os.sep)
    yield tmp_path
    qrhip.set_data_path("../")


@pytest.mark.parametrize("cols", [1, 2, 3, 4, 5, 7])
def test_load_padded_rows(tmp_path, cols):
    rng = np.random.default_rng(cols)
    tex = rng.integers(0, 1 << 24, size=(3, cols), dtype=np.uint32)
    p = tmp_path / "t.bmp"
    p.write_bytes(bmp_bytes(tex))
    got = qrhip.image_load(str(p))
    assert got.shape == (3, cols)
    assert np.array_equal(got, tex)         # rows as stored, top byte 0


def test_load_follows_pixel_offset_and_negative_height(tmp_path):
    tex = np.array([[0x102030, 0xA0B0C0], [0x0000FF, 0xFF0000]], dtype=np.uint32)
    p = tmp_path / "t.bmp"
    p.write_bytes(bmp_bytes(tex, offset=54, gap=b"\xEE" * 10))
    assert np.array_equal(qrhip.image_load(str(p)), tex)
    # a top-down file loads as stored too: QuadRay takes |height|
    p.write_bytes(bmp_bytes(tex, height_sign=-1))
    assert np.array_equal(qrhip.image_load(str(p)), tex)


@pytest.mark.parametrize("bad", [dict(sig=b"BA"), dict(depth=32), dict(depth=8),
                                 dict(planes=2), dict(comp=1), dict(pad_last=False)])
def test_load_rejects_what_quadray_rejects(tmp_path, bad):
    tex = np.arange(6, dtype=np.uint32).reshape(2, 3) * 0x010203
    p = tmp_path / "t.bmp"
    p.write_bytes(bmp_bytes(tex, **bad))
    with pytest.raises(qrhip.QrError, match="failed to load image"):
        qrhip.image_load(str(p))


def test_load_rejects_short_and_missing_files(tmp_path):
    tex = np.arange(12, dtype=np.uint32).reshape(3, 4)
    raw = bmp_bytes(tex)
    p = tmp_path / "t.bmp"
    p.write_bytes(raw[:-1])
    with pytest.raises(qrhip.QrError, match="failed to load image"):
        qrhip.image_load(str(p))
    p.write_bytes(raw[:30])
    with pytest.raises(qrhip.QrError, match="failed to load image"):
        qrhip.image_load(str(p))
    with pytest.raises(qrhip.QrError, match="failed to load image"):
        qrhip.image_load(str(tmp_path / "absent.bmp"))


def test_load_into_small_buffer_fails(tmp_path):
    tex = np.arange(12, dtype=np.uint32).reshape(3, 4)
    p = tmp_path / "t.bmp"
    p.write_bytes(bmp_bytes(tex))
    buf = np.zeros(11, dtype=np.uint32)
    x, y = C.c_int(), C.c_int()
    rc = qrhip.lib().qrhip_image_load(os.fsencode(str(p)), buf.ctypes.data, buf.size,
                                      C.byref(x), C.byref(y))
    assert rc < 0 and (x.value, y.value) == (4, 3)


@pytest.mark.parametrize("cols,rows,top_down", [(5, 3, False), (4, 2, True), (1, 1, False)])
def test_save_header_and_round_trip(tmp_path, cols, rows, top_down):
    rng = np.random.default_rng(cols * 10 + rows)
    px = rng.integers(0, 1 << 32, size=(rows, cols), dtype=np.uint32)
    p = tmp_path / "s.bmp"
    qrhip.image_save(str(p), px, top_down=top_down)
    raw = p.read_bytes()
    pad = (4 - (3 * cols) % 4) % 4
    # QuadRay's header: sizes without row padding, 4000 pixels per metre
    want = b"BM" + struct.pack("<IIIIiiHHIIIIII", 54 + 3 * cols * rows, 0, 54, 40, cols,
                               -rows if top_down else rows, 1, 24, 0, 3 * cols * rows,
                               4000, 4000, 0, 0)
    assert raw[:54] == want
    assert len(raw) == 54 + rows * (3 * cols + pad)
    back = qrhip.image_load(str(p))
    assert np.array_equal(back, px & 0xFFFFFF)      # alpha is not stored


def test_save_strided_rows(tmp_path):
    base = np.arange(7 * 4, dtype=np.uint32).reshape(4, 7) * 0x030507
    view = base[:, 1:6]
    p = tmp_path / "s.bmp"
    qrhip.image_save(str(p), view)
    assert np.array_equal(qrhip.image_load(str(p)), view & 0xFFFFFF)


def test_save_rejects_bad_arguments(tmp_path):
    px = np.zeros((2, 2), dtype=np.uint32)
    L = qrhip.lib()
    path = os.fsencode(str(tmp_path / "s.bmp"))
    assert L.qrhip_image_save(path, px.ctypes.data, 0, 2, 2) < 0
    assert L.qrhip_image_save(path, px.ctypes.data, 2, 0, 2) < 0
    assert L.qrhip_image_save(path, px.ctypes.data, 2, 2, 1) < 0
    assert L.qrhip_image_save(path, None, 2, 2, 2) < 0
    bad = os.fsencode(str(tmp_path / "no" / "dir" / "s.bmp"))
    with pytest.raises(qrhip.QrError, match="failed to save image"):
        qrhip.check(L.qrhip_image_save(bad, px.ctypes.data, 2, 2, 2))


def test_frame_names():
    """rt_Scene::save_frame's names: scrXXX.bmp, scrXXX-Y.bmp from 1000."""
    d = qrhip.Device(qrhip.HOST_ONLY)
    w = qrhip.Workload(1, 1)
    s = qrhip.Scene(d, w.ptr, 8, 4)
    base = None
    try:
        import tempfile
        with tempfile.TemporaryDirectory() as tmp:
            os.mkdir(os.path.join(tmp, "dump"))
            qrhip.set_data_path(tmp + os.sep)
            for index, name in [(0, "scr000.bmp"), (7, "scr007.bmp"), (42, "scr042.bmp"),
                                (999, "scr999.bmp"), (1000, "scr000-0.bmp"),
                                (1234, "scr234-0.bmp"), (2345, "scr345-1.bmp"),
                                (9999, "scr999-8.bmp")]:
                s.save_frame(index)
                assert os.path.exists(os.path.join(tmp, "dump", name)), (index, name)
                os.remove(os.path.join(tmp, "dump", name))
            base = os.listdir(os.path.join(tmp, "dump"))
    finally:
        qrhip.set_data_path("../")
        s.destroy()
        w.close()
        d.close()
    assert base == []


@pytest.mark.parametrize("x_row_sign", [1, -1])
def test_save_frame_writes_top_down_frame(data_dir, x_row_sign):
    d = qrhip.Device(qrhip.HOST_ONLY)
    w = qrhip.Workload(1, 1)
    s = qrhip.Scene(d, w.ptr, 6, 3, x_row=x_row_sign * 8)
    try:
        p = qrhip.lib().qrhip_scene_get_frame(s.h)
        frame = np.arange(18, dtype=np.uint32).reshape(3, 6) * 0x01020305 + 0xFF000000
        for y in range(3):
            for x in range(6):
                p[y * x_row_sign * 8 + x] = int(frame[y, x])
        s.save_frame(5)
        raw = (data_dir / "dump" / "scr005.bmp").read_bytes()
        assert struct.unpack("<ii", raw[18:26]) == (6, -3)
        got = qrhip.image_load(str(data_dir / "dump" / "scr005.bmp"))
        assert np.array_equal(got, frame & 0xFFFFFF)
        with pytest.raises(qrhip.QrError):
            s.save_frame(-1)
    finally:
        s.destroy()
        w.close()
        d.close()


def test_save_frame_without_dump_folder(tmp_path):
    d = qrhip.Device(qrhip.HOST_ONLY)
    w = qrhip.Workload(1, 1)
    s = qrhip.Scene(d, w.ptr, 4, 4)
    qrhip.set_data_path(str(tmp_path) + os.sep)
    try:
        with pytest.raises(qrhip.QrError, match="failed to save image"):
            s.save_frame(1)
    finally:
        qrhip.set_data_path("../")
        s.destroy()
        w.close()
        d.close()


def table_texels(hs):
    t = hs.tables.contents
    return np.ctypeslib.as_array((C.c_uint32 * t.n_texel).from_address(t.texels)).copy()


def test_scene_with_texture_files_equals_texel_arrays(data_dir):
    """A material naming a BMP file under data/textures/ gives the tables and
    the frame of the same material with the texel array in place."""
    arr = HostScene(2, 77, 48, 27, 0, 2)
    fil = HostScene(2, 77, 48, 27, 0, 2, texture_files=True)
    try:
        names = sorted(os.listdir(data_dir / "data" / "textures"))
        assert names and all(n.startswith("qrgen_2_77_") for n in names)
        ta, tf = table_texels(arr), table_texels(fil)
        # file texels carry a zero top byte, which no colour reads
        assert np.array_equal(tf & 0xFFFFFF, ta & 0xFFFFFF)
        assert not np.array_equal(tf, ta)
        assert arr.tables.contents.n_mat == fil.tables.contents.n_mat
        fa = qr_oracle.render(arr.tables, arr.fd)[0]
        ff = qr_oracle.render(fil.tables, fil.fd)[0]
        assert np.array_equal(fa, ff)
    finally:
        arr.close()
        fil.close()


def test_scene_with_missing_texture_file_fails(data_dir):
    w = qrhip.Workload(2, 78, texture_files=True)
    for n in os.listdir(data_dir / "data" / "textures"):
        os.remove(data_dir / "data" / "textures" / n)
    d = qrhip.Device(qrhip.HOST_ONLY)
    try:
        with pytest.raises(qrhip.QrError, match="failed to load image"):
            qrhip.Scene(d, w.ptr, 16, 16).update(0)
    finally:
        d.close()
        w.close()
