"""Shared helpers of the test suite: golden fixtures, host-only scenes,
oracle renders and the parity bar."""
import json
import os

import numpy as np

import qrhip
import qr_oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _index():
    with open(os.path.join(GOLDEN, "index.json")) as f:
        return json.load(f)


def golden_cases(path_tracer=False):
    """Golden cases of the Whitted tracer on the seeded workloads, or of the
    path tracer."""
    return {k: v for k, v in _index().items()
            if bool(v.get("path_tracer", False)) == path_tracer
            and not v.get("tables", False) and "actions" not in v}


def golden_camera_cases():
    """Golden cases with camera actions between renders."""
    return {k: v for k, v in _index().items() if "actions" in v}


def replay_actions(scene, actions, render):
    """After a first update at time 0, each group "T:a,a" of actions at time
    T followed by render(T) (an update, or a render on a device), as
    ref_driver -c does with QuadRay's rt_Scene; returns the last result."""
    out = render(0)
    for grp in actions.split(";"):
        t, acts = grp.split(":")
        for a in acts.split(","):
            scene.update_action(int(t), int(a))
        out = render(int(t))
    return out


def golden_table_cases():
    """Golden cases of QuadRay's own test scenes, stored with their tables."""
    return {k: v for k, v in _index().items() if v.get("tables", False)}


class GoldenTables:
    """A stored test-scene case: .ptr / .fd (the tables and frame description
    this renderer's update stage made of it), .frame (QuadRay's fp32 frame),
    .frame64 (its fp64 build's)."""

    def __init__(self, name):
        with np.load(os.path.join(GOLDEN, name + ".npz")) as z:
            snap = {k: z[k].copy() for k in z.files}
        self.frame = snap.pop("frame")
        self.frame64 = snap.pop("frame64")
        self._t = qrhip.TablesFromSnapshot(snap)
        self.ptr, self.fd = self._t.ptr, self._t.fd

    def oracle(self, want_ids=False, audit=None):
        return qr_oracle.render(self.ptr, self.fd, want_ids=want_ids, audit=audit)


def golden_frame64(name):
    """QuadRay's fp64 frame of a Whitted golden case (the accuracy reference)."""
    with np.load(os.path.join(GOLDEN, name + ".npz")) as z:
        return z["frame64"].copy()


def golden_planes(name):
    with np.load(os.path.join(GOLDEN, name + ".npz")) as z:
        return z["planes"].copy()


def pt_passes(hs, passes, seeds, render):
    """Run path-tracer passes with render(tables, fd, planes, seeds) from a
    zero start; returns (planes [y][x][sample][rgb], per-slot variance of the
    pass samples, last frame)."""
    fd = hs.fd
    ns = {0: 1, 1: 2, 2: 4}[fd.fsaa]
    slots = fd.x_res * fd.y_res * ns
    planes = np.zeros((3, slots), dtype=np.float32)
    prev = np.zeros((3, slots))
    s1 = np.zeros((3, slots))
    s2 = np.zeros((3, slots))
    frame = None
    fd.pt_on = 1
    for p in range(1, passes + 1):
        fd.pts_c = float(p)
        frame = render(hs.tables, fd, planes, seeds)
        cur = planes.astype(np.float64)
        sample = cur * p - prev * (p - 1)
        s1 += sample
        s2 += sample * sample
        prev = cur
    var = np.maximum(s2 / passes - (s1 / passes) ** 2, 0.0)
    shape = (fd.y_res, fd.x_res, ns, 3)
    return planes.T.reshape(shape), var.T.reshape(shape), frame


def golden_frame(name):
    with np.load(os.path.join(GOLDEN, name + ".npz")) as z:
        return z["frame"].copy()


class HostScene:
    """A seeded workload instantiated on a host-only platform: the update
    stage runs and its tables feed the oracle."""

    def __init__(self, config, seed, x_res, y_res, fsaa, depth, ids=False, time=0,
                 texture_files=False, animated=False):
        self.work = qrhip.Workload(config, seed, ids=ids, texture_files=texture_files,
                                   animated=animated)
        self.dev = qrhip.Device(qrhip.HOST_ONLY)
        self.dev.set_fsaa(fsaa)
        self.scene = qrhip.Scene(self.dev, self.work.ptr, x_res, y_res)
        self.scene.set_depth(depth)
        self.scene.update(time)
        self.tables, self.fd = self.scene.tables()

    def oracle(self, want_ids=False, rows=None, audit=None):
        return qr_oracle.render(self.tables, self.fd, want_ids=want_ids, rows=rows,
                                audit=audit)

    def close(self):
        self.scene.destroy()
        self.dev.close()
        self.work.close()


def channels(frame):
    f = frame.astype(np.int64)
    return np.stack([(f >> 16) & 255, (f >> 8) & 255, f & 255], -1)


def frame_stats(a, b):
    """(fraction of identical pixels, fraction within +-1 per channel,
    max channel difference)."""
    d = np.abs(channels(a) - channels(b)).max(-1)
    return float((a == b).mean()), float((d <= 1).mean()), int(d.max())


def assert_frame_parity(got, want, what=""):
    """ARGB8 within +-1 per channel on at least 99.9% of the pixels."""
    assert got.shape == want.shape, (got.shape, want.shape)
    exact, within, dmax = frame_stats(got, want)
    assert within >= 0.999, f"{what}: exact {exact:.6f} within1 {within:.6f} max {dmax}"
    return exact, within, dmax


def assert_frame_exact(got, want, what=""):
    """Bit-identical frames.  The Whitted kernels apply QuadRay's operations
    in QuadRay's order, as the oracle does, so they meet the +-1 bar with
    every pixel exact; any difference is a defect, however small."""
    exact, within, dmax = assert_frame_parity(got, want, what)
    assert exact == 1.0, f"{what}: exact {exact:.6f} within1 {within:.6f} max {dmax}"
    return exact, within, dmax


def surface_boxes(tables):
    """Per-surface culling boxes of a program (7 floats per surface record:
    centre, half extents, frame -- 0 world, 1 the surface's tracing frame;
    NaN where the surface has none): an element step's own local box, else
    the box step right before it that skips just that element."""
    import ctypes as C
    t = tables.contents
    boxes = np.full((t.n_surf, 7), np.nan, dtype=np.float32)
    elems = np.ctypeslib.as_array(C.cast(t.elems, C.POINTER(C.c_int32)),
                                  shape=(t.n_elem * 2,)).reshape(t.n_elem, 2)
    for p in range(t.n_step):
        st = t.steps[p]
        if st.elem >= 0 and st.skip == -2:
            boxes[elems[st.elem, 0]] = list(st.c) + list(st.h) + [1.0]
        elif (st.elem == -1 and st.skip == p + 2 and p + 1 < t.n_step
              and t.steps[p + 1].elem >= 0 and t.steps[p + 1].skip == -1):
            boxes[elems[t.steps[p + 1].elem, 0]] = list(st.c) + list(st.h) + [0.0]
    return boxes


def oracle_rows(cases, threads=16):
    """Oracle rows of many frames at once: cases are (config, seed, x, y, fsaa,
    depth, time, animated, rows); returns one {row: pixels} per case.  Frames
    run in parallel threads (the oracle releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor

    def one(c):
        config, seed, x, y, fsaa, depth, time, animated, rows = c
        hs = HostScene(config, seed, x, y, fsaa, depth, time=time, animated=animated)
        try:
            f = qr_oracle.render(hs.tables, hs.fd, rows=rows, threads=1)[0]
        finally:
            hs.close()
        return {r: f[r].copy() for r in rows}

    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(one, cases))
