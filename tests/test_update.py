"""The update stage's culling programs across updates, on the host.

An update builds the programs when the surface list or a surface's boxes
change kind, and otherwise refits the boxes of the programs it has.  Animated
workloads (QRGEN_ANIMATED: arrays turning and moving with time) exercise the
refit: the refitted programs keep their structure and stay conservative --
every root the oracle accepts lies inside its surface's boxes -- and every
other table equals the one a fresh scene builds at the same time."""
import ctypes as C
import os

import numpy as np
import pytest

import qrhip
import qr_oracle
from helpers import GOLDEN, golden_table_cases, surface_boxes
from test_culling import FUZZ, program_elements


def host_scene(config, seed, animated, x=96, y=54, fsaa=0, depth=4):
    dev = qrhip.Device(qrhip.HOST_ONLY)
    dev.set_fsaa(fsaa)
    w = qrhip.Workload(config, seed, animated=animated)
    s = qrhip.Scene(dev, w.ptr, x, y)
    s.set_depth(depth)
    return dev, w, s


def close(dev, w, s):
    s.destroy()
    w.close()
    dev.close()


def table_bytes(t, name, count, size):
    p = getattr(t, name)
    return C.string_at(C.cast(p, C.c_void_p), count * size) if count else b""


def programs(t):
    return (table_bytes(t, "steps", t.n_step, 32), table_bytes(t, "steps_any", t.n_step_any, 32))


def structure(t):
    return ([(t.steps[p].elem, t.steps[p].skip) for p in range(t.n_step)],
            [(t.steps_any[p].elem, t.steps_any[p].skip) for p in range(t.n_step_any)])


def test_static_scene_refits_to_the_same_programs():
    dev, w, s = host_scene(2, 3, animated=False)
    try:
        s.update(0)
        first = programs(s.tables()[0].contents)
        s.update(500)
        s.update(1000)
        again = programs(s.tables()[0].contents)
        builds, refits, _ = s.program_stats()
    finally:
        close(dev, w, s)
    assert builds == 1 and refits == 0     # host-only: no device refits
    assert again == first


@pytest.mark.parametrize("config,seed", [(2, 4), (4, 5), (7, 6)])
def test_animated_refit_is_conservative(config, seed):
    dev, w, s = host_scene(config, seed, animated=True)
    try:
        s.update(0)
        t0 = s.tables()[0].contents
        shape0 = structure(t0)
        srf0 = table_bytes(t0, "surfaces", t0.n_surf, C.sizeof(qrhip.QrSurface))
        for time in (350, 900, 1600):
            s.update(time)
            t, fd = s.tables()
            assert structure(t.contents) == shape0       # refitted, not rebuilt
            audit = {}
            qr_oracle.render(t, fd, boxes=surface_boxes(t), audit=audit)
            assert audit["checked"] > 100
            assert audit["outside"] == 0
            assert audit["excess"] < 0.75 * FUZZ
        tl = s.tables()[0].contents
        assert table_bytes(tl, "surfaces", tl.n_surf, C.sizeof(qrhip.QrSurface)) != srf0
        builds, _, _ = s.program_stats()
        assert builds == 1
        # the refitted programs' elements are those of the built ones
        assert program_elements(tl.n_step, tl.steps) == list(range(tl.n_elem))
    finally:
        close(dev, w, s)


def test_animated_tables_depend_on_time_alone():
    """A scene updated at several times has, at its last time, the tables
    and the oracle frame of a scene updated at that time only."""
    dev, w, s = host_scene(2, 7, animated=True)
    dev2, w2, s2 = host_scene(2, 7, animated=True)
    try:
        for time in (0, 400, 1250):
            s.update(time)
        s2.update(1250)
        a, fa = s.tables()
        b, fb = s2.tables()
        ac, bc = a.contents, b.contents
        for name, n, size in [("surfaces", ac.n_surf, C.sizeof(qrhip.QrSurface)),
                              ("lights", ac.n_lgt, 48), ("elems", ac.n_elem, 8)]:
            assert table_bytes(ac, name, n, size) == table_bytes(bc, name, n, size), name
        ga = qr_oracle.render(a, fa)[0]
        gb = qr_oracle.render(b, fb)[0]
        assert np.array_equal(ga, gb)
    finally:
        close(dev, w, s)
        close(dev2, w2, s2)


def test_transformed_exactly_when_traced_in_a_node_frame():
    """The kernels keep one local hit per root because every surface record
    has trm != 0 exactly when shift != 0 (the engine refuses a scene where
    that fails): over every workload, animated or not."""
    dev = qrhip.Device(qrhip.HOST_ONLY)
    n = 0
    try:
        for config in range(1, 9):
            for seed in (1, 2, 3):
                for animated in (False, True):
                    w = qrhip.Workload(config, seed, animated=animated)
                    s = qrhip.Scene(dev, w.ptr, 32, 32)
                    try:
                        s.update(333 if animated else 0)
                        t = s.tables()[0].contents
                        for i in range(t.n_surf):
                            r = t.surfaces[i]
                            if r.tag < 0:
                                continue
                            n += 1
                            assert (r.trm != 0) == (r.shift != 0), (config, seed, i)
                    finally:
                        s.destroy()
                        w.close()
    finally:
        dev.close()
    assert n > 1000


def _test_scene_lib():
    path = os.path.join(qr_oracle.REF_DIR, "libqrtestscenes.so")
    if not os.path.exists(path):
        pytest.skip("QuadRay's test scenes are not built (make -C oracle ref)")
    lib = C.CDLL(path)
    lib.qr_test_scene.restype = C.c_void_p
    lib.qr_test_scene.argtypes = [C.c_int]
    return lib


def test_update_stage_reproduces_stored_test_scene_tables():
    """The update stage (matrices on the host, each record's coefficients,
    clip box, bounds and culling inputs from fields.h) makes, byte for
    byte, the tables stored with QuadRay's 18 test scenes, whose frames the
    oracle reproduces bit for bit."""
    lib = _test_scene_lib()
    cases = golden_table_cases()
    assert len(cases) == 18
    dev = qrhip.Device(qrhip.HOST_ONLY)
    try:
        for name, c in sorted(cases.items()):
            n = int(c["scene"][len("scn_test"):])
            dev.set_fsaa(c["fsaa"])
            scn = qrhip.Scene(dev, lib.qr_test_scene(n), c["x_res"], c["y_res"])
            try:
                scn.set_depth(c["depth"])
                scn.update(0)
                got = qrhip.tables_snapshot(*scn.tables())
            finally:
                scn.destroy()
            with np.load(os.path.join(GOLDEN, name + ".npz")) as z:
                for k in got:
                    assert np.array_equal(got[k], z[k]), f"{name}: {k}"
    finally:
        dev.close()
