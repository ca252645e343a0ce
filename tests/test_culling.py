"""The culling program is conservative: every hit the oracle accepts lies in
the culling box of its surface (world box, or the surface's own box in its
tracing frame) grown by the per-ray allowance the renderer applies, so
skipping a box can never change a result -- including the false roots that
fp32 cancellation gives very long rays (shadow rays from far ground hits).
Also checks the programs' structure (nested ranges; list order for the
nearest-hit program; whole groups in list order for the any-hit program)."""
import pytest

import qr_oracle
from helpers import HostScene, surface_boxes

FUZZ = 0.003    # QR_CULL_FUZZ, include/qr_soa.h


@pytest.mark.parametrize("config,seed,x,y,depth", [
    (1, 1, 200, 120, 4), (2, 1, 240, 136, 4), (2, 2, 240, 136, 10),
    (3, 1, 160, 90, 8), (4, 3, 160, 90, 4), (5, 4, 160, 90, 4), (6, 5, 160, 90, 6),
    # custom clippers whose kept region is unbounded (bbox adjustment)
    (7, 1000, 362, 214, 3), (7, 1002, 320, 180, 4), (9, 1, 161, 91, 6),
    (10, 1, 160, 90, 6),
])
def test_accepted_hits_inside_boxes(config, seed, x, y, depth):
    hs = HostScene(config, seed, x, y, 0, depth)
    try:
        audit = {}
        _, _, _, ops, _ = qr_oracle.render(hs.tables, hs.fd,
                                           boxes=surface_boxes(hs.tables),
                                           audit=audit)
    finally:
        hs.close()
    assert audit["checked"] > 1000
    assert audit["outside"] == 0
    # far rays' false roots stay inside the per-ray allowance, with margin
    assert audit["excess"] < 0.75 * FUZZ
    assert audit["ops_culled"] > 0


def program_elements(n_step, steps):
    """Elements of a culling program in visiting order; checks that box ranges
    nest and element steps carry skip -1 or -2."""
    elems, stack = [], []
    for p in range(n_step):
        while stack and stack[-1] == p:
            stack.pop()
        st = steps[p]
        if st.elem >= 0:
            elems.append(st.elem)
            assert st.skip in (-1, -2)
        else:
            assert st.elem == -1 and p < st.skip <= n_step
            assert not stack or st.skip <= stack[-1]   # ranges nest
            stack.append(st.skip)
    return elems


@pytest.mark.parametrize("config", [1, 2, 3, 4, 5])
def test_program_structure(config):
    import ctypes as C
    import numpy as np
    hs = HostScene(config, 9, 64, 36, 0, 4)
    try:
        t = hs.tables.contents
        # nearest-hit walks: every element, in list order
        assert program_elements(t.n_step, t.steps) == list(range(t.n_elem))
        # any-hit walks: every element once; a transform-node group whole and
        # in list order (its members' transforms come from the node's cache)
        seq = program_elements(t.n_step_any, t.steps_any)
        assert sorted(seq) == list(range(t.n_elem))
        el = np.ctypeslib.as_array(C.cast(t.elems, C.POINTER(C.c_int32)),
                                   shape=(t.n_elem * 2,)).reshape(t.n_elem, 2)
        pos = {e: k for k, e in enumerate(seq)}
        groups = 0
        for e in range(t.n_elem):
            last = el[e, 1]
            if last > e:
                groups += 1
                assert [seq[pos[e] + k] for k in range(last - e + 1)] == \
                    list(range(e, last + 1))
        if config in (2, 4, 5):     # the nested-array scenes have groups
            assert groups > 0
    finally:
        hs.close()


def test_test_scene_tables_boxes():
    """QuadRay's own test scenes (stored tables, tests/golden/scn_test*):
    every accepted hit lies in its culling box (custom clippers, arrays,
    every surface kind)."""
    from helpers import GoldenTables, golden_table_cases
    checked = 0
    for name in sorted(golden_table_cases()):
        g = GoldenTables(name)
        audit = {}
        qr_oracle.render(g.ptr, g.fd, boxes=surface_boxes(g.ptr), audit=audit)
        assert audit["outside"] == 0, name
        assert audit["excess"] < 0.75 * FUZZ, name
        checked += audit["checked"]
    assert checked > 1000000
