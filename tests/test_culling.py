"""The culling program is conservative: every hit the oracle accepts lies in
the culling box of its surface (world box, or the surface's own box in its
tracing frame) grown by the per-ray allowance the renderer applies, so
skipping a box can never change a result -- including the false roots that
fp32 cancellation gives very long rays (shadow rays from far ground hits).
Also checks the program's structure (list order, nested ranges)."""
import pytest

import qr_oracle
from helpers import HostScene, surface_boxes

FUZZ = 0.002    # QR_CULL_FUZZ, include/qr_soa.h


@pytest.mark.parametrize("config,seed,x,y,depth", [
    (1, 1, 200, 120, 4), (2, 1, 240, 136, 4), (2, 2, 240, 136, 10),
    (3, 1, 160, 90, 8), (4, 3, 160, 90, 4), (5, 4, 160, 90, 4), (6, 5, 160, 90, 6),
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
    # far rays' false roots stay well inside the per-ray allowance
    assert audit["excess"] < 0.5 * FUZZ
    assert audit["ops_culled"] > 0


@pytest.mark.parametrize("config", [1, 2, 3, 4])
def test_program_structure(config):
    hs = HostScene(config, 9, 64, 36, 0, 4)
    try:
        t = hs.tables.contents
        elems, stack = [], []
        for p in range(t.n_step):
            while stack and stack[-1] == p:
                stack.pop()
            st = t.steps[p]
            if st.elem >= 0:
                elems.append(st.elem)
                assert st.skip in (-1, -2)
            else:
                assert st.elem == -1 and p < st.skip <= t.n_step
                assert not stack or st.skip <= stack[-1]   # ranges nest
                stack.append(st.skip)
        assert elems == list(range(t.n_elem))            # every element, in order
    finally:
        hs.close()
