"""The oracle (CPU restatement) against frames written by QuadRay's own
renderer (tests/golden/make_golden.py), through this package's host update
stage: scene structs -> tables -> oracle -> frame, bit for bit."""
import numpy as np
import pytest

from helpers import HostScene, golden_cases, golden_frame, frame_stats

CASES = golden_cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_reference_frame(name):
    c = CASES[name]
    hs = HostScene(c["config"], c["seed"], c["x_res"], c["y_res"], c["fsaa"],
                   c["depth"], ids=c["ids"])
    try:
        frame, ids, flags, _, _ = hs.oracle(want_ids=c["ids"])
    finally:
        hs.close()
    want = golden_frame(name)
    exact, within, dmax = frame_stats(frame, want)
    assert exact == 1.0, f"exact {exact} within1 {within} max {dmax}"
    if c["ids"]:
        # id frames paint surface k side s with colour ((k << 1) | s) + 1
        decoded = want.astype(np.int64) - 1
        assert np.array_equal(ids[:, :, 0], decoded)
        assert flags is not None


def test_golden_ids_cover_many_surfaces():
    """The id fixtures are not degenerate: many surfaces and both sides."""
    got = golden_frame("whitted_s4_160x90_ids").astype(np.int64) - 1
    ids = np.unique(got[got >= 0])
    assert len(np.unique(ids >> 1)) >= 20
    assert {0, 1} <= set((ids & 1).tolist())
