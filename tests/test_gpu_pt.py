"""HIP path tracer against the oracle's path tracer (same tables, seed plane
and passes: the same paths and arithmetic, so planes agree slot for slot),
and the rt_Scene path (set_pton, render) against QuadRay's golden frames.

The kernels trace a wavefront, bounce by bounce over compacted queues, and
fold each sample's contributions in the depth-first order the oracle (and
QuadRay) adds them, so slots are bit-identical; several passes in one launch
equal the same passes launched one by one; queues that overflow (glass at
depth 12 doubles the rays every bounce) are grown and the run traced
again."""
import numpy as np
import pytest

import qr_oracle
import qrhip
from helpers import (HostScene, assert_frame_parity, golden_cases, golden_frame,
                     pt_passes)

pytestmark = pytest.mark.gpu

PT_CASES = golden_cases(path_tracer=True)


@pytest.mark.parametrize("config,seed,x,y,fsaa,depth,passes", [
    (3, 7, 96, 54, 0, 8, 6),
    (3, 8, 64, 36, 2, 8, 4),
    (6, 9, 64, 36, 0, 5, 4),
    (2, 10, 64, 36, 1, 4, 3),
    (6, 11, 64, 36, 0, 12, 2),
    (10, 12, 96, 54, 2, 8, 3),
])
def test_gpu_pt_matches_oracle(device, config, seed, x, y, fsaa, depth, passes):
    print(f"case pt config {config} seed {seed} {x}x{y} fsaa {fsaa} depth {depth}",
          flush=True)
    hs = HostScene(config, seed, x, y, fsaa, depth)
    try:
        seeds = qrhip.pt_seed_plane(x, y, x, fsaa)
        want, _, fw = pt_passes(hs, passes, seeds.copy(), qr_oracle.render_pt)
        got, _, fg = pt_passes(hs, passes, seeds.copy(), device.render_tables_pt)
    finally:
        hs.close()
    exact = float((got == want).all(-1).mean())
    print(f"slots bit-exact: {exact:.6f}")
    assert exact == 1.0, exact
    assert np.array_equal(fg, fw), "pt frame"


@pytest.mark.parametrize("name", sorted(PT_CASES))
def test_gpu_pt_scene_matches_reference(device, name):
    """set_pton(1) and one render per pass into a frame with QuadRay's row
    pitch, so every sample slot starts from QuadRay's seed."""
    c = PT_CASES[name]
    print(f"case pt scene {name}", flush=True)
    device.set_fsaa(c["fsaa"])
    w = qrhip.Workload(c["config"], c["seed"])
    s = qrhip.Scene(device, w.ptr, c["x_res"], c["y_res"], x_row=c["x_row"])
    try:
        s.set_depth(c["depth"])
        assert s.set_pton(1) == 1
        for t in range(c["passes"]):
            frame = s.render(t)
    finally:
        s.destroy()
        w.close()
    same = float((frame == golden_frame(name)).mean())
    assert same > 0.9, same
    print(f"{name}: {same:.4f} of pixels equal QuadRay's")


@pytest.mark.parametrize("config,seed,x,y,fsaa,depth", [(3, 21, 160, 90, 0, 8),
                                                      (3, 22, 96, 54, 2, 8)])
def test_gpu_pt_passes_in_one_launch(device, config, seed, x, y, fsaa, depth):
    """set_pton(n): n passes traced in one launch (one wavefront of every
    pass's samples) equal n renders of one pass each, and the oracle's n
    passes, bit for bit."""
    print(f"case pt passes in one launch config {config} seed {seed}", flush=True)
    device.set_fsaa(fsaa)
    frames = []
    for per_call, calls in ((6, 1), (1, 6), (2, 3)):
        w = qrhip.Workload(config, seed)
        s = qrhip.Scene(device, w.ptr, x, y)
        try:
            s.set_depth(depth)
            s.set_pton(per_call)
            for t in range(calls):
                frame = s.render(0)
            frames.append(frame)
        finally:
            s.destroy()
            w.close()
    assert np.array_equal(frames[0], frames[1])
    assert np.array_equal(frames[0], frames[2])
    hs = HostScene(config, seed, x, y, fsaa, depth)
    try:
        seeds = qrhip.pt_seed_plane(x, y, x, fsaa)
        _, _, fw = pt_passes(hs, 6, seeds, qr_oracle.render_pt)
    finally:
        hs.close()
    assert np.array_equal(frames[0], fw)
