"""HIP path tracer against the oracle's path tracer (same tables, seed plane
and passes: the same paths and arithmetic, so planes agree slot for slot),
and the rt_Scene path (set_pton, render) against QuadRay's golden frames."""
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
    assert exact > 0.999, exact
    assert_frame_parity(fg, fw, "pt frame")


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
