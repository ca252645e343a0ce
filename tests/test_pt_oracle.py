"""The oracle's path tracer against QuadRay's own (golden planes written by
its fp32 renderer, tests/golden/make_golden.py): same seed plane, same number
of passes.

QuadRay carries each path's random state with the ray, so most sample slots
reproduce bit for bit; the rest diverge where QuadRay's SIMD packets branch
as a whole (a packet whose lanes all meet total inner reflection skips the
path tracer's refraction / reflection pick), after which the two estimates
are independent.  The bar is the Monte-Carlo one: 8x8-pixel block means agree
within the standard error estimated from the per-pass samples."""
import numpy as np
import pytest

import qr_oracle
import qrhip
from helpers import HostScene, golden_cases, golden_frame, golden_planes, pt_passes

PT_CASES = golden_cases(path_tracer=True)


def block_z(a, b, var, passes, block=8):
    """z-scores of 8x8-pixel block mean differences, standard error from the
    per-slot pass-sample variance (two independent estimates)."""
    y, x = a.shape[:2]
    zs = []
    for by in range(0, y, block):
        for bx in range(0, x, block):
            sa = a[by:by + block, bx:bx + block].reshape(-1, 3).astype(np.float64)
            sb = b[by:by + block, bx:bx + block].reshape(-1, 3).astype(np.float64)
            sv = var[by:by + block, bx:bx + block].reshape(-1, 3)
            n = sa.shape[0]
            se = np.sqrt(2.0 * sv.sum(0) / passes) / n
            d = sa.mean(0) - sb.mean(0)
            ok = se > 0
            zs.extend((d[ok] / se[ok]).tolist())
            assert np.all(d[~ok] == 0.0)
    return np.array(zs)


@pytest.mark.parametrize("name", sorted(PT_CASES))
def test_oracle_pt_matches_reference(name):
    c = PT_CASES[name]
    hs = HostScene(c["config"], c["seed"], c["x_res"], c["y_res"], c["fsaa"],
                   c["depth"])
    try:
        seeds = qrhip.pt_seed_plane(c["x_res"], c["y_res"], c["x_row"], c["fsaa"])
        planes, var, frame = pt_passes(hs, c["passes"], seeds, qr_oracle.render_pt)
    finally:
        hs.close()
    ref = golden_planes(name)
    exact = float((planes == ref).all(-1).mean())
    assert exact > 0.9, exact
    z = block_z(planes, ref, var, c["passes"])
    assert np.abs(z).max() < 5.0, np.abs(z).max()
    assert np.sqrt((z * z).mean()) < 1.5
    same_px = float((frame == golden_frame(name)).mean())
    assert same_px > 0.9, same_px
    print(f"{name}: {exact:.4f} of slots bit-exact, block |z| max "
          f"{np.abs(z).max():.2f}, pixels equal {same_px:.4f}")


def test_pt_seed_plane_is_quadray_chain():
    """The product's seed plane equals QuadRay's reset_pseed chain taken at
    QuadRay's slot positions (row pitch, samples innermost)."""
    x, y, x_row = 40, 20, 64
    for fsaa, ns in ((0, 1), (2, 4)):
        chain = qr_oracle.pt_seeds(4 * x_row * y)
        idx = ((np.arange(y)[:, None, None] * x_row + np.arange(x)[None, :, None]) * ns
               + np.arange(ns)[None, None, :]).reshape(-1)
        assert np.array_equal(chain[idx], qrhip.pt_seed_plane(x, y, x_row, fsaa))


def test_pt_estimator_unbiased_against_long_run():
    """The oracle's own running mean after 8 passes lies within its standard
    error of a 64-pass run (checks the running-mean update, not QuadRay)."""
    hs = HostScene(3, 5, 32, 18, 0, 8)
    try:
        seeds = qrhip.pt_seed_plane(32, 18, 32, 0)
        p8, var8, _ = pt_passes(hs, 8, seeds.copy(), qr_oracle.render_pt)
        p64, _, _ = pt_passes(hs, 64, seeds.copy(), qr_oracle.render_pt)
    finally:
        hs.close()
    z = block_z(p8, p64, var8, 8)
    assert np.abs(z).max() < 6.0
