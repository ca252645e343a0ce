"""The update stage's device side through the C ABI: after the first build,
an update of an animated scene runs the transforms on the host and the rest
on the device -- every record, clip box, plane texture transform and culling
input (update.hip's records kernel over fields.h), then the programs' refit.
The device's records, materials and programs equal the host's own
derivation bit for bit, and every frame equals the frame of a fresh scene
built at that time -- for a scene rendered alone and for requests served at
many times."""
import numpy as np
import pytest

import qrhip

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("config,seed,x,y,fsaa", [(2, 11, 160, 90, 2), (8, 7, 128, 72, 0),
                                                   (4, 12, 96, 64, 1), (7, 13, 96, 64, 0),
                                                   (9, 14, 96, 64, 1)])
def test_gpu_refit_frames_equal_fresh_scenes(device, config, seed, x, y, fsaa):
    print(f"case refit config {config} seed {seed}", flush=True)
    times = (0, 600, 1300, 2900)
    device.set_fsaa(fsaa)
    w = qrhip.Workload(config, seed, animated=True)
    s = qrhip.Scene(device, w.ptr, x, y)
    frames = {}
    try:
        s.set_depth(4)
        for t in times:
            frames[t] = s.render(t)
            builds, refits, mismatch = s.program_stats(compare=True)
            assert mismatch == 0, f"time {t}: {mismatch} steps, records or materials differ"
    finally:
        s.destroy()
        w.close()
    assert builds == 1 and refits == len(times) - 1
    assert not np.array_equal(frames[times[0]], frames[times[-1]])   # it moves
    for t in times[1:]:
        device.set_fsaa(fsaa)
        w = qrhip.Workload(config, seed, animated=True)
        s = qrhip.Scene(device, w.ptr, x, y)
        try:
            s.set_depth(4)
            fresh = s.render(t)
        finally:
            s.destroy()
            w.close()
        assert np.array_equal(frames[t], fresh), f"time {t}"


def test_gpu_serve_animated_scenes(device):
    """Requests at different times of animated scenes: each served frame is
    the frame of a fresh scene rendered at its time."""
    device.set_fsaa(0)
    specs = [(7, 31), (2, 32)]
    works = [qrhip.Workload(c, sd, animated=True) for c, sd in specs]
    scenes = [qrhip.Scene(device, wk.ptr, 64, 64) for wk in works]
    for sc in scenes:
        sc.update(0)
    reqs = [qrhip.Server.request(k, k % 2, 80 + 8 * k, 48, k % 3, 3) for k in range(8)]
    for k, r in enumerate(reqs):
        r.time = 250 * k
    srv = qrhip.Server(device, scenes, 8 << 30, 16, 12288)
    got, results = {}, []
    try:
        for r in reqs:
            got[r.tag] = srv.submit(r)
        import time
        deadline = time.time() + 60
        while len(results) < len(reqs) and time.time() < deadline:
            results += srv.poll(wait_ms=200)
    finally:
        srv.destroy()
        builds = [sc.program_stats()[0] for sc in scenes]
        for sc in scenes:
            sc.destroy()
        for wk in works:
            wk.close()
    assert len(results) == len(reqs) and all(res["status"] == 0 for res in results)
    assert builds == [1, 1]          # programs built once, refitted per time
    for r in reqs:
        config, seed = specs[r.scene]
        device.set_fsaa(r.fsaa)
        w = qrhip.Workload(config, seed, animated=True)
        s = qrhip.Scene(device, w.ptr, r.x_res, r.y_res)
        try:
            s.set_depth(r.depth)
            want = s.render(r.time)
        finally:
            s.destroy()
            w.close()
        assert np.array_equal(got[r.tag], want), f"request {r.tag} at time {r.time}"
