"""The serving loop through the C ABI (qrhip_server_*).

Every frame a server returns is the frame of that request alone: the same as
a server tracing only that request as a whole frame, and, for requests with
the scene's own camera, the same as the scene rendered by itself (Whitted) or
with set_pton (path tracer) -- also when the frame is traced in chunks of a
few tiles over many steps beside other requests.  Admission keeps the
requests in flight within the memory budget; requests arriving while others
run join later steps."""
import threading
import time

import numpy as np
import pytest

import qrhip
from helpers import oracle_rows

pytestmark = pytest.mark.gpu

MB = 1 << 20


class Scenes:
    def __init__(self, device, specs, fsaa=0):
        device.set_fsaa(fsaa)
        self.works = [qrhip.Workload(c, s) for c, s in specs]
        self.scenes = [qrhip.Scene(device, w.ptr, 64, 64) for w in self.works]
        for s in self.scenes:
            s.update(0)

    def close(self):
        for s in self.scenes:
            s.destroy()
        for w in self.works:
            w.close()


CHUNK = 12288       # camera samples per step: test frames take several steps


def serve(device, scenes, reqs, budget=8 << 30, max_active=64, step_samples=CHUNK,
          stagger_s=0.0, step_ms=0.0):
    """Submit every request (optionally spaced out from another thread) and
    return {tag: frame}, the results and the server's statistics."""
    srv = qrhip.Server(device, scenes, budget, max_active, step_samples, step_ms)
    results, frames = [], {}
    try:
        if stagger_s > 0:
            def feed():
                for r in reqs:
                    frames[r.tag] = srv.submit(r)
                    time.sleep(stagger_s)
            th = threading.Thread(target=feed)
            th.start()
        else:
            for r in reqs:
                frames[r.tag] = srv.submit(r)
        deadline = time.time() + 60
        while len(results) < len(reqs) and time.time() < deadline:
            results += srv.poll(wait_ms=200)
        if stagger_s > 0:
            th.join()
        stats = srv.stats()
    finally:
        srv.destroy()
    assert len(results) == len(reqs)
    assert all(r["status"] == 0 for r in results)
    return frames, results, stats


def yaw_pose(base, angle):
    """The scene's camera turned about the world z axis through the origin."""
    c, s = np.cos(angle), np.sin(angle)
    rz = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
    b = base.astype(np.float64).reshape(4, 3)
    return (b @ rz.T).astype(np.float32).reshape(12)


def mixed_requests(sc, rng, n_whitted, pt_scene, n_pt, tag0=0):
    reqs = []
    for k in range(n_whitted):
        si = int(rng.integers(len(sc.scenes) - 1))
        x, y = int(rng.integers(24, 120)), int(rng.integers(24, 100))
        pose = None
        if k % 3 == 2:
            pose = yaw_pose(sc.scenes[si].pose(), float(rng.uniform(-0.6, 0.6)))
        reqs.append(qrhip.Server.request(tag0 + k, si, x, y, int(rng.integers(3)),
                                         int(rng.integers(0, 5)), pose=pose))
    for k in range(n_pt):
        reqs.append(qrhip.Server.request(tag0 + n_whitted + k, pt_scene, 40 + 8 * k, 30,
                                         k % 2, 8, spp=3 + k))
    return reqs


def test_serve_results_same_alone_and_in_mix(device):
    rng = np.random.default_rng(3)
    specs = [(7, 301), (7, 302), (2, 303), (3, 304)]
    sc = Scenes(device, specs)
    try:
        reqs = mixed_requests(sc, rng, 10, 3, 2)
        got, results, stats = serve(device, sc.scenes, reqs)
        assert stats["step_frames_max"] > 1              # requests shared steps
        assert max(res["steps"] for res in results) > 1   # frames took several steps
        for r in reqs:
            alone, _, _ = serve(device, sc.scenes, [r], max_active=1, step_samples=0)
            assert np.array_equal(got[r.tag], alone[r.tag]), f"request {r.tag}"
    finally:
        sc.close()
    # with the scene's own camera, a served frame is the scene rendered alone
    for r in reqs:
        if r.use_pose:
            continue
        config, seed = specs[r.scene]
        device.set_fsaa(r.fsaa)
        w = qrhip.Workload(config, seed)
        s = qrhip.Scene(device, w.ptr, r.x_res, r.y_res)
        try:
            s.set_depth(r.depth)
            if r.spp:
                s.set_pton(r.spp)
            want = s.render(0)
        finally:
            s.destroy()
            w.close()
        assert np.array_equal(got[r.tag], want), f"request {r.tag} vs its scene"


def test_serve_admission_within_budget(device):
    rng = np.random.default_rng(4)
    sc = Scenes(device, [(7, 401), (7, 402)])
    try:
        probe = qrhip.Server(device, sc.scenes, 1 << 40, 8, CHUNK)
        reqs = [qrhip.Server.request(k, k % 2, 96, 64, 2, 4) for k in range(6)]
        need = max(probe.footprint(r) for r in reqs)
        big = qrhip.Server.request(99, 0, 4000, 4000, 2, 14)
        big_need = probe.footprint(big)
        probe.destroy()
        budget = int(need * 1.5)            # room for one request at a time
        assert big_need > budget
        got, _, stats = serve(device, sc.scenes, reqs, budget=budget)
        assert stats["step_frames_max"] == 1
        assert 0 < stats["bytes_peak"] <= budget
        free, _, _ = serve(device, sc.scenes, reqs)
        for r in reqs:
            assert np.array_equal(got[r.tag], free[r.tag])
        srv = qrhip.Server(device, sc.scenes, budget, 8, CHUNK)
        try:
            with pytest.raises(qrhip.QrError):
                srv.submit(big)
        finally:
            srv.destroy()
    finally:
        sc.close()


def test_serve_staggered_arrivals_and_times(device):
    """Requests arrive while others are traced, with animation times that
    differ per request (the update stage runs per scene and time)."""
    rng = np.random.default_rng(5)
    sc = Scenes(device, [(7, 501), (2, 502), (3, 503)])
    try:
        reqs = mixed_requests(sc, rng, 16, 2, 2)
        for k, r in enumerate(reqs):
            r.time = 0 if k % 4 else 1000 * k
        got, results, stats = serve(device, sc.scenes, reqs, stagger_s=0.002)
        ref, _, _ = serve(device, sc.scenes, reqs)
        for r in reqs:
            assert np.array_equal(got[r.tag], ref[r.tag]), f"request {r.tag}"
        for res in results:
            assert res["submit_ms"] <= res["admit_ms"] <= res["done_ms"]
    finally:
        sc.close()


def test_serve_time_budget_and_one_scene_at_many_times(device):
    """Steps bounded by predicted device time, requests served by deadline,
    several passes of a path-traced chunk in one run, and requests for one
    animated scene at different times in flight together (each with tables
    of its own): every frame equals the request served alone, and the
    device time charged to the requests adds up to the measured time."""
    print("case serving with a device-time budget", flush=True)
    rng = np.random.default_rng(12)
    works = [qrhip.Workload(8, 2100 + k, animated=True) for k in range(2)]
    works.append(qrhip.Workload(3, 3100))
    scenes = [qrhip.Scene(device, w.ptr, 64, 64) for w in works]
    for s in scenes:
        s.update(0)
    reqs = []
    for k in range(14):
        si = k % 2
        x, y = int(rng.integers(40, 200)), int(rng.integers(30, 160))
        reqs.append(qrhip.Server.request(k, si, x, y, int(rng.integers(3)), 4,
                                         time=40 * (k % 4)))
    for k in range(3):
        reqs.append(qrhip.Server.request(14 + k, 2, 96, 64, k % 2, 8, spp=5 + k))
    try:
        got, results, stats = serve(device, scenes, reqs, step_samples=1 << 24,
                                    step_ms=0.5)
        assert stats["step_frames_max"] > 1
        charged = sum(r["device_ms"] for r in results)
        measured = stats["device_ms_whitted"] + stats["device_ms_pt"]
        assert abs(charged - measured) <= 1e-3 * measured + 1e-6, (charged, measured)
        for r in reqs:
            one, _, _ = serve(device, scenes, [r], max_active=1, step_samples=0)
            assert np.array_equal(got[r.tag], one[r.tag]), f"request {r.tag}"
    finally:
        for s in scenes:
            s.destroy()
        for w in works:
            w.close()
    # each frame is its scene rendered by itself at the request's time
    for r in reqs:
        device.set_fsaa(r.fsaa)
        w = qrhip.Workload(8, 2100 + r.scene, animated=True) if r.scene < 2 \
            else qrhip.Workload(3, 3100)
        s = qrhip.Scene(device, w.ptr, r.x_res, r.y_res)
        try:
            s.set_depth(r.depth)
            if r.spp:
                s.set_pton(r.spp)
            want = s.render(r.time)
        finally:
            s.destroy()
            w.close()
        assert np.array_equal(got[r.tag], want), f"request {r.tag} vs its scene"


def test_serve_rejects_bad_requests(device):
    sc = Scenes(device, [(7, 601)])
    srv = qrhip.Server(device, sc.scenes, 1 << 30, 4, CHUNK)
    try:
        for bad in [qrhip.Server.request(1, 5, 32, 32), qrhip.Server.request(2, 0, 0, 32),
                    qrhip.Server.request(3, 0, 32, 32, fsaa=4),
                    qrhip.Server.request(4, 0, 32, 32, depth=15)]:
            with pytest.raises(qrhip.QrError):
                srv.submit(bad)
    finally:
        srv.destroy()
        sc.close()


def test_serve_at_benchmark_shape_against_oracle(device):
    """Requests of the benchmark's kind: animated Whitted scenes of 16 to
    1024 quadrics resident in HBM, 360p / 720p / 1080p with random AA at
    depth 4 and animation times, path-traced 16-spp requests, yawed cameras,
    arriving while others are traced, in steps of the benchmark's 8 M camera
    samples.  Frames with the scene's own camera equal the oracle on sampled
    rows; every frame equals the same request served alone."""
    rng = np.random.default_rng(8)
    specs = [(8, 2000 + k) for k in range(8)]
    works = [qrhip.Workload(c, sd, animated=True) for c, sd in specs]
    works.append(qrhip.Workload(3, 3000))
    scenes = [qrhip.Scene(device, w.ptr, 64, 64) for w in works]
    for s in scenes:
        s.update(0)
    res = [(640, 360), (1280, 720), (1920, 1080)]
    reqs, tag = [], 0
    for k in range(24):
        si = int(rng.integers(8))
        x, y = res[k % 3]
        pose = None
        if k % 6 == 5:
            pose = yaw_pose(scenes[si].pose(), float(rng.uniform(-0.5, 0.5)))
        reqs.append(qrhip.Server.request(tag, si, x, y, int(rng.integers(3)), 4,
                                         time=40 * int(rng.integers(4)), pose=pose))
        tag += 1
    for k in range(3):
        reqs.append(qrhip.Server.request(tag, 8, 640, 360, 0, 8, spp=16))
        tag += 1
    print(f"case serving {len(reqs)} requests at 360p/720p/1080p", flush=True)
    try:
        got, results, stats = serve(device, scenes, reqs, step_samples=1 << 23,
                                    stagger_s=0.002)
        assert stats["step_frames_max"] > 1
        for r in reqs:
            if r.use_pose or r.spp:
                one, _, _ = serve(device, scenes, [r], max_active=1, step_samples=0)
                assert np.array_equal(got[r.tag], one[r.tag]), f"request {r.tag}"
    finally:
        for s in scenes:
            s.destroy()
        for w in works:
            w.close()
    pick = np.random.default_rng(9)
    own = [r for r in reqs if not r.use_pose and not r.spp]
    cases = [(specs[r.scene][0], specs[r.scene][1], r.x_res, r.y_res, r.fsaa, r.depth,
              r.time, True, sorted(set(int(v) for v in pick.integers(0, r.y_res, 2))))
             for r in own]
    for r, rows in zip(own, oracle_rows(cases)):
        for row, px in rows.items():
            assert np.array_equal(got[r.tag][row], px), f"request {r.tag} row {row}"
    print(f"serving: {len(own)} frames equal the oracle on sampled rows", flush=True)
