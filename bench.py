#!/usr/bin/env python3
"""Throughput of the HIP Whitted tracer on the 1080p 64-quadric scene.

One step = one full frame of the workload traced on the GPU: 1920x1080, 4X
antialiasing (4 camera samples per pixel), reflection/refraction depth 4 with
Fresnel, 64 quadric surfaces plus a ground plane, 4 point lights, seeded
synthetic scene.  The scene tables and the frame stay in HBM for the timed
region (update stage and upload happen before it).  Mrays/s counts camera
samples (W*H*4) per second, the figure QuadRay's CPU renderer is timed on
beside it.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""
import argparse
import csv
import json
import os
import shutil
import subprocess
import sys
import tempfile
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "quadray-engine_amd", "python"))

import numpy as np  # noqa: E402

import qrhip  # noqa: E402

FP32_PEAK_TFLOPS = 157.3     # MI355X vector FP32 (packed), published
X_RES, Y_RES, FSAA, DEPTH, CONFIG = 1920, 1080, 2, 4, 2
SAMPLES = X_RES * Y_RES * 4


def host_cpus():
    """CPUs this process may run on: the affinity mask, bounded by the cgroup
    CPU quota when one is set; with the affinity size, the quota (None when
    unlimited) and the CPU model string."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = float(q) / float(per)
    except (OSError, ValueError):
        try:
            with open("/sys/fs/cgroup/cpu/cpu.cfs_quota_us") as f:
                q = float(f.read())
            with open("/sys/fs/cgroup/cpu/cpu.cfs_period_us") as f:
                per = float(f.read())
            if q > 0:
                quota = q / per
        except (OSError, ValueError):
            pass
    n = aff if quota is None else max(1, min(aff, int(quota)))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return n, aff, quota, model


def host_threads():
    return host_cpus()[0]


def simd_name(simd):
    """QuadRay's SIMD target word (simd_init: native size in 128-bit quads,
    sub-type, size factor) as its name, e.g. 512x2v2."""
    n, t, k = simd & 0xFF, (simd >> 8) & 0xFF, (simd >> 16) & 0xFF
    return f"{n * 128}x{k}v{t}"


def oracle_leg(tables, fd, threads):
    """The oracle's frame, its counted FP32 ops (QuadRay's whole-list walk,
    and the same frame with every ray skipping the surfaces its own box
    tests rule out, box tests included), rays, and wall time (the CPU-port
    baseline when QuadRay's own build is absent)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import qr_oracle
    acct = {}
    t0 = time.perf_counter()
    frame, _, _, ops, rays = qr_oracle.render(tables, fd, threads=threads,
                                              audit=acct)
    return frame, ops, acct["ops_culled"], rays, time.perf_counter() - t0


def reference_leg(seed, threads, frames, warmup=3, simd=None, render_only=True):
    """QuadRay's fp32 renderer built from its sources (oracle/_ref), with its
    own pruning/tiling on, timed on this host's cores: warm-up frames, then
    the median of the timed ones.  simd "n,t,k" forces a SIMD target
    (simd_init); render_only skips the update stage after the first frame
    (RT_OPTS_UPDATE_EXT0), as the GPU's timed launches do."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_f32")
    if not os.path.exists(exe):
        return None
    cmd = [exe, str(CONFIG), str(seed), "-t", str(threads), "-f", str(frames),
           "-w", str(warmup), "-p"]
    if render_only:
        cmd.append("-r")
    if simd is not None:
        cmd += ["-s", simd]
    res = subprocess.run(cmd, check=True, capture_output=True, text=True,
                         timeout=600)
    return json.loads(res.stdout.strip().splitlines()[-1])


def copy_probe(mib):
    """A plain device-to-device copy of known size (the HIP runtime's copy
    kernel): its FETCH_SIZE / WRITE_SIZE readings per dispatch give the factor
    that turns those counters into HBM bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nbytes = mib * (1 << 20)
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(src, 1, C.c_size_t(nbytes)) != 0:
        raise RuntimeError("hipMemset failed")
    for _ in range(3):
        if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) != 0:  # DeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(f"copy probe: {mib} MiB read + {mib} MiB written per dispatch", flush=True)


PROBE_MIB = 256      # device-to-device copy of known size per counter run
PROBE_COPIES = 3


def pmc_traffic(seed, path):
    """HBM bytes per frame: two rocprofv3 counter runs of this script in
    --pmc-child mode (FETCH_SIZE, then WRITE_SIZE), each summed over the
    frame's kernels and scaled by what the same counter reads, in the same
    run, for device-to-device copies of known size.  Runs before this
    process touches the GPU.  Returns (traffic dict, None) or (None, reason)."""
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None, "rocprofv3 not found"
    out = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="qr_pmc_", dir="/tmp")
        try:
            cmd = ["timeout", "-s", "KILL", "120", exe, "--pmc", ctr,
                   "--output-format", "csv", "-d", d, "-o", "p", "--",
                   sys.executable, os.path.abspath(__file__), "--pmc-child",
                   "--steps", "2", "--warmup", "1", "--seed", str(seed),
                   "--path", path]
            env = dict(os.environ, TMPDIR="/tmp")
            r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True,
                               text=True, timeout=150)
            if r.returncode != 0:
                return None, f"rocprofv3 --pmc {ctr} exited {r.returncode}"
            rows = []
            for root, _, files in os.walk(d):
                for f in files:
                    if f.endswith("counter_collection.csv"):
                        with open(os.path.join(root, f)) as fh:
                            rows += list(csv.DictReader(fh))
        finally:
            shutil.rmtree(d, ignore_errors=True)
        copies = [r for r in rows if "copyBuffer" in r["Kernel_Name"]]
        if not copies:
            return None, f"no copy dispatches in the {ctr} run"
        big = max(int(r["Grid_Size"]) for r in copies)
        probe = [float(r["Counter_Value"]) for r in copies if int(r["Grid_Size"]) == big]
        frames = sum(1 for r in rows if "k_resolve" in r["Kernel_Name"])
        kern = sum(float(r["Counter_Value"]) for r in rows
                   if "(anonymous namespace)::k_" in r["Kernel_Name"])
        if len(probe) != PROBE_COPIES or frames == 0 or sum(probe) <= 0:
            return None, f"unexpected dispatches in the {ctr} run"
        moved_kib = PROBE_COPIES * PROBE_MIB * 1024.0     # each copy reads and writes it
        factor = moved_kib / sum(probe)
        out[ctr] = {"kib_per_frame_raw": kern / frames, "scale": round(factor, 4),
                    "bytes_per_frame": kern / frames * factor * 1024.0}
    total = out["FETCH_SIZE"]["bytes_per_frame"] + out["WRITE_SIZE"]["bytes_per_frame"]
    return {"bytes": total, "fetch": out["FETCH_SIZE"], "write": out["WRITE_SIZE"]}, None


BATCH_FRAMES, BATCH_SCENES = 256, 16


def batched_leg(dev, seed):
    """256 independent frames (16 seeded scenes of 8 to 512 quadrics,
    resolutions log-uniform from 64x64 to 1920x1080, AA 1x/2x/4x and depth
    1-4 at random) traced as one batch: one launch per pipeline stage,
    replayed from a captured HIP graph, launch shape chosen by the
    autotuner; against the same 256 frames launched one after another."""
    rng = np.random.default_rng(seed)
    works = [qrhip.Workload(7, 1000 + k) for k in range(BATCH_SCENES)]
    scenes = []
    for w in works:
        s = qrhip.Scene(dev, w.ptr, 64, 64)
        s.update(0)
        scenes.append(s)
    reqs, samples, quadrics = [], 0, []
    for _ in range(BATCH_FRAMES):
        t = rng.uniform()
        x = int(round(64 * (1920 / 64) ** t))
        y = int(round(64 * (1080 / 64) ** t))
        fsaa = int(rng.integers(3))
        reqs.append((scenes[int(rng.integers(BATCH_SCENES))], x, y, fsaa,
                     int(rng.integers(1, 5))))
        samples += x * y * (1, 2, 4)[fsaa]
    for s in scenes:
        t, _ = s.tables()
        quadrics.append(t.contents.n_user - 1)
    b = qrhip.Batch(dev)
    try:
        b.set(reqs)
        b.time(1)                                   # queues settle
        (grid, tile, rpt, cadence), tune_ms = b.tune(reps=1)
        graph_ms = b.time(3, qrhip.BATCH_GRAPH)
        direct_ms = b.time(3, qrhip.BATCH_DIRECT)
        serial_ms = b.time(1, qrhip.BATCH_SERIAL)
    finally:
        b.destroy()
        for s in scenes:
            s.destroy()
        for w in works:
            w.close()
    mr = lambda ms: round(samples / (ms * 1e-3) * 1e-6, 3)
    return {
        "workload": f"{BATCH_FRAMES} frames of {BATCH_SCENES} seeded scenes "
                    f"({min(quadrics)}-{max(quadrics)} quadrics), 64x64..1920x1080 "
                    "log-uniform, AA 1x/2x/4x and depth 1-4 at random, seed "
                    f"{seed}",
        "camera_samples": samples,
        "tuned": {"workgroups_per_cu": grid, "tile_width": tile, "rays_per_thread": rpt,
                  "compaction_cadence": cadence},
        "graph_ms": round(graph_ms, 3), "direct_ms": round(direct_ms, 3),
        "serial_ms": round(serial_ms, 3),
        "graph_mrays_s": mr(graph_ms), "serial_mrays_s": mr(serial_ms),
        "speedup_vs_serial": round(serial_ms / graph_ms, 3),
    }


PT_X, PT_Y, PT_SPP, PT_DEPTH, PT_CONFIG = 1280, 720, 64, 8, 3


def path_tracer_leg(dev, seed, threads, frames=3, cpu=True):
    """The path-traced box (6 walls with an emissive patch, 16 diffuse,
    mirror and glass surfaces) at 1280x720, 64 samples per pixel, depth 8:
    one frame is one rt_Scene render with set_pton(64), its 64 passes traced
    as a wavefront of per-bounce queues; frames timed with HIP events on the
    scene's stream.  The op count comes from the oracle's culled count of
    one pass at the same size, times 64.  QuadRay's own path tracer, its
    pruning on, is timed on this host's cores over a few passes."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import qr_oracle
    L = qrhip.lib()
    dev.set_fsaa(0)
    w = qrhip.Workload(PT_CONFIG, seed)
    s = qrhip.Scene(dev, w.ptr, PT_X, PT_Y)
    try:
        s.set_depth(PT_DEPTH)
        s.set_pton(PT_SPP)
        s.update(0)
        qrhip.check(L.qrhip_scene_launch(s.h))      # warm-up: queues grow
        qrhip.check(L.qrhip_scene_sync(s.h))
        ms = s.time_launches(frames)
        tables, fd = s.tables()
        fd.pt_on = 1
        fd.pts_c = 1.0
        slots = PT_X * PT_Y
        planes = np.zeros((3, slots), dtype=np.float32)
        seeds = qrhip.pt_seed_plane(PT_X, PT_Y, PT_X, 0)
        acct = {}
        t0 = time.perf_counter()
        qr_oracle.render_pt(tables, fd, planes, seeds, threads=threads, audit=acct)
        o_secs = time.perf_counter() - t0
    finally:
        s.destroy()
        w.close()
    samples = PT_X * PT_Y * PT_SPP
    ops = acct["ops_culled"] * PT_SPP
    achieved = ops / (ms * 1e-3) / 1e12
    out = {"workload": f"path-traced box, config {PT_CONFIG} seed {seed}: {PT_X}x{PT_Y}, "
                       f"{PT_SPP} spp, depth {PT_DEPTH}, Russian roulette",
           "ms_per_frame": round(ms, 3),
           "mrays_s": round(samples / (ms * 1e-3) * 1e-6, 3),
           "total_rays_per_s_M": round(acct["rays"] * PT_SPP / (ms * 1e-3) * 1e-6, 3),
           "roofline": {"bound": "valu", "achieved": round(achieved, 3),
                        "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                        "ops_per_frame": int(ops),
                        "ops_note": "the oracle's culled op count of one pass at "
                                    f"{PT_X}x{PT_Y}, times {PT_SPP}"},
           "cpu_baseline": None}
    if cpu:
        exe = os.path.join(ROOT, "oracle", "_ref", "ref_f32")
        if os.path.exists(exe):
            cmd = [exe, str(PT_CONFIG), str(seed), "-x", str(PT_X), "-y", str(PT_Y),
                   "-a", "0", "-d", str(PT_DEPTH), "-q", "6", "-w", "2", "-f", "1",
                   "-t", str(threads), "-p"]
            res = subprocess.run(cmd, check=True, capture_output=True, text=True,
                                 timeout=600)
            ref = json.loads(res.stdout.strip().splitlines()[-1])
            out["cpu_baseline"] = {
                "value": round(ref["msamples_per_s"], 3), "unit": "Mrays/s",
                "cores": threads, "kind": "reference",
                "sample": f"{ref['frames']} path-tracer passes of the same {PT_X}x{PT_Y} "
                          f"scene (1 sample per pixel each) after 2 warm-up passes, "
                          f"median pass {ref['median_s']:.3f} s; QuadRay fp32 with its "
                          f"SIMD buffers, pruning on, {threads} threads",
                "gpu_over_cpu": round(samples / (ms * 1e-3) * 1e-6 / ref["msamples_per_s"], 1)}
        else:
            out["cpu_baseline"] = {
                "value": round(slots / o_secs * 1e-6, 3), "unit": "Mrays/s",
                "cores": threads, "kind": "port",
                "sample": "one pass of the same frame by the scalar C oracle"}
    return out


SERVE_SCENES_WHITTED, SERVE_SCENES_PT = 28, 4
SERVE_RES = [(640, 360), (1280, 720), (1920, 1080)]


def serving_requests(rng, scenes, poses, n):
    """Requests: a scene (Whitted scenes of 16-1024 quadrics, or a path-traced
    box), a camera turned by a random yaw about the scene's centre, an
    animation time, 360p / 720p / 1080p; Whitted depth 4 with random AA, or
    16 path-traced samples per pixel."""
    out = []
    for k in range(n):
        pt = rng.uniform() < 0.15
        si = SERVE_SCENES_WHITTED + int(rng.integers(SERVE_SCENES_PT)) if pt \
            else int(rng.integers(SERVE_SCENES_WHITTED))
        x, y = SERVE_RES[int(rng.integers(3))]
        a = float(rng.uniform(-0.5, 0.5))
        c, s_ = np.cos(a), np.sin(a)
        rz = np.array([[c, -s_, 0.0], [s_, c, 0.0], [0.0, 0.0, 1.0]])
        pose = (poses[si].astype(np.float64).reshape(4, 3) @ rz.T).reshape(12)
        r = qrhip.Server.request(k, si, x, y, 0 if pt else int(rng.integers(3)), 4,
                                 spp=16 if pt else 0, time=int(rng.integers(4)) * 40,
                                 pose=pose)
        out.append(r)
    return out


SERVE_STEP_SAMPLES = 1 << 24      # cap of camera samples a serving step takes
SERVE_STEP_MS = 2.0               # predicted device time a serving step takes
SERVE_REQUESTS = 1000             # requests per rate


def serve_run(srv, reqs, rate, seed=7):
    """Submit reqs to a running server as a Poisson stream of "rate" per
    second (all at once for rate None); latency = done - submit per request."""
    results = []
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(reqs)) if rate else np.zeros(len(reqs))

    def feed():
        t = time.perf_counter()
        for r, g in zip(reqs, gaps):
            t += g
            d = t - time.perf_counter()
            if d > 0:
                time.sleep(d)
            srv.submit(r)
    s0 = srv.stats()
    th = threading.Thread(target=feed)
    th.start()
    deadline = time.time() + 120
    while len(results) < len(reqs) and time.time() < deadline:
        results += srv.poll(wait_ms=100)
    th.join()
    s1 = srv.stats()
    srv.frames = {}
    if len(results) < len(reqs) or any(r["status"] != 0 for r in results):
        raise RuntimeError("serving run did not complete")
    kind = {r.tag: (r.y_res, r.spp > 0) for r in reqs}
    lat = np.array([r["done_ms"] - r["submit_ms"] for r in results])
    lat720 = np.array([r["done_ms"] - r["submit_ms"] for r in results
                       if kind[r["tag"]] == (720, False)])
    span_ms = max(r["done_ms"] for r in results) - min(r["submit_ms"] for r in results)
    pct = lambda a, q: round(float(np.percentile(a, q)), 2) if len(a) else None
    # device time charged to each kind of request, and how busy the device was
    share = {}
    for r in results:
        y, pt = kind[r["tag"]]
        k = f"{'pt' if pt else 'whitted'}_{y}p"
        share[k] = share.get(k, 0.0) + r["device_ms"]
    dev_ms = (s1["device_ms_whitted"] - s0["device_ms_whitted"]) + \
             (s1["device_ms_pt"] - s0["device_ms_pt"])
    total = sum(share.values())
    return {"rate_rps": None if rate is None else round(rate, 2),
            "requests": len(results),
            "frames_s": round(len(results) / (span_ms * 1e-3), 2),
            "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
            "p50_720p_whitted_ms": pct(lat720, 50), "p99_720p_whitted_ms": pct(lat720, 99),
            "n_720p_whitted": int(len(lat720)),
            "steps": int(s1["steps"] - s0["steps"]),
            "step_frames_max": int(s1["step_frames_max"]),
            "device_busy": round(dev_ms / span_ms, 3) if span_ms > 0 else None,
            "device_ms_per_frame": round(dev_ms / len(results), 3),
            "device_share": {k: round(v / total, 3) for k, v in sorted(share.items())}
            if total > 0 else None}


def update_leg(dev, seed):
    """The update stage of an animated 973-quadric scene (arrays turning and
    moving with time) at 1280x720: a scene's first update, which builds its
    culling programs on the host, against later updates at new times, which
    refit the programs' boxes on the device; wall time of the update call and
    the device work it enqueues (synchronised)."""
    L = qrhip.lib()

    def timed_update(s, t):
        a = time.perf_counter()
        s.update(t)
        qrhip.check(L.qrhip_scene_sync(s.h))
        return (time.perf_counter() - a) * 1e3

    build = []
    for k in range(5):
        w = qrhip.Workload(8, 7, animated=True)
        s = qrhip.Scene(dev, w.ptr, 1280, 720)
        try:
            build.append(timed_update(s, 100 * k))
        finally:
            s.destroy()
            w.close()
    w = qrhip.Workload(8, 7, animated=True)
    s = qrhip.Scene(dev, w.ptr, 1280, 720)
    try:
        timed_update(s, 0)
        refit = [timed_update(s, 40 * (k + 1)) for k in range(20)]
        builds, refits, mismatch = s.program_stats(compare=True)
        t = s.tables()[0].contents
        n_surf, n_step = t.n_surf, t.n_step + t.n_step_any
    finally:
        s.destroy()
        w.close()
    return {"workload": "config 8 seed 7 animated: 973 surface records, 1280x720",
            "surfaces": n_surf, "program_steps": n_step,
            "build_ms": round(float(np.median(build)), 3),
            "refit_ms": round(float(np.median(refit)), 3),
            "builds": builds, "device_refits": refits, "program_mismatch": mismatch}


def serving_leg(dev, seed, step_samples=SERVE_STEP_SAMPLES, step_ms=SERVE_STEP_MS):
    """Frames/s and p50/p99 latency of the serving loop on a Poisson stream
    at swept rates, against the same server tracing one whole frame at a time
    at saturation.  Each server is warmed up (buffers at their working sizes,
    per-scene cost rates learnt) by one stream of requests before it is
    measured."""
    rng = np.random.default_rng(seed)
    works = [qrhip.Workload(8, 2000 + k, animated=True) for k in range(SERVE_SCENES_WHITTED)]
    works += [qrhip.Workload(3, 3000 + k) for k in range(SERVE_SCENES_PT)]
    scenes, servers = [], []
    n_req = SERVE_REQUESTS
    try:
        for w in works:
            s = qrhip.Scene(dev, w.ptr, 64, 64)
            s.update(0)
            scenes.append(s)
        poses = [s.pose() for s in scenes]
        reqs = lambda: serving_requests(rng, scenes, poses, n_req)
        plain_srv = qrhip.Server(dev, scenes, 96 << 30, 1, 0)
        servers.append(plain_srv)
        serve_run(plain_srv, reqs(), None)                       # warm-up
        plain = serve_run(plain_srv, reqs(), None)
        srv = qrhip.Server(dev, scenes, 96 << 30, 64, step_samples, step_ms)
        servers.append(srv)
        serve_run(srv, reqs(), None)                             # warm-up
        sweep = []
        for mult in (0.25, 0.5, 0.75, 1.0, 1.5, 2.0):
            sweep.append(serve_run(srv, reqs(), mult * plain["frames_s"]))
        ok = [r for r in sweep if r["p99_720p_whitted_ms"] is not None
              and r["p99_720p_whitted_ms"] <= 16.0]
        best = max(ok, key=lambda r: r["frames_s"]) if ok else None
    finally:
        for v in servers:
            v.destroy()
        for s in scenes:
            s.destroy()
        for w in works:
            w.close()
    return {
        "workload": f"{SERVE_SCENES_WHITTED} animated Whitted scenes of 16-1024 quadrics and "
                    f"{SERVE_SCENES_PT} path-traced boxes resident; requests 360p/720p/1080p, "
                    "random yaw, animation time, Whitted depth 4 with random AA or 16 spp, "
                    f"Poisson arrivals, {n_req} per rate after a warm-up stream, seed {seed}",
        "step_samples": step_samples,
        "step_ms": step_ms,
        "plain_one_frame_at_a_time": plain,
        "sweep": sweep,
        "best_at_p99_720p_16ms": best,
        "speedup_vs_plain": round(best["frames_s"] / plain["frames_s"], 3) if best else None,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-frames", type=int, default=20,
                    help="timed frames of QuadRay's CPU renderer (after 3 warm-up "
                         "frames) in the baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--path", choices=["levels", "onekernel"], default="levels",
                    help="tracing path: level-by-level queues or one kernel")
    ap.add_argument("--profile", action="store_true",
                    help="launches only (for rocprofv3 counter runs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="instead: copy MIB MiB device to device 3 times (the "
                         "known-size kernel that scales HBM byte counters)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 256-frame batch measurement")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the serving-loop measurement")
    ap.add_argument("--serve-step", type=int, default=SERVE_STEP_SAMPLES,
                    help="cap of camera samples a serving step takes")
    ap.add_argument("--serve-step-ms", type=float, default=SERVE_STEP_MS,
                    help="predicted device time a serving step takes (0: samples only)")
    ap.add_argument("--serve-only", action="store_true",
                    help="only the serving measurement")
    ap.add_argument("--no-update", action="store_true",
                    help="skip the update-stage measurement")
    ap.add_argument("--no-pt", action="store_true",
                    help="skip the path-tracer measurement")
    ap.add_argument("--pt-only", action="store_true",
                    help="only the path-tracer frames (for rocprofv3 runs)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter runs for roofline.traffic")
    ap.add_argument("--pmc-child", action="store_true",
                    help="(used under rocprofv3 --pmc) the copy probe, then launches only")
    args = ap.parse_args()

    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.pmc_child:
        copy_probe(PROBE_MIB)
        args.profile = True

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: frames shard across GPUs with no "
              "exchange, and this build measures the single-GPU kernel only",
              file=sys.stderr)
        sys.exit(2)

    if args.serve_only:
        dev = qrhip.Device(0)
        try:
            print(json.dumps({"serving": serving_leg(dev, args.seed, args.serve_step,
                                                     args.serve_step_ms)}), flush=True)
        finally:
            dev.close()
        return

    if args.pt_only:
        dev = qrhip.Device(0)
        try:
            print(json.dumps({"path_tracer": path_tracer_leg(dev, args.seed, host_threads(),
                                                              cpu=False)}), flush=True)
        finally:
            dev.close()
        return

    print(f"bench: workload whitted 1080p seed {args.seed}", file=sys.stderr, flush=True)
    traffic, traffic_note = None, "not measured (--no-traffic)"
    if not (args.profile or args.no_traffic):
        traffic, traffic_note = pmc_traffic(args.seed, args.path)
        print(f"bench: traffic counters done ({traffic_note or 'ok'})", file=sys.stderr,
              flush=True)
    dev = qrhip.Device(0)
    dev.set_fsaa(FSAA)
    dev.set_path(qrhip.PATH_LEVELS if args.path == "levels" else qrhip.PATH_ONE_KERNEL)
    work = qrhip.Workload(CONFIG, args.seed)
    scene = qrhip.Scene(dev, work.ptr, X_RES, Y_RES)
    scene.set_depth(DEPTH)
    L = qrhip.lib()

    # update stage + upload: tables and frame resident in HBM from here on
    scene.update(0)
    tables, fd = scene.tables()
    n_surf = tables.contents.n_surf

    for _ in range(args.warmup):
        qrhip.check(L.qrhip_scene_launch(scene.h))
    qrhip.check(L.qrhip_scene_sync(scene.h))

    # timed region: K launches on the scene's stream, HIP events around them
    t0 = time.perf_counter()
    kernel_ms = scene.time_launches(args.steps)
    qrhip.check(L.qrhip_scene_sync(scene.h))
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    print(f"bench: {args.steps} steps done, {ms_per_step:.3f} ms/step", file=sys.stderr,
          flush=True)

    if args.profile:
        scene.destroy()
        work.close()
        dev.close()
        return

    qrhip.check(L.qrhip_scene_download(scene.h))
    gpu_frame = scene.frame()

    # whole rt_Scene::render calls (update + upload + trace + copy to host)
    rc_t = []
    for k in range(3):
        a = time.perf_counter()
        scene.render(0)
        rc_t.append(time.perf_counter() - a)
    render_call_ms = sorted(rc_t)[1] * 1e3

    threads = host_threads()
    o_frame, ops_full, ops, rays, o_secs = oracle_leg(tables, fd, threads)
    d = np.abs(((gpu_frame[..., None].astype(np.int64) >> [16, 8, 0]) & 255)
               - ((o_frame[..., None].astype(np.int64) >> [16, 8, 0]) & 255)).max(-1)
    parity = {"exact": float((gpu_frame == o_frame).mean()),
              "within1": float((d <= 1).mean()), "max_diff": int(d.max())}
    print(f"bench: oracle leg done ({o_secs:.1f} s)", file=sys.stderr, flush=True)

    achieved = ops / (kernel_ms * 1e-3) / 1e12
    roofline = {"bound": "valu", "achieved": round(achieved, 3),
                "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                "traffic": int(traffic["bytes"]) if traffic else None,
                "ops_per_launch": int(ops), "ops_whole_list": int(ops_full),
                "kernel_ms": round(kernel_ms, 4)}
    if traffic:
        roofline["traffic_detail"] = {
            "unit": "bytes per frame (HBM, from rocprofv3 FETCH_SIZE / WRITE_SIZE)",
            "fetch": int(traffic["fetch"]["bytes_per_frame"]),
            "write": int(traffic["write"]["bytes_per_frame"]),
            "fetch_scale": traffic["fetch"]["scale"],
            "write_scale": traffic["write"]["scale"],
            "gb_per_s": round(traffic["bytes"] / (kernel_ms * 1e-3) * 1e-9, 1)}
    else:
        roofline["traffic_note"] = traffic_note

    cpu = None
    if not args.no_cpu_baseline:
        _, aff, quota, model = host_cpus()
        ref = reference_leg(args.seed, threads, args.cpu_frames)
        if ref is not None:
            e2e = reference_leg(args.seed, threads, args.cpu_frames, render_only=False)
            avx2 = reference_leg(args.seed, threads, args.cpu_frames, simd="2,0,1")
            cpu = {"value": round(ref["msamples_per_s"], 3), "unit": "Mrays/s",
                   "cores": threads, "kind": "reference",
                   "sample": f"{ref['frames']} frames of the same 1920x1080 4X depth-4 "
                             f"scene after 3 warm-up frames, median frame "
                             f"{ref['median_s']:.3f} s, render only (update stage on "
                             f"the first frame only, as the GPU's timed launches); "
                             f"QuadRay fp32 built from its sources, SIMD target "
                             f"{simd_name(ref['simd'])} (its own pick), its list "
                             f"pruning/tiling on, {threads} threads",
                   "cpu_model": model, "affinity_cpus": aff, "cgroup_cpu_quota": quota,
                   "simd": simd_name(ref["simd"]),
                   "end_to_end": {"value": round(e2e["msamples_per_s"], 3),
                                  "median_frame_s": e2e["median_s"],
                                  "note": "update stage on every frame"},
                   "avx2": {"value": round(avx2["msamples_per_s"], 3),
                            "median_frame_s": avx2["median_s"],
                            "simd": simd_name(avx2["simd"]),
                            "note": "forced simd_init(2,0,1)"},
                   "gpu_over_cpu": None}
        else:
            cpu = {"value": round(SAMPLES / o_secs * 1e-6, 3), "unit": "Mrays/s",
                   "cores": threads, "kind": "port",
                   "sample": "one frame of the same workload by the scalar C oracle"}
        print("bench: cpu baseline done", file=sys.stderr, flush=True)

    batched = None
    if not args.no_batch:
        batched = batched_leg(dev, args.seed)
        print("bench: batch leg done", file=sys.stderr, flush=True)

    path_tracer = None
    if not args.no_pt:
        path_tracer = path_tracer_leg(dev, args.seed, threads)
        print("bench: path tracer leg done", file=sys.stderr, flush=True)

    update = None
    if not args.no_update:
        update = update_leg(dev, args.seed)
        print("bench: update leg done", file=sys.stderr, flush=True)

    serving = None
    if not args.no_serve:
        serving = serving_leg(dev, args.seed, args.serve_step, args.serve_step_ms)
        print("bench: serving leg done", file=sys.stderr, flush=True)

    value = SAMPLES / (ms_per_step * 1e-3) * 1e-6
    if cpu is not None and cpu.get("kind") == "reference":
        cpu["gpu_over_cpu"] = round(value / cpu["value"], 1)
    out = {
        "metric": "Mrays/s at 1080p 64-quadric scene",
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic: seeded scene generator, seed {args.seed}",
        "config": {"workload": "whitted_1080p_64_quadrics", "x_res": X_RES,
                   "y_res": Y_RES, "aa": "4x", "depth": DEPTH, "surfaces": n_surf,
                   "lights": 4, "fresnel": True, "seed": args.seed,
                   "path": args.path},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "total_rays_per_s_M": round(rays / (ms_per_step * 1e-3) * 1e-6, 3),
        "render_call_ms": round(render_call_ms, 3),
        "parity_vs_oracle": parity,
        "batched": batched,
        "path_tracer": path_tracer,
        "update": update,
        "serving": serving,
    }
    scene.destroy()
    work.close()
    dev.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
