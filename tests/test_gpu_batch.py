"""Frame batches through the C ABI (qrhip_batch_*).

Every frame of a batch -- scenes of 8 to 512 quadrics, uneven resolutions,
AA modes and depths, traced by one launch per pipeline stage -- equals the
frame its scene renders alone at the same settings, bit for bit.  Replays of
the captured HIP graph (with eager work and synchronisation between them),
the serial mode, queue growth inside a batch and the autotuner change no
frame."""
import numpy as np
import pytest

import qrhip
from helpers import oracle_rows

pytestmark = pytest.mark.gpu


def alone(device, config, seed, x, y, fsaa, depth):
    """The frame a scene renders on its own (rt_Scene::render path)."""
    device.set_fsaa(fsaa)
    w = qrhip.Workload(config, seed)
    s = qrhip.Scene(device, w.ptr, x, y)
    try:
        s.set_depth(depth)
        return s.render(0)
    finally:
        s.destroy()
        w.close()


class BatchCase:
    """Scenes (config, seed) and frame requests over them."""

    def __init__(self, device, scenes, requests):
        self.device = device
        self.works, self.scenes = [], []
        for config, seed in scenes:
            w = qrhip.Workload(config, seed)
            s = qrhip.Scene(device, w.ptr, 64, 64)
            s.update(0)
            self.works.append(w)
            self.scenes.append(s)
        self.keys = [(scenes[k][0], scenes[k][1], x, y, f, d) for k, x, y, f, d in requests]
        self.batch = qrhip.Batch(device)
        self.batch.set([(self.scenes[k], x, y, f, d) for k, x, y, f, d in requests])

    def close(self):
        self.batch.destroy()
        for s in self.scenes:
            s.destroy()
        for w in self.works:
            w.close()


def random_requests(rng, n_scenes, n, lo=16, hi=160, depths=(0, 4)):
    out = []
    for _ in range(n):
        x = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        y = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        out.append((int(rng.integers(n_scenes)), x, y, int(rng.integers(3)),
                    int(rng.integers(depths[0], depths[1] + 1))))
    return out


@pytest.mark.parametrize("seed", [1, 2])
def test_batch_frames_equal_alone(device, seed):
    rng = np.random.default_rng(seed)
    scenes = [(7, 100 * seed + k) for k in range(4)]
    reqs = random_requests(rng, len(scenes), 12)
    print(f"batch seed {seed}: {reqs}", flush=True)
    bc = BatchCase(device, scenes, reqs)
    try:
        got = bc.batch.render(qrhip.BATCH_DIRECT)
    finally:
        bc.close()
    for key, frame in zip(bc.keys, got):
        want = alone(device, *key)
        assert frame.shape == want.shape
        assert np.array_equal(frame, want), f"frame {key} differs from its scene alone"


def test_batch_graph_replays_match_eager(device):
    rng = np.random.default_rng(7)
    scenes = [(7, 701), (7, 702), (2, 703)]
    bc = BatchCase(device, scenes, random_requests(rng, len(scenes), 9))
    try:
        eager = bc.batch.render(qrhip.BATCH_DIRECT)
        for rep in range(3):
            got = bc.batch.render(qrhip.BATCH_GRAPH)
            for k, (a, b) in enumerate(zip(got, eager)):
                assert np.array_equal(a, b), f"graph replay {rep} frame {k}"
            # eager work between replays
            again = bc.batch.render(qrhip.BATCH_DIRECT)
            for k, (a, b) in enumerate(zip(again, eager)):
                assert np.array_equal(a, b), f"eager run after replay {rep} frame {k}"
    finally:
        bc.close()


@pytest.mark.parametrize("depth", [0, 1, 4])
def test_batch_of_one_graph_and_direct(device, depth):
    """A batch of one frame launches no plan kernels (its kernels count their
    own work); outside capture two of its stages run on a second stream,
    under capture everything stays on one stream.  Every replay and every
    direct launch equals the scene rendered alone."""
    print(f"batch of one, depth {depth}", flush=True)
    bc = BatchCase(device, [(2, 41)], [(0, 200, 120, 2, depth)])
    try:
        want = alone(device, *bc.keys[0])
        for rep in range(3):
            got = bc.batch.render(qrhip.BATCH_GRAPH)
            assert np.array_equal(got[0], want), f"graph replay {rep}"
            got = bc.batch.render(qrhip.BATCH_DIRECT)
            assert np.array_equal(got[0], want), f"direct launch {rep}"
    finally:
        bc.close()


def test_batch_serial_equals_direct(device):
    rng = np.random.default_rng(11)
    scenes = [(7, 1101), (7, 1102)]
    bc = BatchCase(device, scenes, random_requests(rng, len(scenes), 6))
    try:
        direct = bc.batch.render(qrhip.BATCH_DIRECT)
        serial = bc.batch.render(qrhip.BATCH_SERIAL)
    finally:
        bc.close()
    for k, (a, b) in enumerate(zip(direct, serial)):
        assert np.array_equal(a, b), f"frame {k}"


def test_batch_queue_growth(device):
    """All-glass frames at depth 8 overflow the first queue sizes inside the
    batch; the batch is traced again with grown queues."""
    reqs = [(0, 96, 54, 0, 8), (1, 64, 64, 2, 3), (0, 40, 72, 1, 8)]
    bc = BatchCase(device, [(6, 21), (7, 22)], reqs)
    try:
        got = bc.batch.render(qrhip.BATCH_DIRECT)
        replay = bc.batch.render(qrhip.BATCH_GRAPH)
    finally:
        bc.close()
    for key, frame, rep in zip(bc.keys, got, replay):
        assert np.array_equal(frame, alone(device, *key)), f"frame {key}"
        assert np.array_equal(rep, frame), f"graph frame {key}"


def test_batch_tune_keeps_frames(device):
    rng = np.random.default_rng(5)
    scenes = [(7, 501), (7, 502)]
    bc = BatchCase(device, scenes, random_requests(rng, len(scenes), 5))
    try:
        before = bc.batch.render(qrhip.BATCH_DIRECT)
        (grid, tile, rpt, cadence), ms = bc.batch.tune(reps=1)
        assert grid in (8, 16, 32, 64) and tile in (4, 8, 16, 32) and ms > 0
        assert rpt in (1, 2, 4) and cadence in (1, 2)
        after = bc.batch.render(qrhip.BATCH_DIRECT)
        graph = bc.batch.render(qrhip.BATCH_GRAPH)
    finally:
        bc.close()
    for k, (a, b, c) in enumerate(zip(before, after, graph)):
        assert np.array_equal(a, b), f"frame {k} after tuning"
        assert np.array_equal(a, c), f"graph frame {k} after tuning"


def test_batch_tile_shapes_equal(device):
    """Every level-0 tile shape gives the same frames."""
    rng = np.random.default_rng(9)
    bc = BatchCase(device, [(7, 901)], random_requests(rng, 1, 4))
    try:
        ref = bc.batch.render()
        for tile in (4, 16, 32):
            bc.batch.config(4, tile)
            got = bc.batch.render()
            for k, (a, b) in enumerate(zip(got, ref)):
                assert np.array_equal(a, b), f"tile width {tile} frame {k}"
    finally:
        bc.close()


@pytest.mark.parametrize("rpt,cadence", [(2, 1), (4, 1), (1, 2), (2, 2), (8, 3), (1, 4)])
def test_batch_rays_per_thread_and_cadence_equal(device, rpt, cadence):
    """Rays per thread and the compaction cadence change no frame: direct,
    graph and serial launches against the default shape, with the queues
    of levels that are not compacted grown on the way (depth up to 4, one
    all-glass scene)."""
    print(f"case batch rpt {rpt} cadence {cadence}", flush=True)
    rng = np.random.default_rng(40 + 7 * rpt + cadence)
    scenes = [(7, 931), (6, 932)]
    bc = BatchCase(device, scenes, random_requests(rng, len(scenes), 6))
    try:
        ref = bc.batch.render()
        bc.batch.config(16, 8, rpt, cadence)
        got = bc.batch.render()
        graph = bc.batch.render(qrhip.BATCH_GRAPH)
        serial = bc.batch.render(qrhip.BATCH_SERIAL)
    finally:
        bc.close()
    for k, (a, b, c, d) in enumerate(zip(ref, got, graph, serial)):
        assert np.array_equal(a, b), f"frame {k}"
        assert np.array_equal(a, c), f"graph frame {k}"
        assert np.array_equal(a, d), f"serial frame {k}"


def test_batch_rejects_bad_requests(device):
    w = qrhip.Workload(7, 3)
    s = qrhip.Scene(device, w.ptr, 32, 32)
    s.update(0)
    b = qrhip.Batch(device)
    try:
        for bad in [(s, 0, 10, 0, 1), (s, 10, 10, 3, 1), (s, 10, 10, 0, 15),
                    (s, 20000, 10, 0, 1)]:
            with pytest.raises(qrhip.QrError):
                b.set([bad])
        with pytest.raises(qrhip.QrError):
            b.config(0, 8)
        with pytest.raises(qrhip.QrError):
            b.config(4, 12)
    finally:
        b.destroy()
        s.destroy()
        w.close()


def test_batch_of_300_frames(device):
    """A large batch: 300 small frames, a sample of which must equal their
    scenes alone (batches plan their stages with the plan kernel; a frame
    alone plans inside the stage kernels)."""
    rng = np.random.default_rng(13)
    scenes = [(7, 1300 + k) for k in range(3)]
    reqs = random_requests(rng, len(scenes), 300, lo=8, hi=24, depths=(1, 3))
    bc = BatchCase(device, scenes, reqs)
    try:
        got = bc.batch.render(qrhip.BATCH_DIRECT)
        graph = bc.batch.render(qrhip.BATCH_GRAPH)
        serial = bc.batch.render(qrhip.BATCH_SERIAL)
        # every frame's batch of one overlaps its stages on the one second
        # stream the device holds, not on one stream per frame
        assert qrhip.lib().qrhip_side_streams() == 1
    finally:
        bc.close()
    for k in range(0, 300, 37):
        want = alone(device, *bc.keys[k])
        assert np.array_equal(got[k], want), f"frame {k}"
        assert np.array_equal(graph[k], want), f"graph frame {k}"
        assert np.array_equal(serial[k], want), f"serial frame {k}"


def test_batch_at_benchmark_shape_against_oracle(device):
    """The benchmark's batch itself: 256 frames of 16 seeded scenes of 8 to
    512 quadrics, resolutions log-uniform from 64x64 to 1920x1080, AA
    1x/2x/4x and depth 1-4 at random (the requests bench.py draws with seed
    1), traced as one batch and replayed from the captured graph: every frame
    equals its scene rendered alone, and two rows of every frame equal the
    oracle's."""
    rng = np.random.default_rng(1)
    scenes = [(7, 1000 + k) for k in range(16)]
    reqs = []
    for _ in range(256):
        t = rng.uniform()
        x = int(round(64 * (1920 / 64) ** t))
        y = int(round(64 * (1080 / 64) ** t))
        fsaa = int(rng.integers(3))
        reqs.append((int(rng.integers(16)), x, y, fsaa, int(rng.integers(1, 5))))
    print("case batch of 256 frames at the benchmark's shape", flush=True)
    bc = BatchCase(device, scenes, reqs)
    try:
        got = bc.batch.render(qrhip.BATCH_DIRECT)
        graph = bc.batch.render(qrhip.BATCH_GRAPH)
    finally:
        bc.close()
    samples = sum(x * y * (1, 2, 4)[f] for _, x, y, f, _ in reqs)
    assert samples > 150e6
    for k, key in enumerate(bc.keys):
        want = alone(device, *key)
        assert np.array_equal(got[k], want), f"frame {k} {key} differs from its scene alone"
        assert np.array_equal(graph[k], want), f"graph frame {k} {key}"
    print("batch: every frame equals its scene alone", flush=True)
    pick = np.random.default_rng(2)
    cases = [(c, sd, x, y, f, d, 0, False,
              sorted(set(int(v) for v in pick.integers(0, y, 2))))
             for (c, sd, x, y, f, d) in bc.keys]
    for k, (case, rows) in enumerate(zip(cases, oracle_rows(cases))):
        for r, px in rows.items():
            assert np.array_equal(got[k][r], px), f"frame {k} {bc.keys[k]} row {r}"
    print("batch: sampled rows equal the oracle", flush=True)
