"""ctypes binding of liboracle.so -- the CPU restatement of QuadRay's tracer.

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never loads it.
"""
import ctypes as C
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "lib", "liboracle.so")
REF_DIR = os.path.join(ORACLE_DIR, "_ref")


class OracleAux(C.Structure):
    _fields_ = [("ids", C.c_void_p), ("flags", C.c_void_p),
                ("ops", C.c_uint64), ("rays", C.c_uint64),
                ("boxes", C.c_void_p), ("box_checked", C.c_uint64),
                ("box_outside", C.c_uint64), ("box_excess", C.c_float),
                ("pad", C.c_float), ("ops_skip", C.c_uint64),
                ("ops_box", C.c_uint64), ("n_dmask", C.c_uint64),
                ("n_twoplane", C.c_uint64), ("n_apex", C.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("liboracle.so is not built: make -C oracle")
        L = C.CDLL(LIB_PATH)
        L.qr_oracle_render.restype = C.c_int
        L.qr_oracle_render.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_int, C.c_int, C.POINTER(OracleAux)]
        _lib = L
    return _lib


def render(tables_ptr, fd, want_ids=False, threads=8, rows=None, boxes=None,
           audit=None):
    """Render a Whitted frame from the tables; returns (frame, ids, flags, ops,
    rays).  Rows are split over threads (ctypes releases the GIL).  With
    "rows" (a list of row indices) only those rows are rendered; the others
    stay zero."""
    ns = {0: 1, 1: 2, 2: 4}[fd.fsaa]
    x_res, y_res = fd.x_res, fd.y_res
    frame = np.zeros((y_res, fd.x_row), dtype=np.uint32)
    ids = np.zeros((y_res, x_res, ns), dtype=np.int32) if want_ids else None
    flags = np.zeros((y_res, x_res, ns), dtype=np.uint8) if want_ids else None
    tp = C.cast(tables_ptr, C.c_void_p)
    fdp = C.cast(C.pointer(fd), C.c_void_p)
    if rows is None:
        chunks = max(1, min(threads, y_res))
        bounds = [(y_res * k // chunks, y_res * (k + 1) // chunks)
                  for k in range(chunks)]
    else:
        bounds = [(int(r), int(r) + 1) for r in rows]
        if any(r0 < 0 or r0 >= y_res for r0, _ in bounds):
            raise ValueError("row outside the frame")
        chunks = max(1, min(threads, len(bounds)))
    auxes = [OracleAux() for _ in bounds]
    if boxes is not None:
        boxes = np.ascontiguousarray(boxes, dtype=np.float32)
    for a in auxes:
        if want_ids:
            a.ids = ids.ctypes.data
            a.flags = flags.ctypes.data
        if boxes is not None:
            a.boxes = boxes.ctypes.data

    def run(k):
        y0, y1 = bounds[k]
        return lib().qr_oracle_render(tp, fdp, frame.ctypes.data, y0, y1,
                                      C.byref(auxes[k]))

    with ThreadPoolExecutor(max_workers=chunks) as ex:
        rcs = list(ex.map(run, range(len(bounds))))
    if any(rcs):
        raise RuntimeError("oracle render failed")
    ops = sum(a.ops for a in auxes)
    rays = sum(a.rays for a in auxes)
    if audit is not None:
        audit["checked"] = sum(a.box_checked for a in auxes)
        audit["outside"] = sum(a.box_outside for a in auxes)
        audit["excess"] = max(a.box_excess for a in auxes)
        # ops under per-ray culling with the tables' culling program
        audit["ops_culled"] = (ops - sum(a.ops_skip for a in auxes)
                               + sum(a.ops_box for a in auxes))
        for k in ("n_dmask", "n_twoplane", "n_apex"):
            audit[k] = sum(getattr(a, k) for a in auxes)
    return frame[:, :x_res], ids, flags, ops, rays


def pt_seeds(n):
    """Initial seed plane of n sample slots (QuadRay's reset_pseed chain)."""
    seeds = np.zeros(n, dtype=np.uint32)
    L = lib()
    L.qr_oracle_pt_seeds.restype = None
    L.qr_oracle_pt_seeds.argtypes = [C.c_void_p, C.c_size_t]
    L.qr_oracle_pt_seeds(seeds.ctypes.data, n)
    return seeds


def render_pt(tables_ptr, fd, planes, seeds, threads=8, audit=None):
    """One path-tracer pass: planes (float32 [3][slots]) and seeds (uint32
    [slots], slot = (y * x_res + x) * samples + k) are updated in place;
    fd.pts_c must count this pass.  Returns the frame.  With "audit" (a
    dict), the pass's counted ops ("ops", "ops_culled" under per-ray culling
    with the tables' programs) and rays are stored there."""
    L = lib()
    L.qr_oracle_render_pt.restype = C.c_int
    L.qr_oracle_render_pt.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_int, C.c_int,
                                      C.POINTER(OracleAux)]
    y_res = fd.y_res
    frame = np.zeros((y_res, fd.x_row), dtype=np.uint32)
    assert planes.dtype == np.float32 and planes.flags.c_contiguous
    assert seeds.dtype == np.uint32 and seeds.flags.c_contiguous
    tp = C.cast(tables_ptr, C.c_void_p)
    fdp = C.cast(C.pointer(fd), C.c_void_p)
    chunks = max(1, min(threads, y_res))
    bounds = [(y_res * k // chunks, y_res * (k + 1) // chunks) for k in range(chunks)]
    auxes = [OracleAux() for _ in bounds]

    def run(k):
        y0, y1 = bounds[k]
        return L.qr_oracle_render_pt(tp, fdp, frame.ctypes.data, planes[0].ctypes.data,
                                     planes[1].ctypes.data, planes[2].ctypes.data,
                                     seeds.ctypes.data, y0, y1, C.byref(auxes[k]))

    with ThreadPoolExecutor(max_workers=chunks) as ex:
        rcs = list(ex.map(run, range(chunks)))
    if any(rcs):
        raise RuntimeError("oracle path tracer failed")
    if audit is not None:
        ops = sum(a.ops for a in auxes)
        audit["ops"] = ops
        audit["rays"] = sum(a.rays for a in auxes)
        audit["ops_culled"] = (ops - sum(a.ops_skip for a in auxes)
                               + sum(a.ops_box for a in auxes))
    return frame[:, :fd.x_res]


def ref_binary(precision="f32"):
    """Path of QuadRay's own renderer built by `make -C oracle ref`, or None."""
    p = os.path.join(REF_DIR, "ref_" + precision)
    return p if os.path.exists(p) else None


def ref_render(config, seed, x_res=None, y_res=None, fsaa=None, depth=None,
               ids=False, threads=1, frames=1, warmup=0, precision="f32",
               out=None, spp=0, planes=None, actions=None):
    """Run QuadRay's renderer on a seeded workload; returns (frame, info).
    actions ("T:a,a;T:a") replays camera actions between renders first."""
    import json
    import tempfile
    exe = ref_binary(precision)
    if exe is None:
        raise RuntimeError("reference renderer not built: make -C oracle ref")
    cmd = [exe, str(config), str(seed), "-t", str(threads), "-f", str(frames),
           "-w", str(warmup)]
    if x_res is not None:
        cmd += ["-x", str(x_res)]
    if y_res is not None:
        cmd += ["-y", str(y_res)]
    if fsaa is not None:
        cmd += ["-a", str(fsaa)]
    if depth is not None:
        cmd += ["-d", str(depth)]
    if ids:
        cmd += ["-i"]
    if spp:
        cmd += ["-q", str(spp)]
    if planes is not None:
        cmd += ["-P", planes]
    if actions is not None:
        cmd += ["-c", actions]
    tmp = None
    if out is None:
        tmp = tempfile.NamedTemporaryFile(suffix=".bin", delete=False)
        out = tmp.name
        tmp.close()
    cmd += ["-o", out]
    res = subprocess.run(cmd, check=True, capture_output=True, text=True)
    info = json.loads(res.stdout.strip().splitlines()[-1])
    frame = np.fromfile(out, dtype=np.uint32).reshape(info["y_res"], info["x_res"])
    if tmp is not None:
        os.unlink(out)
    return frame, info
