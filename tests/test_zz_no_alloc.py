"""qrhip_render_tables_noalloc, the entry point that allocates no device
memory (its buffers come from one static device array per GPU), against
qrhip_render_tables, which allocates: bit-identical frames and first-hit ids
on two calls in a row, the second one smaller (the array keeps the first
call's bytes, which the second must not read), and against the oracle."""
import numpy as np
import pytest

import qrhip
from helpers import HostScene

pytestmark = pytest.mark.gpu


def _both(device, hs, want_ids=True):
    a, ia = device.render_tables(hs.tables, hs.fd, want_ids=want_ids)
    b, ib = device.render_tables_noalloc(hs.tables, hs.fd, want_ids=want_ids)
    return a, ia, b, ib


def test_noalloc_equals_allocating_entry_point(device):
    print("case no-alloc: Whitted 320x180 4X depth 4, then plain 96x54", flush=True)
    hs = HostScene(2, 31, 320, 180, 2, 4)
    try:
        want, ids_w, got, ids_g = _both(device, hs)
        oracle = hs.oracle()[0]
    finally:
        hs.close()
    assert np.array_equal(got, want), int((got != want).sum())
    assert np.array_equal(ids_g, ids_w)
    assert np.array_equal(got, oracle)

    # second call, smaller, with a padded row pitch
    hs = HostScene(1, 32, 96, 54, 0, 2)
    try:
        hs.fd.x_row = 96 + 7
        want, ids_w, got, ids_g = _both(device, hs)
        oracle = hs.oracle()[0]
    finally:
        hs.close()
    assert np.array_equal(got, want), int((got != want).sum())
    assert np.array_equal(ids_g, ids_w)
    assert np.array_equal(got, oracle)
    print("no-alloc: both calls equal the allocating entry point and the oracle")


def test_noalloc_rejects_frames_beyond_its_scratch(device):
    """A frame whose buffers exceed QRHIP_NOALLOC_BYTES fails on the host,
    before any device work, with a message saying why."""
    hs = HostScene(1, 33, 64, 36, 0, 2)
    try:
        fd = hs.fd
        fd.x_res, fd.y_res, fd.x_row, fd.fsaa = 4096, 2048, 4096, 2   # 4X ids: 128 MiB
        with pytest.raises(qrhip.QrError, match="no-alloc scratch"):
            device.render_tables_noalloc(hs.tables, fd, want_ids=True)
    finally:
        hs.close()
