"""The allocation-free entry point g4mi_sort_index_noalloc against the
allocating g4mi_sort_index it mirrors, through the C-ABI: the same target
slice for every particle (bit for bit) and the same moved count, on two
calls in a row (the second with fewer particles, so stale scratch contents
would show), in the FP64 and the FP32 form; a beam above the scratch's
maximum is refused with an error."""
import math
import sys

import numpy as np
import pytest

import g4mi
import harness as H

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("precision", [g4mi.FP64, g4mi.FP32])
def test_sort_index_noalloc_matches_allocating_entry(precision):
    slen = 2 * math.pi
    for counts, spread in (([4096] * 6, 1.7), ([1000, 24, 776], 2.3)):
        _log(f"sort_index_noalloc: precision {precision}, counts {counts}")
        sim = H.make_state(counts)
        rng = np.random.default_rng(len(counts))
        sim.theta += rng.uniform(-spread, spread, size=len(sim.theta)) * slen
        ctx = H.to_gpu(sim, precision=precision)
        a1, m1 = ctx.sort_index(slen)
        a2, m2 = ctx.sort_index_noalloc(slen)
        a3, m3 = ctx.sort_index_noalloc(slen)      # the same scratch again
        assert np.array_equal(a1, a2) and m1 == m2, (m1, m2)
        assert np.array_equal(a2, a3) and m2 == m3
        assert 0 < m1 < len(a1)
        ctx.close()


def test_sort_index_noalloc_refuses_beams_above_its_scratch():
    n = (1 << 22) + 4
    off = np.array([0, n // 2, n], dtype=np.int64)
    ctx = g4mi.Context(off, 15, 0.2e-3, 1e-10, precision=g4mi.FP64)
    with pytest.raises(g4mi.G4miError, match="maximum"):
        ctx.sort_index_noalloc(2 * math.pi)
    ctx.close()
