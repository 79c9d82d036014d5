"""The allocation-free residual-terms export (jst_batch_residual_terms_noalloc) against the
allocating one (jst_batch_residual_terms) through the C ABI: bitwise the same terms and dt,
on two calls in a row, the second on a smaller case, so that the fixed work-item array is
reused with fewer items than the first call left in it."""
from __future__ import annotations

import json

import numpy as np
import pytest

from conftest import config_dict, ellipse_mesh

pytestmark = pytest.mark.gpu

NAMES = [f"R{k}" for k in range(5)] + ["S"]


@pytest.fixture(scope="module")
def jst():
    from purinecfd_amd import jst as mod
    mod.lib()
    return mod


def _terms(b, cid, ni, nj):
    return {n: b.get(cid, n, ni, nj) for n in NAMES}


@pytest.mark.parametrize("variant", [10, 9, 1])  # own lengths, 1-row and 24-row segments
def test_noalloc_export_equals_the_allocating_one(jst, variant):
    cfg = config_dict()
    cfg["simulation"].update({"AOA": 3.0, "Ma": 0.4, "CFL": 1.0})
    big = (65, 256, 1.05)     # 64 x 256 cells: 5 column groups
    small = (13, 40, 1.2)     # 12 x 40 cells: 1 column group
    b = jst.Batch(1 << 30, 4)
    b.set_tiling(variant, 0)
    ids = []
    for rings, nj, st in (big, small):
        ids.append((b.add(json.dumps(cfg), ellipse_mesh(rings, nj, st), 10 ** 6), rings - 1, nj))
    b.step(7)
    for cid, ni, nj in ids:  # the large case first, then the smaller one
        dt_a = b.residual_terms_noalloc(cid)
        got = _terms(b, cid, ni, nj)
        dt_b = b.residual_terms(cid)
        want = _terms(b, cid, ni, nj)
        assert dt_a == dt_b
        for n in NAMES:
            assert np.array_equal(got[n], want[n]), (variant, cid, n)
        assert np.any(got["R1"] != 0.0)
    # the export leaves the run untouched: both cases still step
    b.step(3)
    for cid, _, _ in ids:
        assert b.info(cid)["step"] == 10
    b.close()
