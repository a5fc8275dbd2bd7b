"""hp_compute_potential_noalloc, the entry point that allocates nothing,
against the existing compute_potential on the same records: the same
permutation of the records and the same pe, bit for bit (the same kernels on
the same tree, in the same order), on two calls in a row, the second
smaller; the size limit is a clear error; the caller's stream orders it."""
import numpy as np
import pytest

import halopot
from halopot import catalog
from halopot.hipmem import DeviceArray, Stream, synchronize
from conftest import with_ids

pytestmark = pytest.mark.gpu


def _records(n, seed):
    cat = catalog.small_batch(seed=seed, n_halos=1, lo=n, hi=n + 1, contam=0.05)
    po = with_ids(cat.segment(0)[:n])
    return po, cat.params


def _same(got, want):
    np.testing.assert_array_equal(got["flags"], want["flags"])  # the reference's permutation
    for f in ("pos", "r2", "ke", "pe"):
        np.testing.assert_array_equal(got[f], want[f], err_msg=f)


def test_noalloc_matches_compute_potential_on_two_calls():
    first, prm = _records(halopot.NOALLOC_MAX_RECORDS, seed=31)
    first["pos"][[5, 77], 0] = np.nan                      # left out of the tree, pe = 0
    second, _ = _records(3001, seed=32)
    halopot.set_config(prm)
    for po in (first, second):
        print(f"[noalloc] {po.shape[0]} records", flush=True)
        want = po.copy()
        halopot.compute_potential(want)
        d = DeviceArray.from_numpy(po)
        try:
            halopot.compute_potential_noalloc(d.address, po.shape[0])
            synchronize()
            got = d.download(np.empty_like(po))
        finally:
            d.free()
        _same(got, want)
        if po is first:
            assert np.all(got["pe"][np.isin(got["flags"], [5, 77])] == 0)


def test_noalloc_size_limit_and_empty_call():
    d = DeviceArray.zeros(48)
    try:
        halopot.compute_potential_noalloc(d.address, 0)
        with pytest.raises(halopot.HalopotError, match="at most 32768"):
            halopot.compute_potential_noalloc(d.address, halopot.NOALLOC_MAX_RECORDS + 1)
    finally:
        d.free()


def test_noalloc_on_a_caller_stream():
    po, prm = _records(9000, seed=33)
    halopot.set_config(prm)
    want = po.copy()
    halopot.compute_potential(want)
    src = DeviceArray.from_numpy(po)
    dst = DeviceArray.from_numpy(np.zeros_like(po))
    s = Stream()
    try:
        s.copy(dst.address, src.address, po.nbytes)  # queued, not waited for
        halopot.compute_potential_noalloc(dst.address, po.shape[0], stream=s.handle)
        s.synchronize()
        _same(dst.download(np.empty_like(po)), want)
    finally:
        s.destroy()
        src.free()
        dst.free()
