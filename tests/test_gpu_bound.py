"""GPU parity of the batched bound pass (src/properties.c:537-550: potential,
potential-only records dropped, sort by r2, calculate_corevel,
compute_kinetic_energy, bound test !(pe < ke) of src/properties.c:403)
against the restatement of that driver, call by call.

Tolerances:
  * kept counts and the record order (ties in r2 included: the order the
    reference's in-place tree build and potential-only strip leave before its
    stable sort): identical;
  * n_core (the selected prefix): identical;
  * corevel / bulkvel: 1e-6 relative (the GPU restates the running Welford
    moments with exact prefix sums; the two differ in the last bits);
  * ke: 1e-5 relative;
  * bound flags: identical except where |pe - ke| / pe < 1e-4, the stated
    margin (pe carries the 1e-5 FP32-vs-FP64 difference).
"""
import numpy as np
import pytest

import halopot
from halopot import catalog
from conftest import with_ids
from gpu_util import rel, run_plan

pytestmark = pytest.mark.gpu

MARGIN = 1e-4


def _check_batch(restatement, cat, opts=None):
    cat = cat.copy()
    cat.po["flags"] = np.arange(cat.n_records, dtype=np.int32)
    prm = cat.params
    po_g, h_g, bound_g, st = run_plan(cat, opts)
    po_o = cat.po.copy()
    h_o = cat.halos.copy()
    kept_o, bound_o = restatement.bound_pass_batch(po_o, h_o, prm, nthreads=8)
    n_mismatch = n_margin = 0
    for h in range(cat.n_halos):
        f = int(cat.halos["first"][h])
        k = int(kept_o[h])
        assert int(h_g["kept"][h]) == k, h
        g = po_g[f:f + k]
        o = po_o[f:f + k]
        np.testing.assert_array_equal(g["r2"], o["r2"])
        # the same records in the same order, ties in r2 included
        np.testing.assert_array_equal(g["flags"], o["flags"], err_msg=f"call {h}")
        if k > 1:
            assert np.max(rel(g["pe"], o["pe"])) < 1e-5, h
        hg, ho = h_g[h], h_o[h]
        assert hg["n_core"] == ho["n_core"], h
        np.testing.assert_allclose(hg["corevel"], ho["corevel"], rtol=1e-6, atol=1e-4)
        np.testing.assert_allclose(hg["bulkvel"], ho["bulkvel"], rtol=1e-6, atol=1e-4)
        np.testing.assert_allclose(hg["pos"], ho["pos"], rtol=1e-6, atol=1e-4)
        np.testing.assert_allclose(hg["min_vel_err"], ho["min_vel_err"], rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(hg["min_bulkvel_err"], ho["min_bulkvel_err"], rtol=1e-5,
                                   atol=1e-6)
        ke_g, ke_o = g["ke"], o["ke"]
        np.testing.assert_allclose(ke_g, ke_o, rtol=1e-5, atol=1e-3)
        bg = bound_g[f:f + k].astype(bool)
        bo = bound_o[f:f + k].astype(bool)
        mism = bg != bo
        if mism.any():
            pe = o["pe"][mism]
            margin = np.abs(pe - ke_o[mism]) / np.maximum(pe, 1e-300)
            assert np.all(margin < MARGIN), (h, margin.max())
            n_margin += int(mism.sum())
        n_mismatch += int(mism.sum())
        assert int(h_g["n_bound"][h]) == int(bg.sum())
    return n_mismatch, st


def test_bound_small_batch(restatement):
    cat = catalog.small_batch(seed=11, n_halos=50, lo=2, hi=4000, contam=0.05,
                              potential_only=0.2, sub_fraction=0.5)
    _check_batch(restatement, cat)


def test_bound_c1(restatement):
    _check_batch(restatement, catalog.config1())


def test_bound_repeat_is_deterministic():
    cat = catalog.small_batch(seed=3, n_halos=20, lo=15, hi=3000)
    a = run_plan(cat)
    b = run_plan(cat, repeats=3)
    np.testing.assert_array_equal(a[0]["pe"], b[0]["pe"])
    np.testing.assert_array_equal(a[0]["ke"], b[0]["ke"])
    np.testing.assert_array_equal(a[2], b[2])


def test_plain_kinetic_energy(restatement, request):
    cat = catalog.config1()
    prm = cat.params
    halopot.set_config(prm)
    po = with_ids(cat.po)
    po["ke"][::7] = -1.0  # potential-only records are left alone
    vc = np.array([100.0, -50.0, 20.0], np.float32)
    pc = cat.halos["pos"][0, :3].copy()
    g = po.copy()
    halopot.compute_kinetic_energy(g, vc, pc, 1.0)
    o = po.copy()
    restatement.kinetic(o, vc, pc, 1.0, prm)
    np.testing.assert_array_equal(g["ke"], o["ke"])
    try:
        ref = request.getfixturevalue("reference")
    except pytest.skip.Exception:
        return
    r = po.copy()
    ref.compute_kinetic_energy(r, vc, pc, 1.0, prm)
    np.testing.assert_array_equal(g["ke"], r["ke"])


def test_bound_sort_inside_calls_at_every_size(restatement):
    """The sort by r2 inside each call at the sizes where its method changes
    (one wavefront up to 64 records, one workgroup up to 1,024, radix passes
    over tiles of 4,096 above), with r2 full of ties and potential-only
    records spread through each call: the kept records come out in r2 order,
    ties in the order the reference's compute_potential and potential-only
    strip leave them (then a stable sort, as glibc's merge-sort qsort): the
    restatement's order, itself pinned to the compiled reference's; the kept
    counts are exact."""
    src = catalog.small_batch(seed=5, n_halos=40, lo=800, hi=4000, potential_only=0.1)
    sizes = [1, 2, 63, 64, 65, 300, 1023, 1024, 1025, 4096, 4097, 9000]
    assert sum(sizes) <= src.n_records
    cat = src.copy()
    first = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    halos = np.zeros(len(sizes), cat.halos.dtype)
    for name in cat.halos.dtype.names:
        halos[name] = cat.halos[name][:len(sizes)]
    halos["first"], halos["count"] = first, sizes
    cat.halos = halos
    cat.po = np.ascontiguousarray(cat.po[:sum(sizes)])
    rng = np.random.default_rng(7)
    cat.po["r2"] = np.round(cat.po["r2"] * 8.0) / 8.0           # many exact ties
    cat.po["ke"] = np.where(rng.random(cat.po.shape[0]) < 0.15, -1.0, cat.po["ke"])
    cat.po["flags"] = np.arange(cat.po.shape[0], dtype=np.int32)
    po_g, h_g, _, _ = run_plan(cat)
    po_o, h_o = cat.po.copy(), cat.halos.copy()
    kept_o, _ = restatement.bound_pass_batch(po_o, h_o, cat.params, nthreads=8)
    for h, (f, n) in enumerate(zip(first, sizes)):
        seg = cat.po[f:f + n]
        k = int((seg["ke"] >= 0).sum())
        assert int(h_g["kept"][h]) == k == int(kept_o[h]), (h, n)
        got = po_g[f:f + k]
        assert np.all(np.diff(got["r2"]) >= 0), (h, n)
        if n >= 63:
            assert np.unique(got["r2"]).size < k  # the case has ties
        np.testing.assert_array_equal(got["flags"], po_o["flags"][f:f + k],
                                      err_msg=f"call {h} of {n}")
