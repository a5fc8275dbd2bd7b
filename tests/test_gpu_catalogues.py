"""Full-size parity of the batched path on the catalogues the bench reports,
each on the bench's own path (an autotuned plan replayed from hipGraphs):

  c2  2,000 calls, 0.53M records (power-law sizes up to 210,000);
  c3  the 2.13M-record host call and its 30 sub-halo calls: the deep tree,
      whose split levels run past the depth the graphs were recorded with;
  c5  the one-GPU catalogue, 50,000 calls, 6.71M records.

Checked against the oracle:
  * walk decisions: the batch's P2P pair and P2M term counts equal the
    restatement's (summed over its per-call walks) under every launch
    setting the autotuner can choose (the trees, and so every decision, do
    not depend on the setting; one of them differs per call-size bucket),
    with the sibling-pair walk and with the node-by-node walk;
  * the bound pass against the reference's own compute_potential and
    compute_kinetic_energy (oracle/_ref) inside the restated driver of
    src/properties.c:537-550: kept counts and n_core identical; the kept
    records and their r2 order identical; pe within 1e-5 relative (FP32
    pair terms, compensated sums); corevel, bulkvel within 1e-6 relative;
    ke within 1e-5; bound flags identical except where |pe - ke| / pe < 1e-4,
    the stated margin;
  * pe against the FP64 evaluation of the same trees and decisions
    (HP_ACC_F64): median 1e-6, max 1e-5 -- the path's "<= 1e-4 relative
    error vs FP64" met with room.
"""
import numpy as np
import pytest

import halopot
from halopot import catalog
from halopot.hipmem import D2D, DeviceArray, hip, synchronize
from gpu_util import rel

pytestmark = pytest.mark.gpu

MARGIN = 1e-4
SETTINGS = [(bt, wm) for bt in (128, 256, 512, 1024) for wm in (32, 64)]
SETTINGS.append(([1024, 512, 256, 128], [64, 32, 64, 32]))  # a different one per call-size bucket
THREADS = 16


@pytest.fixture(scope="module", params=["c2", "c3", "c5"])
def cat(request):
    print(f"\n[catalogue {request.param}]", flush=True)
    c = catalog.make(request.param)
    c.po["flags"] = np.arange(c.n_records, dtype=np.int32)  # record ids (the path never reads flags)
    return c


class Device:
    """Pristine records on the device, restored into the working copy
    (device to device, waited for) before every pass."""

    def __init__(self, cat):
        self.cat = cat
        self.src = DeviceArray.from_numpy(cat.po)
        self.po = DeviceArray(cat.po.nbytes)
        self.hsrc = DeviceArray.from_numpy(cat.halos)
        self.h = DeviceArray(cat.halos.nbytes)
        self.b = DeviceArray.zeros(cat.n_records)

    def restore(self):
        assert hip().hipMemcpy(self.po.ptr, self.src.ptr, self.cat.po.nbytes, D2D) == 0
        assert hip().hipMemcpy(self.h.ptr, self.hsrc.ptr, self.cat.halos.nbytes, D2D) == 0
        synchronize()

    def results(self):
        c = self.cat
        return (self.po.download(np.empty_like(c.po)), self.h.download(np.empty_like(c.halos)),
                self.b.download(np.empty(c.n_records, np.uint8)))

    def free(self):
        for d in (self.src, self.po, self.hsrc, self.h, self.b):
            d.free()


def _kept_mask(halos, kept):
    count = halos["count"].astype(np.int64)
    call = np.repeat(np.arange(halos.shape[0]), count)
    pos = np.arange(int(count.sum())) - halos["first"][call]
    return pos < np.asarray(kept)[call], call


def test_walk_counts_every_setting(restatement, cat):
    prm = cat.params
    halopot.set_config(prm)
    p2p, p2m = restatement.potential_batch(cat.po.copy(), cat.halos, prm, nthreads=THREADS)
    want = (int(p2p.sum()), int(p2m.sum()))
    print(f"restatement: {want[0]} P2P pairs, {want[1]} P2M terms", flush=True)
    dev = Device(cat)
    try:
        for walk in (halopot.WALK_PAIRS, halopot.WALK_NODES):
            plan = halopot.Plan(cat.halos, halopot.options(count_work=1, walk=walk))
            try:
                for bt, wm in SETTINGS if walk == halopot.WALK_PAIRS else SETTINGS[:1]:
                    plan.set_tuning(bt, wm)
                    dev.restore()
                    plan.potential(dev.po.address)
                    st = plan.stats()
                    assert (st["p2p"], st["p2m"]) == want, (walk, bt, wm, st["p2p"], st["p2m"])
            finally:
                plan.close()
    finally:
        dev.free()


def test_bound_pass_vs_reference(restatement, reference, cat):
    prm = cat.params
    po_r, h_r = cat.po.copy(), cat.halos.copy()
    kept_r = restatement.bound_pass_batch_with(po_r, h_r, prm, reference, nthreads=THREADS)

    halopot.set_config(prm)
    dev = Device(cat)
    plan = halopot.Plan(cat.halos, halopot.options(use_graph=1, timing=1))
    try:
        dev.restore()
        tuning = plan.tune(dev.po.address, reps=1)
        print(f"tuning: {tuning}", flush=True)
        for _ in range(3):  # the first run records the graphs, the others replay them
            dev.restore()
            plan.run(dev.po.address, dev.h.address, dev.b.address)
        synchronize()
        po_g, h_g, b_g = dev.results()
    finally:
        plan.close()
        dev.free()

    np.testing.assert_array_equal(h_g["kept"], kept_r)
    np.testing.assert_array_equal(h_g["n_core"], h_r["n_core"])
    mask, call = _kept_mask(cat.halos, kept_r)
    np.testing.assert_array_equal(po_g["r2"][mask], po_r["r2"][mask])
    ids_g, ids_r = po_g["flags"][mask], po_r["flags"][mask]
    np.testing.assert_array_equal(ids_g, ids_r)  # the same records in the same order

    first, count = cat.halos["first"][call[mask]], cat.halos["count"][call[mask]]
    assert np.all((ids_g >= first) & (ids_g < first + count))
    r = rel(po_g["pe"][mask], po_r["pe"][mask])
    print(f"pe vs reference BH: median {np.median(r):.3g}, max {r.max():.3g}", flush=True)
    assert r.max() < 1e-5
    np.testing.assert_allclose(po_g["ke"][mask], po_r["ke"][mask], rtol=1e-5, atol=1e-3)
    for f in ("corevel", "bulkvel", "pos"):
        np.testing.assert_allclose(h_g[f], h_r[f], rtol=1e-6, atol=1e-4, err_msg=f)
    for f in ("min_vel_err", "min_bulkvel_err"):
        np.testing.assert_allclose(h_g[f], h_r[f], rtol=1e-5, atol=1e-6, err_msg=f)

    bg = b_g[mask].astype(bool)
    pe, ke = po_r["pe"][mask], po_r["ke"][mask]
    br = ~(pe < ke)
    mism = bg != br
    if mism.any():
        margin = np.abs(pe[mism] - ke[mism]) / np.maximum(pe[mism], 1e-300)
        assert np.all(margin < MARGIN), margin.max()
    print(f"bound flags: {int(mism.sum())} of {int(mask.sum())} differ, all inside the margin",
          flush=True)
    nb = np.bincount(call[mask], weights=bg, minlength=cat.n_halos).astype(np.int64)
    np.testing.assert_array_equal(h_g["n_bound"], nb)


def test_pe_vs_fp64_same_trees(cat):
    prm = cat.params
    halopot.set_config(prm)
    dev = Device(cat)
    out = {}
    try:
        for acc in (halopot.ACC_COMPENSATED, halopot.ACC_F64):
            plan = halopot.Plan(cat.halos, halopot.options(accumulate=acc))
            try:
                dev.restore()
                plan.potential(dev.po.address)
                synchronize()
                out[acc] = dev.po.download(np.empty_like(cat.po))
            finally:
                plan.close()
    finally:
        dev.free()
    g, f = out[halopot.ACC_COMPENSATED], out[halopot.ACC_F64]
    np.testing.assert_array_equal(g["flags"], f["flags"])
    live = f["pe"] != 0
    r = rel(g["pe"][live], f["pe"][live])
    print(f"pe vs FP64 of the same trees: median {np.median(r):.3g}, p99 "
          f"{np.quantile(r, 0.99):.3g}, max {r.max():.3g}", flush=True)
    assert np.median(r) < 1e-6 and r.max() < 1e-5
    assert np.all(g["pe"][~live] == 0)
