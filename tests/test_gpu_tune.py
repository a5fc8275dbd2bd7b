"""The autotuner (hp_plan_tune): it picks, per call-size bucket (calls of at
most 1,024, 8,192, 65,536 and more records), among launch settings that
change no tree, walk decision or sum, so a tuned plan must give the untuned
plan's results bit for bit; under HP_METHOD_AUTO it may also move calls of
at most 10 points between the tree walk and the direct kernel, which agree
to float rounding (the tree of such a call is a single leaf)."""
import numpy as np
import pytest

import halopot
from halopot import catalog
from halopot.hipmem import DeviceArray, synchronize

pytestmark = pytest.mark.gpu


def _run(cat, plan):
    d_po = DeviceArray.from_numpy(cat.po)
    d_h = DeviceArray.from_numpy(cat.halos)
    d_b = DeviceArray.zeros(cat.n_records)
    try:
        plan.run(d_po.address, d_h.address, d_b.address)
        synchronize()
        return (d_po.download(np.empty_like(cat.po)), d_h.download(np.empty_like(cat.halos)),
                d_b.download(np.empty(cat.n_records, np.uint8)))
    finally:
        for d in (d_po, d_h, d_b):
            d.free()


@pytest.mark.parametrize("method", [halopot.METHOD_TREE, halopot.METHOD_AUTO])
def test_tuned_plan_gives_the_same_results(method):
    cat = catalog.small_batch(seed=61, n_halos=50, lo=3, hi=3000, potential_only=0.1)
    cat.po["flags"] = np.arange(cat.n_records, dtype=np.int32)
    halopot.set_config(cat.params)
    base = halopot.Plan(cat.halos, halopot.options(method=method))
    try:
        want = _run(cat, base)
    finally:
        base.close()
    plan = halopot.Plan(cat.halos, halopot.options(method=method))
    d_tmp = DeviceArray.from_numpy(cat.po)
    try:
        t = plan.tune(d_tmp.address, reps=1)
        assert len(t["big_threshold"]) == len(t["wave_max"]) == halopot.TUNE_BUCKETS
        assert all(b in (128, 256, 512, 1024) for b in t["big_threshold"])
        assert all(w in (32, 64) for w in t["wave_max"])
        assert t["direct_max"] in (0, 10)
        assert t["ms_potential"] > 0
        got = _run(cat, plan)
        got2 = _run(cat, plan)  # the tuned setting is kept
    finally:
        plan.close()
        d_tmp.free()
    np.testing.assert_array_equal(got[0]["flags"], want[0]["flags"])
    np.testing.assert_array_equal(got2[0]["pe"], got[0]["pe"])
    tiny = np.repeat(cat.halos["count"] <= 10, cat.halos["count"])
    np.testing.assert_array_equal(got[0]["pe"][~tiny], want[0]["pe"][~tiny])
    np.testing.assert_allclose(got[0]["pe"][tiny], want[0]["pe"][tiny], rtol=1e-6)
    np.testing.assert_array_equal(got[1]["n_bound"], want[1]["n_bound"])


def _buckets_batch():
    """Calls in every bucket: 40 small ones, then calls of about 3,000,
    20,000 and 90,000 records (the last two above 10^4)."""
    parts = [catalog.small_batch(seed=71, n_halos=40, lo=12, hi=900, potential_only=0.05)]
    for seed, n in ((72, 3000), (73, 20000), (74, 90000)):
        parts.append(catalog.small_batch(seed=seed, n_halos=1, lo=n, hi=n + 1))
    halos = catalog._concat([c.halos for c in parts])
    halos["first"] = np.concatenate([[0], np.cumsum(halos["count"])[:-1]])
    cat = catalog.Catalogue("buckets", parts[0].params, catalog._concat([c.po for c in parts]),
                            halos)
    cat.po["flags"] = np.arange(cat.n_records, dtype=np.int32)
    return cat


def test_per_bucket_settings_give_the_same_results():
    """A different setting in every bucket (and the tuner's own choice) on
    calls up to 90,000 records: the same walk decisions (P2P and P2M
    counts), the same record order and bound counts, pe bit for bit."""
    cat = _buckets_batch()
    sizes = cat.halos["count"]
    assert {halopot.tune_bucket(int(n)) for n in sizes} == {0, 1, 2, 3}
    assert (sizes > 10_000).sum() >= 2
    halopot.set_config(cat.params)
    results = []
    settings = [None, ([1024, 512, 256, 128], [64, 32, 64, 32]),
                ([128, 1024, 128, 512], [32, 64, 32, 64]), "tuned"]
    for setting in settings:
        plan = halopot.Plan(cat.halos, halopot.options(count_work=1))
        d_tmp = DeviceArray.from_numpy(cat.po)
        try:
            if setting == "tuned":
                t = plan.tune(d_tmp.address, reps=1)
                print(f"tuned per bucket: big {t['big_threshold']}, wave {t['wave_max']}", flush=True)
            elif setting is not None:
                plan.set_tuning(*setting)
            out = _run(cat, plan)
            st = plan.stats()
        finally:
            plan.close()
            d_tmp.free()
        results.append((out, (st["p2p"], st["p2m"])))
    (want, wcounts) = results[0]
    for got, counts in results[1:]:
        assert counts == wcounts
        np.testing.assert_array_equal(got[0]["flags"], want[0]["flags"])
        np.testing.assert_array_equal(got[0]["pe"], want[0]["pe"])
        np.testing.assert_array_equal(got[1]["n_bound"], want[1]["n_bound"])
        np.testing.assert_array_equal(got[1]["n_core"], want[1]["n_core"])


def test_set_tuning_rejects_other_values():
    cat = catalog.small_batch(seed=75, n_halos=5, lo=20, hi=200)
    plan = halopot.Plan(cat.halos)
    try:
        with pytest.raises(halopot.HalopotError):
            plan.set_tuning([1024, 1024, 1024, 300], 64)
        with pytest.raises(halopot.HalopotError):
            plan.set_tuning(1024, [64, 64, 48, 64])
    finally:
        plan.close()
