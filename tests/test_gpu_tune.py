"""The autotuner (hp_plan_tune): it picks among launch settings that change
no tree, walk decision or sum, so a tuned plan must give the untuned plan's
results bit for bit; under HP_METHOD_AUTO it may also move calls of at most
10 points between the tree walk and the direct kernel, which agree to float
rounding (the tree of such a call is a single leaf)."""
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
        assert t["big_threshold"] in (128, 256, 512)
        assert t["wave_max"] in (32, 64)
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
