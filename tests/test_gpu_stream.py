"""Ordering with the caller's work: the library's stream is a blocking one,
so records that the caller fills with a device-to-device hipMemcpy on the
null stream (a copy that can return before it is done) are complete when a
plan reads them.  Each pass over a freshly copied array must equal the pass
over the same records uploaded and waited for."""
import numpy as np
import pytest

import halopot
from halopot import catalog
from halopot.hipmem import D2D, DeviceArray, hip, synchronize
from gpu_util import run_plan

pytestmark = pytest.mark.gpu


def test_plan_reads_records_after_a_null_stream_copy():
    cat = catalog.small_batch(seed=3, n_halos=300, lo=100, hi=20000)
    want, _, _, _ = run_plan(cat, potential_only=True)
    halopot.set_config(cat.params)
    po = np.ascontiguousarray(cat.po.copy())
    zeros = np.zeros_like(po)
    d_src = DeviceArray.from_numpy(po)
    d_dst = DeviceArray.from_numpy(zeros)
    plan = halopot.Plan(np.ascontiguousarray(cat.halos.copy()))
    try:
        for _ in range(3):
            d_dst.upload(zeros)
            synchronize()
            assert hip().hipMemcpy(d_dst.ptr, d_src.ptr, po.nbytes, D2D) == 0  # not waited for
            plan.potential(d_dst.address)
            synchronize()
            got = d_dst.download(np.empty_like(po))
            np.testing.assert_array_equal(got["pe"], want["pe"])
            np.testing.assert_array_equal(got["pos"], po["pos"])
    finally:
        plan.close()
        d_src.free()
        d_dst.free()
