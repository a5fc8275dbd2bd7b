"""Ordering with the caller's work.

* The library's stream is a blocking one, so records that the caller fills
  with a device-to-device hipMemcpy on the null stream (a copy that can
  return before it is done) are complete when a plan reads them.
* A caller stream passed to hp_plan_potential / hp_plan_run (here a
  non-blocking one, which the null stream does not order) is waited for
  before the plan reads the records, and waits for the plan before its own
  later work.
Each pass over a freshly copied array must equal the pass over the same
records uploaded and waited for."""
import numpy as np
import pytest

import halopot
from halopot import catalog
from halopot.hipmem import D2D, DeviceArray, Stream, hip, synchronize
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
            np.testing.assert_array_equal(got["pos"], want["pos"])  # the same permutation
    finally:
        plan.close()
        d_src.free()
        d_dst.free()


def test_plan_orders_with_a_caller_stream():
    cat = catalog.small_batch(seed=5, n_halos=300, lo=100, hi=20000, potential_only=0.1)
    want_pot, _, _, _ = run_plan(cat, potential_only=True)
    want_run, want_h, want_b, _ = run_plan(cat)
    halopot.set_config(cat.params)
    po = np.ascontiguousarray(cat.po.copy())
    zeros = np.zeros_like(po)
    d_src = DeviceArray.from_numpy(po)
    d_dst = DeviceArray.from_numpy(zeros)
    d_out = DeviceArray.from_numpy(zeros)
    d_h = DeviceArray.from_numpy(cat.halos)
    d_b = DeviceArray.zeros(cat.n_records)
    s = Stream()
    plan = halopot.Plan(np.ascontiguousarray(cat.halos.copy()))
    try:
        for rep in range(3):
            d_dst.upload(zeros)
            d_out.upload(zeros)
            d_h.upload(cat.halos)
            synchronize()
            s.copy(d_dst.address, d_src.address, po.nbytes)      # queued on the caller stream
            if rep < 2:
                plan.potential(d_dst.address, stream=s.handle)
                want = want_pot
            else:
                plan.run(d_dst.address, d_h.address, d_b.address, stream=s.handle)
                want = want_run
            s.copy(d_out.address, d_dst.address, po.nbytes)      # after the plan, on the stream
            s.synchronize()
            got = d_out.download(np.empty_like(po))
            np.testing.assert_array_equal(got["pe"], want["pe"])
            np.testing.assert_array_equal(got["pos"], want["pos"])
            if rep == 2:
                np.testing.assert_array_equal(d_b.download(np.empty(cat.n_records, np.uint8)),
                                              want_b)
                np.testing.assert_array_equal(d_h.download(np.empty_like(cat.halos))["n_bound"],
                                              want_h["n_bound"])
    finally:
        plan.close()
        s.destroy()
        for d in (d_src, d_dst, d_out, d_h, d_b):
            d.free()
