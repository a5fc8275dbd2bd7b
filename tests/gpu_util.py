"""Helpers for the GPU tests: run a plan on device-resident copies of a
catalogue and bring the results back.  Device memory comes from the HIP
runtime libhalopot links (halopot.hipmem)."""
import numpy as np

import halopot
from halopot import catalog
from halopot.hipmem import DeviceArray, synchronize


def run_plan(cat, opts=None, potential_only=False, repeats=1):
    halopot.set_config(cat.params)
    po = np.ascontiguousarray(cat.po.copy())
    halos = np.ascontiguousarray(cat.halos.copy())
    plan = halopot.Plan(halos, opts)
    d_po = DeviceArray.from_numpy(po)
    d_h = DeviceArray.from_numpy(halos)
    d_b = DeviceArray.zeros(max(1, po.shape[0]))
    try:
        for _ in range(repeats):
            d_po.upload(po)
            d_h.upload(halos)
            if potential_only:
                plan.potential(d_po.address)
            else:
                plan.run(d_po.address, d_h.address, d_b.address)
        synchronize()
        st = plan.stats()
        po_out = d_po.download(np.empty_like(po))
        h_out = d_h.download(np.empty_like(halos))
        bound = d_b.download(np.empty(po.shape[0], np.uint8))
    finally:
        plan.close()
        d_po.free()
        d_h.free()
        d_b.free()
    return po_out, h_out, bound, st


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.abs(a - b)
    s = np.maximum(np.abs(b), 1e-300)
    return d / s
