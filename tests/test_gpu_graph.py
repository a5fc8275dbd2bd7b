"""The hipGraph replay of hp_plan_run (use_graph) against the eager path.

The first graph run is eager and records the graphs; later runs replay
them.  Every replay is checked against the eager result (bit for bit: the
kernels are the same and their sums are ordered by the tree, not by
scheduling), with eager work and synchronisations in between.  A batch
whose trees need more split levels than were recorded is finished eagerly
from the device state and must give the eager answer too.  The graphs are
recorded again when a configuration global they take (PARTICLE_MASS,
FORCE_RES, SCALE_NOW, the cosmology) changes between runs of one plan."""
import dataclasses

import numpy as np
import pytest

import halopot
from halopot import catalog
from halopot.hipmem import DeviceArray, synchronize

pytestmark = pytest.mark.gpu


def _eager(cat):
    halopot.set_config(cat.params)
    plan = halopot.Plan(cat.halos, halopot.options())
    d_po = DeviceArray.from_numpy(cat.po)
    d_h = DeviceArray.from_numpy(cat.halos)
    d_b = DeviceArray.zeros(cat.n_records)
    try:
        plan.run(d_po.address, d_h.address, d_b.address)
        synchronize()
        return (d_po.download(np.empty_like(cat.po)), d_h.download(np.empty_like(cat.halos)),
                d_b.download(np.empty(cat.n_records, np.uint8)))
    finally:
        plan.close()


def _same(a, b):
    np.testing.assert_array_equal(a[0]["pe"], b[0]["pe"])
    np.testing.assert_array_equal(a[0]["ke"], b[0]["ke"])
    np.testing.assert_array_equal(a[0]["flags"], b[0]["flags"])
    np.testing.assert_array_equal(a[1]["corevel"], b[1]["corevel"])
    np.testing.assert_array_equal(a[1]["n_bound"], b[1]["n_bound"])
    np.testing.assert_array_equal(a[2], b[2])


def test_graph_replays_match_eager():
    cat = catalog.small_batch(seed=41, n_halos=40, lo=5, hi=6000, potential_only=0.1)
    cat.po["flags"] = np.arange(cat.n_records, dtype=np.int32)
    want = _eager(cat)
    halopot.set_config(cat.params)
    plan = halopot.Plan(cat.halos, halopot.options(use_graph=1, timing=1))
    d_po = DeviceArray.from_numpy(cat.po)
    d_h = DeviceArray.from_numpy(cat.halos)
    d_b = DeviceArray.zeros(cat.n_records)
    try:
        for rep in range(4):
            d_po.upload(cat.po)
            d_h.upload(cat.halos)
            plan.run(d_po.address, d_h.address, d_b.address)
            synchronize()
            got = (d_po.download(np.empty_like(cat.po)), d_h.download(np.empty_like(cat.halos)),
                   d_b.download(np.empty(cat.n_records, np.uint8)))
            _same(got, want)
            st = plan.stats()
            assert st["nodes"] > 0 and st["ms_walk"] > 0
            # eager work of the program between replays
            other = catalog.small_batch(seed=100 + rep, n_halos=5, lo=20, hi=500)
            _eager(other)
    finally:
        plan.close()


def test_graph_deeper_tree_falls_back():
    base = catalog.small_batch(seed=42, n_halos=6, lo=2000, hi=4000)
    base.po["flags"] = np.arange(base.n_records, dtype=np.int32)
    deep = base.copy()
    # same calls, but one halo collapsed onto a geometric cluster: many more levels
    f, c = int(deep.halos["first"][0]), int(deep.halos["count"][0])
    k = np.arange(c)
    cen = deep.halos["pos"][0, :3]
    deep.po["pos"][f:f + c, 0] = cen[0] + (0.5 ** (k % 40)) * 0.3
    deep.po["pos"][f:f + c, 1] = cen[1] + (k % 7) * 1e-3
    deep.po["pos"][f:f + c, 2] = cen[2] + (k % 5) * 1e-3
    want_base, want_deep = _eager(base), _eager(deep)
    halopot.set_config(base.params)
    plan = halopot.Plan(base.halos, halopot.options(use_graph=1))
    d_po = DeviceArray.from_numpy(base.po)
    d_h = DeviceArray.from_numpy(base.halos)
    d_b = DeviceArray.zeros(base.n_records)
    try:
        for cat, want in ((base, want_base), (base, want_base), (deep, want_deep),
                          (deep, want_deep), (base, want_base)):
            d_po.upload(cat.po)
            d_h.upload(cat.halos)
            plan.run(d_po.address, d_h.address, d_b.address)
            synchronize()
            got = (d_po.download(np.empty_like(cat.po)), d_h.download(np.empty_like(cat.halos)),
                   d_b.download(np.empty(cat.n_records, np.uint8)))
            _same(got, want)
    finally:
        plan.close()


def test_graph_rerecords_on_config_change():
    cat = catalog.small_batch(seed=43, n_halos=30, lo=5, hi=5000, potential_only=0.1)
    cat.po["flags"] = np.arange(cat.n_records, dtype=np.int32)
    p0 = cat.params
    p1 = dataclasses.replace(p0, scale_now=0.5, particle_mass=2.5 * p0.particle_mass,
                             force_res=0.7 * p0.force_res, Om=0.3, Ol=0.7)
    other = cat.copy()
    other.params = p1
    want0, want1 = _eager(cat), _eager(other)
    assert not np.array_equal(want0[0]["pe"], want1[0]["pe"])
    assert not np.array_equal(want0[0]["ke"], want1[0]["ke"])
    plan = halopot.Plan(cat.halos, halopot.options(use_graph=1))
    d_po = DeviceArray.from_numpy(cat.po)
    d_h = DeviceArray.from_numpy(cat.halos)
    d_b = DeviceArray.zeros(cat.n_records)
    try:
        for prm, want in ((p0, want0), (p0, want0), (p1, want1), (p1, want1), (p0, want0)):
            halopot.set_config(prm)
            d_po.upload(cat.po)
            d_h.upload(cat.halos)
            plan.run(d_po.address, d_h.address, d_b.address)
            synchronize()
            got = (d_po.download(np.empty_like(cat.po)), d_h.download(np.empty_like(cat.halos)),
                   d_b.download(np.empty(cat.n_records, np.uint8)))
            _same(got, want)
    finally:
        plan.close()
        halopot.set_config(p0)
