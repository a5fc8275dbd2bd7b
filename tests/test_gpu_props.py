"""GPU parity of the halo property pass (hp_plan_props; src/properties.c:
384-510 for bound = 0 then bound = 1, src/properties.c:552-553) against the
restatement, which the CPU tests pin bit for bit to the reference's own
compiled code.

Both sides read the same input: the GPU bound pass's own output records
(sorted by r2, pe and ke set) and halos (kept, centre).  Tolerances:
  * counts and indices (n_mdelta, dens_tot, n_vir, n_half) and every field
    made from them alone (m, mgrav, alt_m, vmax, rvmax, halfmass_radius):
    identical;
  * r and rvir (cbrt on the device vs glibc): 2e-7 relative;
  * the fields built on f64 sums (J, Xoff, Voff, vrms, energy, kin_to_pot,
    spin, bullock_spin): 1e-5 relative, the reference's running means and
    sequential sums being restated as sums of exact differences;
  * NFW scale radii and chi2, pseudo-evolution masses: 1e-5 relative;
  * shapes (b_to_a, c_to_a, A, inertia tensor, in both windows), every
    window: the device sums a window of at most 256 records or fewer than 64
    bound records in the reference's order, record by record (the iteration
    is ill-conditioned there: the reference's own result moves when only the
    order of its sums changes, tests/test_props_oracle.py), and those agree
    to 1e-6 of the field's scale (axis ratios: absolute, the near-zero ratio
    of a flat window being the last bits of a vanishing eigenvalue); larger
    windows are summed per tile and agree to 1e-4 of the field's scale.
"""
import numpy as np
import pytest

import halopot
from halopot import catalog
from halopot.hipmem import DeviceArray, synchronize

pytestmark = pytest.mark.gpu

EXACT = ("m", "mgrav", "alt_m", "vmax", "rvmax", "halfmass_radius")
CBRT = ("r", "rvir")
SUMS = ("J", "Xoff", "Voff", "vrms", "energy", "kin_to_pot", "spin", "bullock_spin")
FITS = ("rs", "klypin_rs", "chi2", "m_pe_b", "m_pe_d")
SHAPE = {1: ("b_to_a", "c_to_a", "A", "inertia_tensor"),
         0: ("b_to_a2", "c_to_a2", "A2", "inertia_tensor2")}
EXACT_RECORDS, EXACT_BOUND = 256, 64  # windows the device sums in the reference's order
RATIOS = ("b_to_a", "c_to_a", "b_to_a2", "c_to_a2")


def run_props(cat, defs=None, opts=None):
    halopot.set_config(cat.params)
    po = np.ascontiguousarray(cat.po.copy())
    halos = np.ascontiguousarray(cat.halos.copy())
    md = halopot.mass_definitions(defs)
    plan = halopot.Plan(halos, opts)
    d_po = DeviceArray.from_numpy(po)
    d_h = DeviceArray.from_numpy(halos)
    d_b = DeviceArray.zeros(max(1, po.shape[0]))
    d_p = DeviceArray.zeros(max(1, halos.shape[0]) * catalog.PROPS_DTYPE.itemsize)
    try:
        plan.run(d_po.address, d_h.address, d_b.address)
        plan.props(d_po.address, d_h.address, d_p.address, md)
        synchronize()
        po_g = d_po.download(np.empty_like(po))
        h_g = d_h.download(np.empty_like(halos))
        props = d_p.download(np.empty(halos.shape[0], catalog.PROPS_DTYPE))
    finally:
        plan.close()
        for d in (d_po, d_h, d_b, d_p):
            d.free()
    return po_g, h_g, props, md.as_array()


def _window_counts(po, halos, props, particle_mass):
    """Records and bound records of each call's two shape windows:
    [0, np_alt[2]) and [0, dens_tot), each cut at the kept count
    (alt_m[2] = np_alt[2] * PARTICLE_MASS)."""
    size = np.zeros((halos.shape[0], 2), np.int64)
    bound = np.zeros((halos.shape[0], 2), np.int64)
    for h in range(halos.shape[0]):
        f, k = int(halos["first"][h]), int(halos["kept"][h])
        b = ~(po["pe"][f:f + k] < po["ke"][f:f + k].astype(np.float64))
        n0 = int(round(float(props["alt_m"][h][2]) / particle_mass))
        size[h] = min(n0, k), min(int(props["dens_tot"][h]), k)
        bound[h] = int(b[:size[h, 0]].sum()), int(b[:size[h, 1]].sum())
    return size, bound


def check(restatement, cat, defs=None, opts=None):
    po_g, h_g, got, th = run_props(cat, defs, opts)
    want = np.zeros(cat.n_halos, catalog.PROPS_DTYPE)
    restatement.halo_props_batch(po_g, h_g, cat.params, th, want, nthreads=8)
    for f in FITS:
        np.testing.assert_allclose(got[f], want[f], rtol=1e-5, atol=0, err_msg=f)
    size, nbound = _window_counts(po_g, h_g, want, cat.params.particle_mass)
    exact = (size <= EXACT_RECORDS) | (nbound < EXACT_BOUND)
    for w, fields in SHAPE.items():
        tol = np.where(exact[:, w], 1e-6, 1e-4)[:, None]
        for f in fields:
            a = got[f].reshape(cat.n_halos, -1).astype(np.float64)
            b = want[f].reshape(cat.n_halos, -1).astype(np.float64)
            scale = np.maximum(np.max(np.abs(b), axis=1, keepdims=True), 1e-30)
            if f in RATIOS:
                scale = np.ones_like(scale)
            bad = np.nonzero(np.any(np.abs(a - b) > tol * scale, axis=1))[0]
            assert bad.size == 0, (f, bad[:5], a[bad[:5]], b[bad[:5]])
        assert np.all(np.isfinite(got[SHAPE[w][0]]))
    for f in catalog.PROPS_INT_FIELDS:
        np.testing.assert_array_equal(got[f], want[f], err_msg=f)
    for f in EXACT:
        np.testing.assert_array_equal(got[f], want[f], err_msg=f)
    for f in CBRT:
        np.testing.assert_allclose(got[f], want[f], rtol=2e-7, atol=0, err_msg=f)
    for f in SUMS:
        w = want[f].astype(np.float64)
        # an absolute floor of 1e-5 of the field's scale over the batch for
        # values that cancel to near zero (a component of J, a tiny offset)
        floor = 1e-5 * max(1e-30, float(np.max(np.abs(w))))
        np.testing.assert_allclose(got[f], w, rtol=1e-5, atol=floor, err_msg=f)
    return got


def test_props_catalogue(restatement):
    got = check(restatement, catalog.config2())
    assert (got["n_mdelta"] > 0).sum() > 100 and (got["vmax"] > 0).sum() > 100


def test_props_small_calls_and_potential_only(restatement):
    """Calls of 2 to 3,000 records with 10% potential-only records (dropped
    before the pass), including calls with no bound record."""
    cat = catalog.small_batch(seed=71, n_halos=300, lo=2, hi=3000, potential_only=0.1)
    got = check(restatement, cat)
    assert (got["n_mdelta"] == 0).any()


def test_props_other_epoch_and_definitions(restatement):
    cat = catalog.small_batch(seed=72, n_halos=60, lo=20, hi=20000)
    cat.params = catalog.Params(force_res=0.002, particle_mass=cat.params.particle_mass,
                                Om=0.3, Ol=0.7, scale_now=0.5)
    check(restatement, cat, defs=("m200c", "vir", "178b", "1000c", "200b"))


def test_props_per_call_plan(restatement):
    cat = catalog.small_batch(seed=73, n_halos=20, lo=5, hi=2000)
    check(restatement, cat, opts=halopot.options(per_call=1))


def test_props_match_the_reference_on_the_catalogue(restatement):
    """The GPU property pass against the reference's own compiled code,
    call by call, on the GPU's bound-pass output."""
    import oracle

    if not oracle.ReferenceProps.available():
        pytest.skip("compiled reference property pass (oracle/_ref) not present")
    ref = oracle.ReferenceProps()
    cat = catalog.small_batch(seed=74, n_halos=120, lo=20, hi=20000)
    po_g, h_g, got, th = run_props(cat)
    for h in range(cat.n_halos):
        f, k = int(h_g["first"][h]), int(h_g["kept"][h])
        r = ref.halo_props(po_g[f:f + k].copy(), k, h_g["pos"][h], cat.params, th)
        for name in EXACT:
            np.testing.assert_array_equal(got[h][name], r[name], err_msg=f"{h} {name}")
        for name in SUMS + ("r",):
            np.testing.assert_allclose(got[h][name], r[name], rtol=1e-5,
                                       atol=1e-5 * float(np.max(np.abs(r[name]))) + 1e-30,
                                       err_msg=f"{h} {name}")
