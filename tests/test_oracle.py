"""The oracle, pinned: the restatement of the reference's potential and bound
pass (oracle/restate.c) against

  * the reference's own outputs: golden vectors from the compiled reference
    (tests/golden/, written by make_golden.py) and, where oracle/_ref is
    built, the live reference on regenerated inputs;
  * what the mathematics fixes: a tree of at most 10 points is one leaf, so
    Barnes-Hut equals the direct sum; with every monopole refused the walk is
    the direct sum; every particle's partners and monopoles account for every
    other particle exactly once; two-particle and coincident-point values;
    closed forms of the kinetic energy;
  * a brute-force O(n^2) restatement of calculate_corevel's selection.

CPU only.
"""
import pathlib

import numpy as np
import pytest

from halopot import catalog
from conftest import by_id, with_ids

GOLD = pathlib.Path(__file__).resolve().parent / "golden"


def _rel(a, b):
    return np.abs(np.asarray(a, float) - b) / np.maximum(np.abs(b), 1e-300)


def direct_float_diff(po, prm):
    """Textbook direct sum with the reference's pair arithmetic
    (src/potential.c:22-36): float differences, double squares and root,
    softened at FORCE_RES.  Non-finite records are left out (pe 0)."""
    x = po["pos"][:, :3].astype(np.float32)
    ok = np.isfinite(x).all(axis=1)
    out = np.zeros(x.shape[0])
    xs = x[ok]
    d = (xs[:, None, :] - xs[None, :, :]).astype(np.float32).astype(np.float64)
    r = np.sqrt((d * d).sum(axis=2))
    r = np.maximum(r, prm.force_res)
    np.fill_diagonal(r, np.inf)
    out[ok] = (prm.particle_mass * (1.0 / r)).sum(axis=1)
    return out


# ------------------------------------------------------------ golden vectors

def test_golden_c1(restatement):
    g = np.load(GOLD / "c1_reference.npz")
    cat = catalog.config1()
    np.testing.assert_array_equal(cat.po["pos"], g["pos"])  # generator unchanged
    po = np.zeros(g["pos"].shape[0], catalog.POTENTIAL_DTYPE)
    po["pos"], po["r2"] = g["pos"], g["r2"]
    q = with_ids(po)
    restatement.compute_potential(q, cat.params)
    r = _rel(by_id(q), g["pe"])
    assert r.max() < 1e-12
    assert np.mean(r == 0) > 0.9  # mostly bit-identical: only f64 summation order differs
    k = po.copy()
    k["ke"] = g["ke_in"]
    restatement.kinetic(k, g["vel_cen"], g["pos_cen"], 1.0, cat.params)
    np.testing.assert_array_equal(k["ke"], g["ke"])


def test_golden_micro(restatement):
    g = np.load(GOLD / "micro_reference.npz")
    prm = catalog.Params(force_res=g["params"][0], particle_mass=g["params"][1])
    off = g["offsets"]
    for a, b in zip(off[:-1], off[1:]):
        po = np.zeros(b - a, catalog.POTENTIAL_DTYPE)
        po["pos"] = g["pos"][a:b]
        q = with_ids(po)
        restatement.compute_potential(q, prm)
        want = g["pe"][a:b]
        got = by_id(q)
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=0)


def test_golden_batch(restatement):
    g = np.load(GOLD / "batch_reference.npz")
    prm = catalog.Params(force_res=g["params"][0], particle_mass=g["params"][1])
    for f, c in zip(g["first"], g["count"]):
        po = np.zeros(c, catalog.POTENTIAL_DTYPE)
        po["pos"] = g["pos"][f:f + c]
        q = with_ids(po)
        restatement.compute_potential(q, prm)
        assert _rel(by_id(q), g["pe"][f:f + c]).max() < 1e-12


def test_live_reference_c3_subhalo_calls(restatement, reference):
    """Regenerated inputs through the compiled reference, on calls the
    fixtures do not hold (the deep-tree catalogue's sub-halo calls)."""
    cat = catalog.small_batch(seed=77, n_halos=8, lo=500, hi=20000, contam=0.02)
    for h in range(cat.n_halos):
        seg = with_ids(cat.segment(h))
        a = seg.copy()
        reference.compute_potential(a, cat.params)
        b = seg.copy()
        restatement.compute_potential(b, cat.params)
        assert _rel(by_id(b), by_id(a)).max() < 1e-12


def _nonfinite_batch():
    cat = catalog.small_batch(seed=78, n_halos=10, lo=11, hi=3000)
    rng = np.random.default_rng(5)
    for h in range(0, cat.n_halos, 2):
        f, c = int(cat.halos["first"][h]), int(cat.halos["count"][h])
        bad = f + rng.choice(c, size=max(1, c // 7), replace=False)
        cat.po["pos"][bad, rng.integers(0, 3, bad.size)] = np.nan
    return cat


@pytest.mark.parametrize("which", ["c1", "batch", "nonfinite"])
def test_permutation_is_the_references(restatement, reference, which):
    """compute_potential permutes po (src/fast3tree.c:762-773 partition,
    875-887 non-finite removal); the restatement leaves every record where
    the compiled reference does."""
    if which == "c1":
        cat = catalog.config1()
    elif which == "batch":
        cat = catalog.small_batch(seed=79, n_halos=20, lo=11, hi=20000)
    else:
        cat = _nonfinite_batch()
    for h in range(cat.n_halos):
        seg = with_ids(cat.segment(h))
        a, b = seg.copy(), seg.copy()
        reference.compute_potential(a, cat.params)
        restatement.compute_potential(b, cat.params)
        np.testing.assert_array_equal(b["flags"], a["flags"], err_msg=f"call {h}")
        np.testing.assert_array_equal(b["pe"] == 0, a["pe"] == 0)
        assert _rel(b["pe"][a["pe"] != 0], a["pe"][a["pe"] != 0]).max(initial=0) < 1e-12


# --------------------------------------------------------------- mathematics

@pytest.mark.parametrize("n", [1, 2, 3, 7, 10])
def test_single_leaf_is_direct_sum(restatement, n):
    rng = np.random.default_rng(n)
    prm = catalog.Params()
    po = np.zeros(n, catalog.POTENTIAL_DTYPE)
    po["pos"][:, :3] = rng.uniform(0.0, 0.05, (n, 3))  # some pairs inside FORCE_RES
    q = with_ids(po)
    st = restatement.compute_potential(q, prm)
    assert st["nodes"] == 1 and st["p2m"] == 0
    np.testing.assert_allclose(by_id(q), direct_float_diff(po, prm), rtol=1e-14, atol=0)


def test_two_particles():
    import oracle

    O = oracle.Restatement()
    prm = catalog.Params(particle_mass=2.5e9)
    for d in (0.5, 0.001):  # outside and inside the softening length
        po = np.zeros(2, catalog.POTENTIAL_DTYPE)
        po["pos"][1, 1] = d
        O.compute_potential(po, prm)
        want = 2.5e9 / max(np.float64(np.float32(d)), 0.003)
        np.testing.assert_allclose(po["pe"], [want, want], rtol=1e-15)


def test_coincident_points(restatement):
    prm = catalog.Params()
    for n, self_term in ((15, True), (20, True), (21, False), (40, False)):
        po = np.zeros(n, catalog.POTENTIAL_DTYPE)
        po["pos"][:, :3] = 1.5
        restatement.compute_potential(po, prm)
        # all points equal: the root cannot split (src/fast3tree.c:760-761); a
        # leaf of more than 20 points has no self term (src/potential.c:144-146)
        want = (n - 1) * prm.particle_mass / prm.force_res if self_term else 0.0
        np.testing.assert_allclose(po["pe"], want, rtol=1e-14)


def test_nonfinite_records_get_zero(restatement):
    prm = catalog.Params()
    rng = np.random.default_rng(3)
    po = np.zeros(30, catalog.POTENTIAL_DTYPE)
    po["pos"][:, :3] = rng.uniform(0, 1, (30, 3))
    po["pos"][4, 0] = np.nan
    po["pos"][9, 2] = np.inf
    q = with_ids(po)
    st = restatement.compute_potential(q, prm)
    pe = by_id(q)
    assert st["nfinite"] == 28 and pe[4] == 0 and pe[9] == 0
    keep = np.ones(30, bool)
    keep[[4, 9]] = False
    sub = po[keep].copy()
    q2 = with_ids(sub)
    restatement.compute_potential(q2, prm)
    np.testing.assert_allclose(pe[keep], by_id(q2), rtol=1e-14)


def test_never_accept_is_direct_sum(restatement):
    cat = catalog.config1()
    po = cat.po[::5].copy()  # 2,200 records, a deep enough tree
    q = with_ids(po)
    restatement.never_accept(True)
    try:
        st = restatement.compute_potential(q, cat.params)
    finally:
        restatement.never_accept(False)
    assert st["p2m"] == 0
    assert _rel(by_id(q), direct_float_diff(po, cat.params)).max() < 1e-12


def test_every_particle_accounted_once(restatement):
    """P2P partners + points behind accepted monopoles = all others, except a
    degenerate (> 20 point) sink leaf, which skips its own points."""
    cat = catalog.config1()
    po = with_ids(cat.po)
    out = restatement.compute_potential_mass(po, cat.params)
    n = po.shape[0]
    tree = out["tree"]
    leaf_n = np.zeros(n, np.int64)
    for nd in tree[tree["left"] < 0]:
        leaf_n[nd["first"]:nd["first"] + nd["n"]] = nd["n"]
    expect = np.where(leaf_n <= 20, n - 1, n - leaf_n)
    np.testing.assert_array_equal(out["np2p"] + out["nmass"], expect)
    assert out["p2m"] > 0


def test_tree_is_the_reference_split(restatement):
    """Every internal node splits its tight box at the float midpoint of the
    widest side (z first, x / y only if strictly wider), left <= lim."""
    cat = catalog.config1()
    po = cat.po.copy()
    tree = restatement.build_tree(po, cat.params)
    pos = po["pos"][:, :3]
    for nd in tree:
        p = pos[nd["first"]:nd["first"] + nd["n"]]
        np.testing.assert_array_equal(nd["lo"], p.min(axis=0))
        np.testing.assert_array_equal(nd["hi"], p.max(axis=0))
        if nd["left"] < 0:
            assert nd["n"] <= 10 or np.all(p == p[0])
            continue
        w = (nd["hi"] - nd["lo"]).astype(np.float32)
        dim = 2
        for k in (0, 1):
            if w[k] > w[dim]:
                dim = k
        lim = np.float32(np.float32(0.5) * np.float32(nd["hi"][dim] + nd["lo"][dim]))
        L = tree[nd["left"]]
        assert np.all(p[:L["n"], dim] <= lim) and np.all(p[L["n"]:, dim] > lim)


def test_reference_bh_error_vs_fp64(restatement):
    """The reference's Barnes-Hut (ERR_TOL 1, monopoles) is ~2e-3 from the
    exact sum; its own arithmetic is ~1e-8 from an FP64 evaluation of the same
    tree.  The second is what '<= 1e-4 vs FP64' can be held to."""
    cat = catalog.config1()
    po = with_ids(cat.po)
    a = po.copy()
    restatement.compute_potential(a, cat.params)
    b = po.copy()
    restatement.compute_potential_bh64(b, cat.params)
    d = restatement.direct64(cat.po, cat.params)
    e_direct = _rel(by_id(a), d)
    e_bh64 = _rel(by_id(a), by_id(b))
    assert 1e-4 < np.median(e_direct) < 1e-2
    assert np.median(e_bh64) < 1e-7 and e_bh64.max() < 1e-6


# ------------------------------------------------------------ kinetic energy

def test_kinetic_closed_forms(restatement):
    prm = catalog.Params()
    assert restatement.hubble_scaling(prm, 0.0) == pytest.approx(1.0, abs=1e-15)
    gc = catalog.GC
    po = np.zeros(4, catalog.POTENTIAL_DTYPE)
    cen = np.array([10.0, 20.0, 30.0], np.float32)
    vc = np.array([100.0, 0.0, -50.0], np.float32)
    po["pos"][:, :3] = cen
    po["pos"][:, 3:] = vc
    po["pos"][0, 3:] += [3.0, 4.0, 0.0]       # peculiar only: |dv|^2 = 25
    po["pos"][1, 0] += 1.0                     # Hubble flow only: dv = 100 km/s
    po["pos"][2, 3:] += [3.0, 4.0, 0.0]
    po["pos"][2, 1] += 0.25                    # both: (3, 4 + 25, 0)
    po["ke"][3] = -1.0                         # potential-only: untouched
    restatement.kinetic(po, vc, cen, 1.0, prm)
    conv = 0.5 / gc  # ke = (float)((double)ke * conv), src/potential.c:209
    assert po["ke"][0] == np.float32(25.0 * conv)
    assert po["ke"][1] == np.float32(1e4 * conv)
    assert po["ke"][2] == np.float32((9.0 + 29.0 ** 2) * conv)
    assert po["ke"][3] == -1.0


def test_kinetic_scale_factor(restatement):
    """At a = 0.5 the Hubble term is H(z=1) a (x - x_c), conversion a / (2 Gc)."""
    prm = catalog.Params(scale_now=0.5)
    po = np.zeros(1, catalog.POTENTIAL_DTYPE)
    po["pos"][0, 2] = 0.5
    restatement.kinetic(po, np.zeros(3, np.float32), np.zeros(3, np.float32), 0.5, prm)
    e = np.sqrt(prm.Om * 8.0 + prm.Ol)
    dv = 100.0 * e * 0.5 * 0.5
    assert po["ke"][0] == pytest.approx(0.5 * 0.5 / catalog.GC * dv * dv, rel=1e-6)


# ------------------------------------------------------------- core velocity

def _corevel_brute(po, rvir_dens):
    """src/properties.c:165-214 restated as an O(n^2) search: two-pass mean
    and variance of every prefix, then the running-minimum selection."""
    n = po.shape[0]
    r2 = po["r2"].astype(np.float32)
    thr = rvir_dens * (4.0 * np.pi / 3.0)
    rv = -1
    for j in range(n - 1, -1, -1):
        if float(j * j) > float(np.float32(r2[j] * r2[j]) * r2[j]) * (thr * thr):
            rv = j
            break
    if rv < 1:
        return None
    cm = -1
    for j in range(n - 1, -1, -1):
        if float(r2[j]) * 100.0 < float(r2[rv]):
            cm = j
            break
    cm = max(cm, 100)
    v = po["pos"][:, 3:].astype(np.float64)
    best = None
    core = bulk = None
    for i in range(rv):
        m = v[:i + 1].mean(axis=0)
        var = float(((v[:i + 1] - m) ** 2).sum())
        val = var / ((i - 3) * i) if i > 3 else 0.0
        if i < 10 or val < best:
            best = val
            if i < cm:
                core = (i, m.copy(), val)
            bulk = (m.copy(), val)
    return core, bulk


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_corevel_brute_force(restatement, seed):
    rng = np.random.default_rng(seed)
    prm = catalog.Params(particle_mass=1e10)
    n = int(rng.integers(150, 400))
    po = np.zeros(n, catalog.POTENTIAL_DTYPE)
    r = np.sort(rng.uniform(0.0, 1.0, n) ** 2 * 0.6)
    po["r2"] = (r * r).astype(np.float32)
    po["pos"][:, 3:] = rng.normal(0, 50.0, (n, 3)) + (1.0 / (0.05 + r))[:, None] * 20.0
    h = np.zeros(1, catalog.HALO_DTYPE)
    restatement.corevel(h, po, prm)
    want = _corevel_brute(po, prm.particle_rvir_dens)
    assert want is not None
    (ic, mc, vc), (mb, vb) = want
    assert int(h["n_core"][0]) == ic
    np.testing.assert_allclose(h["corevel"][0], mc, rtol=1e-6, atol=1e-5)
    np.testing.assert_allclose(h["bulkvel"][0], mb, rtol=1e-6, atol=1e-5)
    np.testing.assert_allclose(h["min_vel_err"][0], vc, rtol=1e-6)
    np.testing.assert_allclose(h["pos"][0, 3:], mc, rtol=1e-6, atol=1e-5)


def test_corevel_no_virial_prefix(restatement):
    prm = catalog.Params(particle_mass=1e9)
    po = np.zeros(50, catalog.POTENTIAL_DTYPE)
    po["r2"] = 100.0  # far too diffuse: rvir_max < 1, the halo is left alone
    po["pos"][:, 3:] = 7.0
    h = np.zeros(1, catalog.HALO_DTYPE)
    h["corevel"] = [1, 2, 3]
    restatement.corevel(h, po, prm)
    np.testing.assert_array_equal(h["corevel"][0], [1, 2, 3])
    assert h["n_core"][0] == 0


def test_bound_pass_driver_matches_reference_path(restatement, reference):
    """The bound pass with the compiled reference's own compute_potential and
    compute_kinetic_energy against the full restatement."""
    cat = catalog.small_batch(seed=21, n_halos=25, lo=15, hi=3000, potential_only=0.15,
                              sub_fraction=0.5)
    a_po, a_h = cat.po.copy(), cat.halos.copy()
    ka = restatement.bound_pass_batch_with(a_po, a_h, cat.params, reference, nthreads=4)
    b_po, b_h = cat.po.copy(), cat.halos.copy()
    kb, _ = restatement.bound_pass_batch(b_po, b_h, cat.params, nthreads=4)
    np.testing.assert_array_equal(ka, kb)
    np.testing.assert_array_equal(a_h["n_core"], b_h["n_core"])
    np.testing.assert_allclose(a_h["corevel"], b_h["corevel"], rtol=1e-6, atol=1e-4)
    assert np.abs(a_h["n_bound"] - b_h["n_bound"]).max() <= 1
