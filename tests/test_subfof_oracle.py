"""The phase-space sub-FOF step on the CPU (_find_subfofs_better2,
src/groupies.c:287-314): norm_sd, the 6D GROUPIES fast3tree rebuilt in place,
next-closest distances of the sample, find_median_r and the linking at that
length (_find_subfofs_at_r with src/fof.c).

* The restatement (orc_subfof, oracle/restate.c) against the reference's own
  code, compiled from src/groupies.c (with src/fast3tree.c inside),
  src/fof.c and src/jacobi.c into oracle/_ref/libref_props.so and called
  through oracle/ref_props.c: the particle array after norm_sd and the
  rebuild (byte for byte: the order the reference's linking loop and its
  sample walk), the sample's distances, the linking length, and every
  sub-FOF of _find_subfofs_better2 as a set of particle ids.  On seeded FOF
  groups of 200 to 73,000 particles (the larger ones sample 10,000 with
  rand_r), with EXACT_LL_CALC, below the top level, and on hand-made groups
  (one to three particles, one leaf, duplicates, no velocity spread).
* The two search steps alone against the reference: next-closest distances
  of random queries, find_median_r on arrays full of ties under many seeds.
* Closed forms: the next-closest distance is the distance to the nearest
  other point (brute force) when no point repeats; the components are the
  connected components of the pair graph (brute force).
"""
import numpy as np
import pytest

import oracle
from halopot import catalog


@pytest.fixture(scope="module")
def ref():
    if not oracle.ReferenceProps.available():
        pytest.skip("compiled reference (oracle/_ref) not present")
    r = oracle.ReferenceProps()
    r.set_subfof_config()
    return r


def _sets_from_labels(label, ids, min_p=10):
    d = {}
    for lab, i in zip(label, ids):
        if lab >= 0:
            d.setdefault(int(lab), []).append(int(i))
    return sorted(tuple(sorted(v)) for v in d.values() if len(v) >= min_p)


def _ref_sets(r):
    return sorted(tuple(sorted(r["final"]["id"][a:a + b])) for a, b in r["fofs"])


def _compare(ref, restatement, p, **kw):
    top = kw.get("top", True)
    r = ref.subfof(p, top=top)
    o = restatement.subfof(p, top=top, exact_ll=kw.get("exact_ll", 0))
    assert r["tree_order"].tobytes() == o["tree_order"].tobytes()
    np.testing.assert_array_equal(r["sample_r"], o["sample_r"])
    assert r["target_r"] == o["target_r"]
    assert _ref_sets(r) == _sets_from_labels(o["label"], o["tree_order"]["id"])
    assert np.array_equal(r["fofs"], r["fofs_at_r"])   # the harness's own sequence
    # the reference's output order: particles sorted by union-find root, the
    # sub-FOFs in root order
    assert r["final"].tobytes() == o["final"].tobytes()
    assert np.array_equal(r["fofs"], o["fofs"])
    return r, o


def test_restatement_matches_reference_on_fof_groups(ref, restatement):
    fg = catalog.fof_groups(n_groups=60, lo=200, hi=30000)
    nsub = 0
    for f, n in fg.groups:
        r, _ = _compare(ref, restatement, fg.particles[f:f + n].copy())
        nsub += len(r["fofs"])
    assert nsub > len(fg.groups) and (fg.groups["num_p"] > 10000).sum() >= 2


def test_restatement_matches_reference_on_sampled_groups(ref, restatement):
    """Groups above 10,000 particles: the sample is drawn with rand_r from the
    tree order, so it matches only if the in-place partition does."""
    fg = catalog.fof_groups(seed=99, n_groups=4, lo=12000, hi=40000)
    for f, n in fg.groups:
        r, o = _compare(ref, restatement, fg.particles[f:f + n].copy())
        assert r["sample_r"].shape[0] == 10000


def test_restatement_matches_reference_exact_ll_and_lower_levels(ref, restatement):
    fg = catalog.fof_groups(seed=5, n_groups=6, lo=300, hi=12000)
    ref.set_subfof_config(exact_ll_calc=1, initial_metric_scaling=1.5)
    try:
        for f, n in fg.groups:
            p = fg.particles[f:f + n].copy()
            r = ref.subfof(p, top=True)
            o = restatement.subfof(p, top=True, scaling=1.5, exact_ll=1)
            assert r["tree_order"].tobytes() == o["tree_order"].tobytes()
            assert r["target_r"] == o["target_r"] and r["sample_r"].shape[0] == n
            assert _ref_sets(r) == _sets_from_labels(o["label"], o["tree_order"]["id"])
            assert r["final"].tobytes() == o["final"].tobytes()
            assert np.array_equal(r["fofs"], o["fofs"])
    finally:
        ref.set_subfof_config()
    for f, n in fg.groups[:3]:
        _compare(ref, restatement, fg.particles[f:f + n].copy(), top=False)


def _group(x):
    p = np.zeros(len(x), catalog.PARTICLE_DTYPE)
    p["id"] = np.arange(len(x)) * 7 + 3
    p["pos"] = np.asarray(x, np.float32)
    return p


def test_restatement_matches_reference_on_edge_groups(ref, restatement):
    rng = np.random.default_rng(11)
    cases = [rng.normal(size=(k, 6)) for k in (1, 2, 3, 10, 40, 41, 81)]
    # duplicates: a third of the points repeated, some several times
    x = rng.normal(size=(300, 6))
    x[100:200] = x[rng.integers(0, 100, 100)]
    cases.append(x)
    # no velocity spread (sig_v = 0: norm_sd leaves the centred coordinates)
    x = rng.normal(size=(200, 6))
    x[:, 3:] = 5.0
    cases.append(x)
    # a dense clump and a thin cloud, and points on a lattice (ties everywhere)
    cases.append(np.concatenate([rng.normal(0, 0.01, (150, 6)), rng.normal(0, 3, (350, 6))]))
    g = np.stack(np.meshgrid(*[np.arange(3)] * 6, indexing="ij"), -1).reshape(-1, 6)
    cases.append(g.astype(np.float64) + 10.0)
    for x in cases:
        _compare(ref, restatement, _group(x))


def test_next_closest_matches_reference(ref, restatement):
    rng = np.random.default_rng(12)
    pts = _group(np.concatenate([rng.normal(size=(3000, 6)), rng.normal(size=(20, 6))]))
    pts["pos"][3000:] = pts["pos"][:20]                       # duplicates
    q = np.concatenate([pts["pos"][::7], rng.normal(size=(300, 6)).astype(np.float32)])
    a = ref.next_closest(pts.copy(), q)
    b = restatement.next_closest(pts, q)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("seed", [1, 2, 3, 77, 1000, 2 ** 31 + 5])
def test_find_median_r_matches_reference(ref, restatement, seed):
    rng = np.random.default_rng(seed % 1000)
    for n in (1, 2, 3, 17, 500, 10000):
        for ties in (0, 5, 50):
            v = rng.uniform(0, 1, n).astype(np.float32)
            if ties:
                v = np.round(v * ties).astype(np.float32) / ties
            for frac in (0.7, 0.5, 0.05, 0.999):
                assert ref.find_median_r(v, frac, seed) == restatement.find_median_r(v, frac, seed)


def test_next_closest_is_the_nearest_other_point(restatement):
    """With no repeated point, fast3tree_find_next_closest_distance is the
    distance to the nearest other point (its leaf-box start and the
    sqrt(r * r) round trips move it by at most an ulp or two)."""
    rng = np.random.default_rng(13)
    pts = _group(rng.normal(size=(2500, 6)) * np.array([1, 1, 1, 3, 3, 3]))
    got = restatement.next_closest(pts, pts["pos"])
    x = pts["pos"].astype(np.float64)
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    want = np.sqrt(d2.min(1))
    assert np.all(np.abs(got - want) <= 4e-7 * want)


def test_components_are_the_pair_graph_components(restatement):
    """The linking joins every pair closer than the linking length; with the
    restatement's labels that is the connected components of that graph
    (brute force over all pairs, float distances as the reference's)."""
    rng = np.random.default_rng(14)
    x = np.concatenate([rng.normal(0, 0.3, (200, 6)), rng.normal(3, 0.3, (150, 6)),
                        rng.normal(0, 4, (250, 6))])
    o = restatement.subfof(_group(x))
    t = o["tree_order"]["pos"]
    r2 = np.float32(o["target_r"]) * np.float32(o["target_r"])
    d = np.zeros((t.shape[0], t.shape[0]), np.float32)
    for j in range(6):                       # the reference's float sum order
        dx = t[:, None, j] - t[None, :, j]
        d = d + dx * dx
    adj = d < r2
    np.fill_diagonal(adj, False)
    n = t.shape[0]
    lab = np.arange(n)
    changed = True
    while changed:                           # min-label propagation
        nl = np.where(adj, lab[None, :], n).min(1)
        nl = np.minimum(nl, lab)
        changed = not np.array_equal(nl, lab)
        lab = nl
    want = np.where(adj.any(1), lab, -1)
    np.testing.assert_array_equal(o["label"], want)
    assert (want >= 0).sum() > 300
