"""GPU parity of compute_potential (libhalopot through its C ABI) against the
oracle: the restatement of src/potential.c, the compiled reference where it
was built, and the FP64 evaluations.

Tolerances:
  * tree: identical (node boxes, counts, centres bit for bit; record order in
    tree order identical, the order the reference's in-place partitions
    leave; dmin within one float ulp, the f64 sums of the big nodes being
    reduced by tiles);
  * compute_potential permutes the records as the reference's does;
  * interaction counts (P2P pairs, P2M terms): identical to the restatement's,
    i.e. every acceptance decision is the reference's;
  * pe vs the restatement / reference Barnes-Hut: max relative 1e-5
    (FP32 pair terms, compensated sums);
  * pe vs the FP64 evaluation of the same tree: median 1e-6, max 1e-5.
"""
import numpy as np
import pytest

import halopot
from halopot import catalog
from conftest import by_id, with_ids
from gpu_util import rel, run_plan

pytestmark = pytest.mark.gpu


def _compare_trees(gnodes, perm, onodes, opo):
    """Walks both trees from the root in parallel."""
    stack = [(0, 0)]
    seen = 0
    while stack:
        g, o = stack.pop()
        G, O = gnodes[g], onodes[o]
        seen += 1
        assert G["count"] == O["n"], (g, o)
        assert G["first"] == O["first"], (g, o)
        np.testing.assert_array_equal(G["lo"], O["lo"])
        np.testing.assert_array_equal(G["hi"], O["hi"])
        np.testing.assert_array_equal(G["com"], O["com"])
        assert G["mass"] == O["m"]
        d = np.float32(O["dmin"])
        assert abs(int(np.float32(G["dmin"]).view(np.int32)) - int(d.view(np.int32))) <= 1
        assert (G["left"] < 0) == (O["left"] < 0), (g, o)
        if G["left"] >= 0:
            stack.append((G["left"], O["left"]))
            stack.append((G["left"] + 1, O["left"] + 1))
    # every oracle node is reached; the GPU array may also hold unused
    # (empty) ids reserved by subtree workgroups
    assert seen == len(onodes)
    reach = np.zeros(len(gnodes), bool)
    stack = [0]
    while stack:
        g = stack.pop()
        reach[g] = True
        if gnodes[g]["left"] >= 0:
            stack += [gnodes[g]["left"], gnodes[g]["left"] + 1]
    assert np.all(gnodes["count"][~reach] == 0)
    np.testing.assert_array_equal(perm, opo["flags"])


@pytest.mark.parametrize("which", ["c1", "small"])
def test_tree_identical(restatement, which):
    cat = catalog.config1() if which == "c1" else catalog.small_batch(n_halos=1, lo=2000, hi=2001)
    prm = cat.params
    halopot.set_config(prm)
    po = with_ids(cat.segment(0))
    gnodes, perm = halopot.debug_tree(po)
    opo = po.copy()
    onodes = restatement.build_tree(opo, prm)
    _compare_trees(gnodes, perm, onodes, opo)


@pytest.mark.parametrize("n", [11, 37, 64, 65, 256, 257, 300, 512, 513, 1024, 1025])
def test_tree_identical_small_calls(restatement, n):
    """Calls around the sizes where subtrees move from one wavefront (<= 64
    points) to a workgroup, between the workgroup kernel's three sizes (256,
    512 and 1,024 points), and to the split levels above 1,024."""
    cat = catalog.config1()
    prm = cat.params
    halopot.set_config(prm)
    po = with_ids(cat.segment(0)[:n])
    gnodes, perm = halopot.debug_tree(po)
    opo = po.copy()
    onodes = restatement.build_tree(opo, prm)
    _compare_trees(gnodes, perm, onodes, opo)


def test_tree_identical_coincident_points(restatement):
    """31 of 40 points coincide: the split that cannot separate them leaves a
    31-point leaf (src/fast3tree.c:805-809), whose self term the walk skips
    (more than 20 points, src/potential.c:144-146)."""
    cat = catalog.config1()
    prm = cat.params
    halopot.set_config(prm)
    po = with_ids(cat.segment(0)[:40])
    po["pos"][10:, :3] = po["pos"][0, :3]
    gnodes, perm = halopot.debug_tree(po)
    opo = po.copy()
    onodes = restatement.build_tree(opo, prm)
    _compare_trees(gnodes, perm, onodes, opo)
    assert np.any((gnodes["left"] < 0) & (gnodes["count"] == 31))
    g = po.copy()
    halopot.compute_potential(g)
    o = po.copy()
    restatement.compute_potential(o, prm)  # permutes o
    np.testing.assert_array_equal(g["flags"], o["flags"])
    np.testing.assert_allclose(g["pe"], o["pe"], rtol=1e-5)


def test_c1_potential_parity(restatement, request):
    cat = catalog.config1()
    prm = cat.params
    halopot.set_config(prm)
    po = with_ids(cat.po)
    g = po.copy()
    halopot.compute_potential(g)
    # the records are permuted as the reference's tree build permutes po
    # (src/fast3tree.c:762-773), every field travelling with its record
    o = po.copy()
    restatement.compute_potential(o, prm)
    np.testing.assert_array_equal(g["flags"], o["flags"])
    for f in ("pos", "r2", "ke"):
        np.testing.assert_array_equal(g[f], po[f][g["flags"]])
    r = rel(g["pe"], o["pe"])
    assert np.max(r) < 1e-5, np.max(r)
    b64 = po.copy()
    restatement.compute_potential_bh64(b64, prm)
    r64 = rel(by_id(g), by_id(b64))
    assert np.median(r64) < 1e-6 and np.max(r64) < 1e-5, (np.median(r64), np.max(r64))
    try:
        ref = request.getfixturevalue("reference")
    except pytest.skip.Exception:
        return
    a = po.copy()
    ref.compute_potential(a, prm)
    np.testing.assert_array_equal(g["flags"], a["flags"])
    rr = rel(g["pe"], a["pe"])
    assert np.max(rr) < 1e-5, np.max(rr)


@pytest.mark.parametrize("acc", [halopot.ACC_COMPENSATED, halopot.ACC_NAIVE, halopot.ACC_F64])
def test_batch_counts_and_accumulation(restatement, acc):
    cat = catalog.small_batch(seed=7, n_halos=60, lo=2, hi=5000)
    prm = cat.params
    for h in range(cat.n_halos):  # ids local to each call, as with_ids gives them
        f, c = int(cat.halos["first"][h]), int(cat.halos["count"][h])
        cat.po["flags"][f:f + c] = np.arange(c, dtype=np.int32)
    po_out, _, _, st = run_plan(cat, halopot.options(accumulate=acc, count_work=1),
                                potential_only=True)
    want_p2p = want_p2m = 0
    for h in range(cat.n_halos):
        seg = with_ids(cat.segment(h))
        o = seg.copy()
        s = restatement.compute_potential(o, prm)
        want_p2p += s["p2p"]
        want_p2m += s["p2m"]
        f = int(cat.halos["first"][h])
        out = po_out[f:f + seg.shape[0]]
        np.testing.assert_array_equal(out["flags"], o["flags"])  # the reference's permutation
        got = by_id(out)
        if acc == halopot.ACC_F64:
            # every term in FP64 on the same tree: the FP64 evaluation of the
            # reference's decisions (its own P2M quotient is a float)
            b = seg.copy()
            restatement.compute_potential_bh64(b, prm)
            r = rel(got, by_id(b))
        else:
            r = rel(got, by_id(o))
        tol = {halopot.ACC_COMPENSATED: 1e-5, halopot.ACC_NAIVE: 1e-4, halopot.ACC_F64: 1e-12}[acc]
        assert np.max(r, initial=0) < tol, (h, np.max(r))
    assert st["p2p"] == want_p2p
    assert st["p2m"] == want_p2m


def test_edge_cases(restatement):
    prm = catalog.Params(particle_mass=1e9)
    halopot.set_config(prm)
    rng = np.random.default_rng(5)
    cases = []
    # two particles: pe = m / max(d, FORCE_RES) each
    two = np.zeros(2, catalog.POTENTIAL_DTYPE)
    two["pos"][1, 0] = 0.5
    cases.append(two)
    # closer than the softening
    near = two.copy()
    near["pos"][1, 0] = 0.001
    cases.append(near)
    # 15 coincident points (<= 20: every pair at FORCE_RES) and 30 (> 20:
    # a degenerate leaf, no self term)
    for n in (15, 30):
        c = np.zeros(n, catalog.POTENTIAL_DTYPE)
        c["pos"][:, :3] = 3.25
        cases.append(c)
    # non-finite points are left out and get pe = 0
    nf = np.zeros(40, catalog.POTENTIAL_DTYPE)
    nf["pos"][:, :3] = rng.uniform(0, 1, (40, 3))
    nf["pos"][3, 1] = np.nan
    nf["pos"][17, 0] = np.inf
    cases.append(nf)
    # single particle, all-NaN
    one = np.zeros(1, catalog.POTENTIAL_DTYPE)
    cases.append(one)
    for c in cases:
        g = with_ids(c)
        halopot.compute_potential(g)
        o = with_ids(c)
        restatement.compute_potential(o, prm)
        np.testing.assert_array_equal(g["flags"], o["flags"])
        np.testing.assert_allclose(g["pe"], o["pe"], rtol=1e-6, atol=0)
    assert cases[0]["pe"].sum() == 0  # inputs untouched
    g = with_ids(cases[0])
    halopot.compute_potential(g)
    np.testing.assert_allclose(g["pe"], [1e9 / 0.5] * 2, rtol=1e-6)
    g = with_ids(cases[2])
    halopot.compute_potential(g)
    np.testing.assert_allclose(g["pe"], 14 * 1e9 / 0.003, rtol=1e-6)
    g = with_ids(cases[3])
    halopot.compute_potential(g)
    assert np.all(g["pe"] == 0)
    g = with_ids(cases[4])
    halopot.compute_potential(g)
    pe = by_id(g)
    assert pe[3] == 0 and pe[17] == 0 and np.all(pe[[0, 1, 2, 4]] > 0)
    # non-finite records end at the tail (src/fast3tree.c:875-887)
    assert set(g["flags"][-2:]) == {3, 17}


def _shell_chain(shells=100, per_shell=24, core=8, seed=3):
    """Points on nested shells along x: shell k holds per_shell points with x
    in [0.52, 0.98] * 2^(k - 49) (one at 0.98), y and z in +-0.48 of that
    scale; a core of `core` points lies below shell 0.  Every split of the
    chain is the plane between two shells, so the tree is shells + a few
    levels deep, and each shell node is close enough to the core to be
    opened by the core's leaves: the walk of the first 64 points keeps
    every shell pending at once."""
    rng = np.random.default_rng(seed)
    s0 = 2.0 ** -50
    parts = [np.column_stack([rng.uniform(0.05, 0.45, core) * s0,
                              rng.uniform(-0.2, 0.2, (core, 2)) * s0])]
    for k in range(shells):
        s = 2.0 ** (k + 1) * s0
        x = rng.uniform(0.52, 0.95, per_shell) * s
        x[0] = 0.98 * s
        parts.append(np.column_stack([x, rng.uniform(-0.48, 0.48, (per_shell, 2)) * s]))
    pos = np.concatenate(parts)
    po = catalog._records(pos, np.zeros_like(pos), np.zeros(3))
    return with_ids(po)


def _pending_pairs(nodes, npts=64):
    """Deepest stack of pending sibling pairs in the pair walk of the first
    npts points in tree order (the restatement's tree; acceptance in FP64,
    the shells being far from the threshold)."""
    leaves = np.nonzero(nodes["left"] < 0)[0]
    leaf_of = np.empty(int(nodes["n"][0]), np.int64)
    for lf in leaves:
        leaf_of[nodes["first"][lf]:nodes["first"][lf] + nodes["n"][lf]] = lf
    lanes = leaf_of[:npts]
    com = nodes["com"].astype(np.float64)
    dmin2 = nodes["dmin"].astype(np.float64) ** 2

    def opened(i, m):
        if nodes["left"][i] < 0:
            return set()
        out = set()
        for ln in m:
            lf = lanes[ln]
            inside = nodes["first"][i] <= nodes["first"][lf] < nodes["first"][i] + nodes["n"][i]
            if inside or not np.sum((com[lf] - com[i]) ** 2) > dmin2[i]:
                out.add(ln)
        return out

    stack, deepest = [], 0
    c, m = int(nodes["left"][0]), set(range(npts))
    while True:
        o0, o1 = opened(c, m), opened(c + 1, m)
        if o1:
            stack.append((int(nodes["left"][c + 1]), o1))
            deepest = max(deepest, len(stack))
        if o0:
            c, m = int(nodes["left"][c]), o0
        elif stack:
            c, m = stack.pop()
        else:
            return deepest


@pytest.mark.parametrize("walk", [halopot.WALK_PAIRS, halopot.WALK_NODES])
def test_deep_tree_walks(restatement, walk):
    """A tree about 100 levels deep whose first wavefront keeps more sibling
    pairs pending than the pair walk's 64-entry stack holds: that wavefront
    finishes on the node-by-node walk, the others on pairs.  Both walks take
    the restatement's decisions (equal P2P and P2M counts) and give its pe."""
    prm = catalog.Params(particle_mass=1e9)
    halopot.set_config(prm)
    po = _shell_chain()
    o = po.copy()
    s = restatement.compute_potential(o, prm, tree=True)
    nodes = s["tree"]
    assert int(nodes["depth"].max()) > 90
    assert _pending_pairs(nodes) > 64
    cat = catalog.Catalogue("shell_chain", prm, po.copy(),
                            catalog._halo_rows([0], [po.shape[0]], [[0, 0, 0]], [[0, 0, 0]], [0]))
    out, _, _, st = run_plan(cat, halopot.options(count_work=1, walk=walk), potential_only=True)
    assert (st["p2p"], st["p2m"]) == (s["p2p"], s["p2m"])
    np.testing.assert_array_equal(out["flags"], o["flags"])
    r = rel(out["pe"], o["pe"])
    assert np.max(r) < 1e-5, np.max(r)


def test_per_call_plan_on_many_small_calls():
    """A per-call plan reserves its buffers for its largest call but loads
    the whole batch's tables as well (each record's call is written for the
    whole batch): 400 calls of 20 to 200 records, 100 times the largest
    call.  The per-call run gives the batched plan's record order, bound
    flags and counts, and its potentials to float rounding."""
    cat = catalog.small_batch(seed=91, n_halos=400, lo=20, hi=200, potential_only=0.05)
    cat.po["flags"] = np.arange(cat.n_records, dtype=np.int32)
    assert cat.n_records > 50 * int(cat.halos["count"].max())
    want = run_plan(cat)
    got = run_plan(cat, halopot.options(per_call=1))
    np.testing.assert_array_equal(got[0]["flags"], want[0]["flags"])
    np.testing.assert_array_equal(got[2], want[2])
    np.testing.assert_array_equal(got[1]["n_bound"], want[1]["n_bound"])
    np.testing.assert_allclose(got[0]["pe"], want[0]["pe"], rtol=1e-6, atol=0)
