"""Record assembly (calc_particle_radii / _reset_potentials, src/groupies.c:
212-268): each call's struct potential records built from the FOF group's
particle copies and the halo hierarchy.

CPU:
  * the restatement (orc_particle_radii) against the reference's own
    calc_particle_radii (oracle/_ref/libref_props.so through
    oracle/ref_props.c): every record byte for byte, on seeded FOF groups and
    on hand-made hierarchies (a child link back to an ancestor, a halo in two
    sibling lists, empty halos, host ratio 0 and 1e9);
  * the record count of a tree-shaped hierarchy is num_child_particles of the
    halo, or of its parent when the parent's subtree joins as potential-only
    records (src/groupies.c:260-265), and the flags follow src/groupies.c:
    228-231;
  * hp_gather_calls rejects bad input on the host (before any device work).
GPU: hp_gather_calls against the restatement, byte for byte, and the
gathered batch through the bound pass.
"""
import ctypes as C

import numpy as np
import pytest

import halopot
import oracle
from halopot import catalog


@pytest.fixture(scope="module")
def refprops():
    if not oracle.ReferenceProps.available():
        pytest.skip("compiled reference (oracle/_ref) not present")
    return oracle.ReferenceProps()


@pytest.fixture(scope="module")
def groups():
    return catalog.hierarchy(n_groups=120)


def _tree(rows, npart):
    """rows: (num_p, sub_of, children) per halo; particles laid out in order."""
    t = np.zeros(len(rows), catalog.TREE_DTYPE)
    start = 0
    for h, (n, parent, kids) in enumerate(rows):
        t["p_start"][h] = start
        start += n
        t["num_p"][h] = n
        t["sub_of"][h] = parent
        t["child"][h] = kids[0] if kids else -1
        t["next_cochild"][h] = -1
    for h, (n, parent, kids) in enumerate(rows):
        for a, b in zip(kids[:-1], kids[1:]):
            t["next_cochild"][a] = b
    assert start <= npart
    t["pos"][:, :3] = np.arange(len(rows))[:, None] * 0.01
    return t


def _ncp(t):
    """num_child_particles as src/properties.c:139-153 computes it."""
    out = np.zeros(t.shape[0], np.int64)
    seen = set()

    def f(h, stack):
        if h in stack:
            return 0
        n = int(t["num_p"][h])
        c = int(t["child"][h])
        while c > -1:
            n += f(c, stack | {h})
            c = int(t["next_cochild"][c])
        return n
    for h in range(t.shape[0]):
        out[h] = f(h, seen)
    return out


def _particles(n, seed=3):
    rng = np.random.default_rng(seed)
    p = np.zeros(n, catalog.PARTICLE_DTYPE)
    p["id"] = np.arange(n)
    p["pos"] = rng.normal(0.0, 1.0, (n, 6)).astype(np.float32)
    return p


def test_restatement_matches_reference_on_groups(restatement, refprops, groups):
    cap = 2 * groups.particles.shape[0]
    for b in groups.base:
        a = restatement.particle_radii(groups.particles, groups.tree, int(b), 0.3, cap)
        r = refprops.particle_radii(groups.particles, groups.tree, int(b), 0.3, cap)
        assert a.tobytes() == r.tobytes(), b


def test_record_counts_and_flags(restatement, groups):
    t = groups.tree
    ncp = t["num_child_particles"]
    cap = 2 * groups.particles.shape[0]
    n_parent = 0
    for b in groups.base:
        b = int(b)
        po = restatement.particle_radii(groups.particles, t, b, 0.3, cap)
        par = int(t["sub_of"][b])
        joined = par > -1 and ncp[par] * 0.3 < ncp[b]
        n_parent += joined
        assert po.shape[0] == (ncp[par] if joined else ncp[b]), b
        own = slice(0, int(t["num_p"][b]))
        assert np.all(po["flags"][own] == 1)
        np.testing.assert_array_equal(po["pos"][own],
                                      groups.particles["pos"][t["p_start"][b]:][:t["num_p"][b]])
        assert np.all(po["ke"][: ncp[b]] == 0) and np.all(po["ke"][ncp[b]:] == -1)
        assert np.all(po["pe"] == 0)
        d = po["pos"][:, :3] - t["pos"][b, :3]
        r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        np.testing.assert_array_equal(po["r2"], r2.astype(np.float32))
        small = po["flags"][: ncp[b]] == 2
        assert np.all(po["flags"][ncp[b]:] == 0)
        assert small.sum() <= ncp[b] - t["num_p"][b]
    assert n_parent > 10


def test_restatement_matches_reference_on_odd_hierarchies(restatement, refprops):
    parts = _particles(4000)
    cases = []
    # 0 host with children 1, 2; 1 has child 3; 3's child list points back at 1
    t = _tree([(500, -1, [1, 2]), (300, 0, [3]), (40, 0, []), (200, 1, [1])], 4000)
    cases.append(t)
    # halo 2 listed under both 0 and 1 (its particles are gathered twice)
    t = _tree([(600, -1, [1, 2]), (400, 0, [2]), (100, 0, [])], 4000)
    cases.append(t)
    # empty halos anywhere
    t = _tree([(800, -1, [1, 2, 3]), (0, 0, []), (300, 0, [4]), (0, 0, []), (0, 2, [])], 4000)
    cases.append(t)
    for t in cases:
        t["num_child_particles"] = _ncp(t)
        for ratio in (0.0, 0.3, 1e9):
            for b in range(t.shape[0]):
                a = restatement.particle_radii(parts, t, b, ratio, 8000)
                r = refprops.particle_radii(parts, t, b, ratio, 8000)
                assert a.tobytes() == r.tobytes(), (b, ratio)


def _gather(t, base, npart, cap):
    L = halopot.lib()
    b = np.ascontiguousarray(base, np.int64)
    calls = np.zeros(max(1, b.shape[0]), catalog.HALO_DTYPE)
    n = C.c_int64()
    # non-null device addresses that are never touched: the host checks
    # reject these inputs before any device work
    fake = 1 << 20
    return L.hp_gather_calls(t.ctypes.data, t.shape[0], b.ctypes.data, b.shape[0], fake, npart,
                             0.3, fake, cap, calls.ctypes.data, C.byref(n))


def test_gather_rejects_bad_input_on_the_host():
    # checked before any device work, so these run without a GPU
    t = _tree([(10, -1, [1]), (5, 0, [])], 100)
    t["num_child_particles"] = _ncp(t)
    err = halopot.lib().hp_last_error
    assert _gather(t, [2], 100, 100) == -1 and b"out of range" in err()    # base
    assert _gather(t, [0], 100, 10) == -1 and b"do not fit" in err()       # capacity
    assert _gather(t, [0], 12, 100) == -1 and b"outside" in err()          # particles
    bad = t.copy()
    bad["child"][1] = 7
    assert _gather(bad, [0], 100, 100) == -1 and b"out of range" in err()  # link
    cyc = _tree([(10, -1, [1, 2]), (5, 0, []), (5, 0, [])], 100)
    cyc["next_cochild"][2] = 1                                    # siblings back to the first
    assert _gather(cyc, [0], 100, 100) == -1 and b"sibling" in err()
    loop = _tree([(10, -1, [1, 2, 3]), (5, 0, []), (5, 0, []), (5, 0, [])], 100)
    loop["next_cochild"][3] = 2                                   # 2 -> 3 -> 2 -> ...
    assert _gather(loop, [0], 100, 100) == -1 and b"does not end" in err()


@pytest.mark.gpu
def test_gpu_gather_matches_restatement(restatement, groups):
    from halopot.hipmem import DeviceArray

    cap = 2 * groups.particles.shape[0]
    d_parts = DeviceArray.from_numpy(groups.particles)
    d_po = DeviceArray.zeros(cap * 48)
    try:
        calls, n = halopot.gather_calls(groups.tree, groups.base, d_parts.address,
                                        groups.particles.shape[0], d_po.address, cap)
        po = d_po.download(np.empty(n, catalog.POTENTIAL_DTYPE))
    finally:
        d_parts.free()
        d_po.free()
    first = 0
    for i, b in enumerate(groups.base):
        want = restatement.particle_radii(groups.particles, groups.tree, int(b), 0.3, cap)
        assert calls["first"][i] == first and calls["count"][i] == want.shape[0]
        assert po[first:first + want.shape[0]].tobytes() == want.tobytes(), b
        np.testing.assert_array_equal(calls["pos"][i], groups.tree["pos"][b])
        assert calls["is_sub"][i] == (groups.tree["sub_of"][b] > -1)
        first += want.shape[0]
    assert n == first


@pytest.mark.gpu
def test_gpu_gathered_batch_through_the_bound_pass(restatement, groups):
    """Gather on the device, then the bound pass on the gathered records: the
    kept counts are the records that are not potential-only, and the result
    agrees with the restatement's bound pass on the restatement's records."""
    from gpu_util import rel
    from halopot.hipmem import DeviceArray, synchronize

    halopot.set_config(groups.params)
    cap = 2 * groups.particles.shape[0]
    d_parts = DeviceArray.from_numpy(groups.particles)
    d_po = DeviceArray.zeros(cap * 48)
    try:
        calls, n = halopot.gather_calls(groups.tree, groups.base, d_parts.address,
                                        groups.particles.shape[0], d_po.address, cap)
        calls["corevel"] = groups.tree["pos"][groups.base, 3:]
        calls["bulkvel"] = groups.tree["pos"][groups.base, 3:]
        d_h = DeviceArray.from_numpy(calls)
        d_b = DeviceArray.zeros(max(1, n))
        plan = halopot.Plan(calls)
        try:
            plan.run(d_po.address, d_h.address, d_b.address)
            synchronize()
        finally:
            plan.close()
        h_g = d_h.download(np.empty_like(calls))
        d_h.free()
        d_b.free()
    finally:
        d_parts.free()
        d_po.free()
    for i, b in enumerate(groups.base[:200]):
        po = restatement.particle_radii(groups.particles, groups.tree, int(b), 0.3, cap)
        row = calls[i:i + 1].copy()
        kept, _ = restatement.bound_pass(po, row, groups.params)
        assert h_g["kept"][i] == kept == np.count_nonzero(
            restatement.particle_radii(groups.particles, groups.tree, int(b), 0.3, cap)["ke"] >= 0)
        assert h_g["n_core"][i] == row["n_core"][0], b
        assert np.max(rel(h_g["corevel"][i], row["corevel"][0])) < 1e-5, b
