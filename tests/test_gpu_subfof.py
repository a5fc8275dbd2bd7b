"""The phase-space sub-FOF step on the GPU (hp_find_subfofs, replacing
_find_subfofs_better2, src/groupies.c:287-314) against the restatement,
which tests/test_subfof_oracle.py pins bit for bit to the reference's own
compiled code, and against that compiled code directly.

Every comparison is exact:
  * the linking length of each group (find_median_r's result);
  * each group's particles as the call leaves them, byte for byte: normalised
    by norm_sd, in the reference's tree order, linked, and sorted by the
    reference's union-find roots with its partition sort (the particles with
    no neighbour first).  This checks the normalisation, the in-place tree
    order, the components, the roots' numbers and the partition at once;
  * the sub-FOFs (components of at least MIN_HALO_PARTICLES) as (first,
    num_p) in the reference's root order, against the restatement and
    against the reference's own _find_subfofs_better2.
The host-side argument checks run without a GPU.
"""
import ctypes as C

import numpy as np
import pytest

import halopot
import oracle
from halopot import catalog


def _batch(parts_list):
    """Concatenate groups; returns (particles, groups)."""
    parts = np.concatenate(parts_list) if parts_list else np.zeros(0, catalog.PARTICLE_DTYPE)
    g = np.zeros(len(parts_list), catalog.FOF_GROUP_DTYPE)
    off = 0
    for i, p in enumerate(parts_list):
        g["first"][i] = off
        g["num_p"][i] = p.shape[0]
        off += p.shape[0]
    return np.ascontiguousarray(parts), g


def _run_gpu(parts, groups, **kw):
    from halopot.hipmem import DeviceArray

    d = DeviceArray.from_numpy(parts)
    try:
        g, subs = halopot.find_subfofs(d.address, parts.shape[0], groups, **kw)
        out = d.download(np.empty_like(parts))
    finally:
        d.free()
    return g, subs, out


def _expected(restatement, p, top=True, scaling=1.0, exact_ll=0, min_p=10):
    o = restatement.subfof(p, top=top, scaling=scaling, exact_ll=exact_ll, min_p=min_p)
    return o["target_r"], o["final"], [(int(a), int(b)) for a, b in o["fofs"]]


def _check_batch(restatement, parts_list, top=True, scaling=1.0, exact_ll=0, min_p=10):
    parts, groups = _batch(parts_list)
    g, subs, out = _run_gpu(parts, groups, exact_ll_calc=exact_ll,
                            initial_metric_scaling=scaling, min_halo_particles=min_p)
    g["top"] = int(top)
    if not top:
        groups_nt = groups.copy()
        g2 = np.zeros(len(groups_nt), halopot.FOF_GROUP_DTYPE)
        g2["first"], g2["num_p"], g2["top"] = groups_nt["first"], groups_nt["num_p"], 0
        g, subs, out = _run_gpu(parts, g2, exact_ll_calc=exact_ll,
                                initial_metric_scaling=scaling, min_halo_particles=min_p)
    nsub = 0
    for i, p in enumerate(parts_list):
        f, n = int(groups["first"][i]), int(groups["num_p"][i])
        tr, arranged, want = _expected(restatement, p, top, scaling, exact_ll, min_p)
        assert g["target_r"][i] == tr, (i, n)
        assert out[f:f + n].tobytes() == arranged.tobytes(), (i, n)
        got = [(int(s["first"]) - f, int(s["num_p"]))
               for s in subs[g["first_sub"][i]:g["first_sub"][i] + g["num_sub"][i]]]
        assert got == want, (i, n)
        nsub += len(got)
    assert subs.shape[0] == nsub
    return g, subs, out


@pytest.mark.gpu
def test_gpu_subfof_matches_restatement_on_fof_groups(restatement):
    fg = catalog.fof_groups(n_groups=60, lo=200, hi=30000)
    big = catalog.fof_groups(seed=99, n_groups=4, lo=12000, hi=40000)
    plist = [fg.particles[f:f + n].copy() for f, n in fg.groups]
    plist += [big.particles[f:f + n].copy() for f, n in big.groups]
    g, subs, _ = _check_batch(restatement, plist)
    assert subs.shape[0] > len(plist)


@pytest.mark.gpu
def test_gpu_subfof_small_min_halo_particles(restatement):
    """MIN_HALO_PARTICLES 2: every linked component is a sub-FOF, so the
    list holds many small roots in the reference's order."""
    fg = catalog.fof_groups(seed=17, n_groups=12, lo=200, hi=8000)
    plist = [fg.particles[f:f + n].copy() for f, n in fg.groups]
    _, subs, _ = _check_batch(restatement, plist, min_p=2)
    assert subs.shape[0] > 10 * len(plist)


@pytest.mark.gpu
def test_gpu_subfof_large_group_and_dense_clump(restatement):
    """Two paths the seeded batches do not reach.  A group of 150,000
    particles creates 13,647 small FOFs, more than the merge replay holds in
    LDS (12,288), so its union-find table lives in global memory.  A clump of
    6,000 particles inside a sparse cloud links every pair of the clump
    (about 18M pairs, and thousands per leaf chunk of the walk): more than
    the pair list's first allocation (8 per particle) and more than any
    earlier batch of this file left it, so the linking walk runs again with
    room for every pair."""
    big = catalog.fof_groups(seed=41, n_groups=1, lo=150000, hi=150001)
    assert big.particles.shape[0] == 150000
    _check_batch(restatement, [big.particles.copy()])
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.normal(0, 1.0, (14700, 6)), rng.normal(0, 1e-3, (6000, 6))])
    clump = np.zeros(len(x), catalog.PARTICLE_DTYPE)
    clump["id"] = np.arange(len(x))
    clump["pos"] = x.astype(np.float32)
    _, subs, _ = _check_batch(restatement, [clump])
    assert subs["num_p"].max() >= 6000


@pytest.mark.gpu
def test_gpu_subfof_shared_walk_equals_ordered_walk():
    """The default next-closest search (one walk per wavefront, the rare
    close calls repeated in the reference's order) against every search in
    the reference's order: identical linking lengths and results."""
    fg = catalog.fof_groups(seed=31, n_groups=80, lo=200, hi=40000)
    parts, groups = fg.particles, fg.groups
    a = _run_gpu(parts, groups)
    b = _run_gpu(parts, groups, ordered_search=1)
    np.testing.assert_array_equal(a[0]["target_r"], b[0]["target_r"])
    assert a[1].tobytes() == b[1].tobytes() and a[2].tobytes() == b[2].tobytes()


@pytest.mark.gpu
def test_gpu_subfof_edge_groups(restatement):
    rng = np.random.default_rng(21)

    def grp(x):
        p = np.zeros(len(x), catalog.PARTICLE_DTYPE)
        p["id"] = rng.integers(0, 1 << 40, len(x))
        p["pos"] = np.asarray(x, np.float32)
        return p

    cases = [grp(rng.normal(size=(k, 6))) for k in (1, 2, 3, 10, 40, 41, 81, 257)]
    x = rng.normal(size=(300, 6))
    x[100:200] = x[rng.integers(0, 100, 100)]
    cases.append(grp(x))
    x = rng.normal(size=(200, 6))
    x[:, 3:] = 5.0
    cases.append(grp(x))
    cases.append(grp(np.concatenate([rng.normal(0, 0.01, (150, 6)), rng.normal(0, 3, (350, 6))])))
    lat = np.stack(np.meshgrid(*[np.arange(3)] * 6, indexing="ij"), -1).reshape(-1, 6)
    cases.append(grp(lat.astype(np.float64) + 10.0))
    _check_batch(restatement, cases)


@pytest.mark.gpu
def test_gpu_subfof_empty_groups_and_mixed_levels(restatement):
    """Empty groups between others (linking length 0, no sub-FOF; the
    reference asserts on an empty sample) and top-level and lower-level
    groups in one batch (sigma_x scaled by INITIAL_METRIC_SCALING only for
    the top-level ones)."""
    fg = catalog.fof_groups(seed=8, n_groups=5, lo=300, hi=5000)
    plist = [fg.particles[f:f + n].copy() for f, n in fg.groups]
    empty = np.zeros(0, catalog.PARTICLE_DTYPE)
    plist = [empty, plist[0], empty, plist[1], plist[2], empty, plist[3], plist[4]]
    parts, groups = _batch(plist)
    top = np.array([1, 1, 0, 0, 1, 1, 0, 1], np.int32)
    g = np.zeros(len(plist), halopot.FOF_GROUP_DTYPE)
    g["first"], g["num_p"], g["top"] = groups["first"], groups["num_p"], top
    got, subs, out = _run_gpu(parts, g, initial_metric_scaling=1.3)
    for i, p in enumerate(plist):
        f, n = int(groups["first"][i]), int(groups["num_p"][i])
        if n == 0:
            assert got["target_r"][i] == 0.0 and got["num_sub"][i] == 0
            continue
        tr, arranged, want = _expected(restatement, p, bool(top[i]), 1.3, 0)
        assert got["target_r"][i] == tr, i
        assert out[f:f + n].tobytes() == arranged.tobytes(), i
        sl = subs[got["first_sub"][i]:got["first_sub"][i] + got["num_sub"][i]]
        assert [(int(x["first"]) - f, int(x["num_p"])) for x in sl] == want, i


@pytest.mark.gpu
def test_gpu_subfof_exact_ll_and_lower_levels(restatement):
    fg = catalog.fof_groups(seed=5, n_groups=6, lo=300, hi=12000)
    plist = [fg.particles[f:f + n].copy() for f, n in fg.groups]
    _check_batch(restatement, plist, exact_ll=1, scaling=1.5)
    _check_batch(restatement, plist, top=False, scaling=1.5)


@pytest.mark.gpu
def test_gpu_subfof_matches_reference():
    """Against the reference's own _find_subfofs_better2 (oracle/_ref): the
    linking length, the particle array it leaves byte for byte and its
    sub-FOF list in its order."""
    if not oracle.ReferenceProps.available():
        pytest.skip("compiled reference (oracle/_ref) not present")
    ref = oracle.ReferenceProps()
    ref.set_subfof_config()
    fg = catalog.fof_groups(seed=123, n_groups=40, lo=200, hi=20000)
    plist = [fg.particles[f:f + n].copy() for f, n in fg.groups]
    parts, groups = _batch(plist)
    g, subs, out = _run_gpu(parts, groups)
    for i, p in enumerate(plist):
        r = ref.subfof(p)
        assert g["target_r"][i] == r["target_r"], i
        f = int(groups["first"][i])
        mine = [(int(s["first"]) - f, int(s["num_p"]))
                for s in subs[g["first_sub"][i]:g["first_sub"][i] + g["num_sub"][i]]]
        assert mine == [(int(a), int(b)) for a, b in r["fofs"]], i
        assert out[f:f + p.shape[0]].tobytes() == r["final"].tobytes(), i


@pytest.mark.gpu
def test_gpu_subfof_rejects_non_finite(restatement):
    rng = np.random.default_rng(3)
    p = np.zeros(100, catalog.PARTICLE_DTYPE)
    p["pos"] = rng.normal(size=(100, 6))
    p["pos"][17, 4] = np.nan
    parts, groups = _batch([p])
    with pytest.raises(halopot.HalopotError, match="non-finite"):
        _run_gpu(parts, groups)


def test_subfof_rejects_bad_input_on_the_host():
    # checked before any device work, so these run without a GPU
    L = halopot.lib()
    g = np.zeros(2, halopot.FOF_GROUP_DTYPE)
    g["first"] = [0, 50]
    g["num_p"] = [60, 50]
    subs = np.zeros(10, halopot.SUBFOF_DTYPE)
    n = C.c_int64()
    fake = C.c_void_p(1 << 20)
    cfg = halopot.SubfofConfig(0.7, 1.0, 10, 0, 0)

    def call(groups, npart, cfg=cfg):
        return L.hp_find_subfofs(fake, npart, groups.ctypes.data, groups.shape[0],
                                 C.byref(cfg), subs.ctypes.data, subs.shape[0], C.byref(n))
    assert call(g, 200) == -1 and b"overlap" in L.hp_last_error()
    g["num_p"] = [50, 50]
    assert call(g, 90) == -1 and b"outside" in L.hp_last_error()
    bad = halopot.SubfofConfig(1.0, 1.0, 10, 0, 0)
    assert call(g, 200, bad) == -1 and b"FOF_FRACTION" in L.hp_last_error()
