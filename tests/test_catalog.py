"""The seeded synthetic catalogues: layouts, determinism, the configurations'
shapes, and the physics pieces they are built from (NFW inversion, Jeans
dispersion, the reference's r2 rounding)."""
import math

import numpy as np
import pytest

from halopot import catalog


def test_nfw_inversion_round_trip():
    c = np.array([2.0, 6.0, 10.0, 25.0])
    p = np.linspace(0.001, 0.999, 50)[:, None] * catalog.nfw_m(c)[None, :]
    x = catalog.nfw_invert_mass(p)
    np.testing.assert_allclose(catalog.nfw_m(x), p, rtol=1e-10)
    assert np.all(x <= c[None, :] * (1 + 1e-9))


def test_jeans_dispersion_against_quadrature():
    """sigma^2 / (G M_s / r_s) of an isotropic NFW sphere,
    x (1+x)^2 int_x^inf m(t) / (t^3 (1+t)^2) dt: the table against adaptive
    quadrature; it peaks near x = 0.76 (Lokas & Mamon 2001)."""
    from scipy.integrate import quad

    xs = np.array([1e-3, 0.1, 0.76, 3.0, 30.0])
    want = []
    for x in xs:
        v, _ = quad(lambda t: catalog.nfw_m(t) / (t ** 3 * (1 + t) ** 2), x, np.inf, limit=400)
        want.append(x * (1 + x) ** 2 * v)
    np.testing.assert_allclose(catalog.jeans_sigma2(xs), want, rtol=2e-4)
    grid = np.linspace(0.3, 1.5, 121)
    assert abs(grid[np.argmax(catalog.jeans_sigma2(grid))] - 0.76) < 0.05


def test_config1_shape_and_determinism():
    a = catalog.config1()
    b = catalog.config1()
    assert a.n_records == 11000 and a.n_halos == 1
    assert a.po.tobytes() == b.po.tobytes()
    assert a.po.dtype.itemsize == 48
    r = np.sqrt(((a.po["pos"][:10000, :3] - a.halos["pos"][0, :3]) ** 2).sum(axis=1))
    assert r.max() <= 1.0 + 1e-5  # truncated at r_vir
    # half-mass radius of an NFW c=10 halo: m(x) = m(10) / 2 at x ~ 3.3
    x_half = float(catalog.nfw_invert_mass(catalog.nfw_m(10.0) / 2))
    assert abs(np.median(r) - x_half * 0.1) < 0.02


def test_r2_rounds_like_the_reference():
    """_reset_potentials (src/groupies.c:220-225): float difference, float
    squares summed left to right."""
    cat = catalog.config1()
    p = cat.po["pos"][:, :3]
    c = cat.halos["pos"][0, :3].astype(np.float32)
    for i in range(0, 11000, 997):
        acc = np.float32(0.0)
        for k in range(3):
            d = np.float32(p[i, k] - c[k])
            acc = np.float32(acc + np.float32(d * d))
        assert cat.po["r2"][i] == acc


def test_config2_batch():
    cat = catalog.config2()
    assert cat.n_halos == 2000
    n = cat.halos["count"]
    assert n.min() >= 21 and n.max() <= 210_000
    assert np.all(cat.halos["first"][1:] == np.cumsum(n)[:-1])
    # the n^-1.9 law on [20, 200000] with 5% extra: ~0.5M records
    assert 3e5 < cat.n_records < 1.5e6


def test_config5_totals():
    pytest.importorskip("scipy")
    cat = catalog.config5(n_halos=2000, total=16_777_216 // 25)  # scaled-down draw
    assert cat.n_records == round(16_777_216 // 25 * 0.4)
    assert cat.halos["count"].min() >= 20


def test_particle_rvir_dens_float_rounding():
    prm = catalog.Params(particle_mass=1e9)
    v = prm.particle_rvir_dens
    assert v == float(np.float32(v))
    # 359.3 Om rho_crit / m at a = 1 (src/groupies.c:97-100 with Om = 0.27)
    assert v == pytest.approx(359.3 * 0.27 * catalog.CRITICAL_DENSITY / 1e9, rel=2e-3)


def test_small_batch_potential_only_tail():
    cat = catalog.small_batch(seed=5, n_halos=10, lo=20, hi=200, potential_only=0.3)
    for h in range(cat.n_halos):
        seg = cat.segment(h)
        ke = seg["ke"]
        k = int(np.argmax(ke < 0)) if np.any(ke < 0) else seg.shape[0]
        assert np.all(ke[:k] == 0) and np.all(ke[k:] == -1.0)
