"""Seeded synthetic halo catalogues for the potential + bound pass.

Every catalogue is a list of ``compute_potential`` calls laid end to end: one
array of ``struct potential`` records (src/potential.h:7-21) plus one
``hp_halo`` descriptor per call, exactly as the po assembly
(src/groupies.c:212-270) would hand them to the bound pass
(src/properties.c:537-550).  ``r2`` is filled as ``_reset_potentials`` does
(src/groupies.c:220-225): float differences to the centre, float squares
summed left to right.

Physics conventions (units of the reference: comoving Mpc/h, km/s, Msun/h):
  * NFW radii come from inverting the enclosed-mass profile truncated at
    r_vir with the principal branch of Lambert W; directions are isotropic.
  * Velocities are isotropic Gaussians with the isotropic Jeans dispersion of
    an untruncated NFW profile (outer boundary condition sigma -> 0 as
    r -> infinity), evaluated from a tabulated integral.
  * G is the reference's Gc = 4.30117902e-9 (km/s)^2 Mpc/Msun
    (src/universal_constants.h:4).
  * Virial radii, where a config leaves them free, follow the reference's
    virial density at a = 1: 359.3 * Om * rho_crit (src/groupies.c:97-100).

The RNG is numpy's Philox-4x32-10 counter generator, seeded per config.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

GC = 4.30117902e-9            # src/universal_constants.h:4
CRITICAL_DENSITY = 2.77519737e11  # src/universal_constants.h:5

POTENTIAL_DTYPE = np.dtype(
    {
        "names": ["pos", "r2", "pe", "ke", "flags"],
        "formats": [("<f4", (6,)), "<f4", "<f8", "<f4", "<i4"],
        "offsets": [0, 24, 32, 40, 44],
        "itemsize": 48,
    }
)

# hp_halo_props (include/halopot.h): the halo property pass's results
_PROPS_LAYOUT = [("m", 1), ("mgrav", 1), ("r", 1), ("vmax", 1), ("rvmax", 1), ("vrms", 1),
                 ("Xoff", 1), ("Voff", 1), ("alt_m", 4), ("halfmass_radius", 1), ("J", 3),
                 ("energy", 1), ("kin_to_pot", 1), ("spin", 1), ("bullock_spin", 1),
                 ("rvir", 1), ("rs", 1), ("klypin_rs", 1), ("chi2", 1), ("b_to_a", 1),
                 ("c_to_a", 1), ("A", 3), ("b_to_a2", 1), ("c_to_a2", 1), ("A2", 3),
                 ("inertia_tensor", 6), ("inertia_tensor2", 6), ("m_pe_b", 1), ("m_pe_d", 1)]
PROPS_FLOAT_FIELDS = tuple(f for f, _ in _PROPS_LAYOUT)
PROPS_INT_FIELDS = ("n_mdelta", "dens_tot", "n_vir", "n_half")
PROPS_DTYPE = np.dtype([(f, "<f4") if n == 1 else (f, "<f4", (n,)) for f, n in _PROPS_LAYOUT]
                       + [(f, "<i8") for f in PROPS_INT_FIELDS])
assert PROPS_DTYPE.itemsize == 224

# struct particle (src/particle.h:5-8): a FOF group's particle copies
PARTICLE_DTYPE = np.dtype({"names": ["id", "pos"], "formats": ["<i8", ("<f4", (6,))],
                           "offsets": [0, 8], "itemsize": 32})

# hp_tree_halo (include/halopot.h): one halo of the hierarchy
TREE_DTYPE = np.dtype(
    {
        "names": ["p_start", "num_p", "num_child_particles", "child", "next_cochild", "sub_of",
                  "pos"],
        "formats": ["<i8"] * 6 + [("<f4", (6,))],
        "offsets": [0, 8, 16, 24, 32, 40, 48],
        "itemsize": 72,
    }
)

# mass definitions of the halo finder's defaults (src/config.template.h:9-13)
DEFAULT_MASS_DEFINITIONS = ("vir", "200b", "200c", "500c", "2500c")

HALO_DTYPE = np.dtype(
    {
        "names": [
            "first", "count", "pos", "corevel", "bulkvel", "min_vel_err",
            "min_bulkvel_err", "n_core", "is_sub", "status", "kept", "n_bound",
        ],
        "formats": [
            "<i8", "<i8", ("<f4", (6,)), ("<f4", (3,)), ("<f4", (3,)), "<f4",
            "<f4", "<i8", "<i4", "<i4", "<i8", "<i8",
        ],
        "offsets": [0, 8, 16, 40, 52, 64, 68, 72, 80, 84, 88, 96],
        "itemsize": 104,
    }
)

SEEDS = {"fof": 0x5EED0007, "c1": 0x5EED0001, "c2": 0x5EED0002, "c3": 0x5EED0003, "c4": 0x5EED0001,
         "c5": 0x5EED0005, "tree": 0x5EED0006}


@dataclasses.dataclass
class Params:
    """Configuration globals the path reads (src/config.template.h:7-25)."""

    force_res: float = 0.003
    particle_mass: float = 1e9
    Om: float = 0.27
    Ol: float = 0.73
    W0: float = -1.0
    WA: float = 0.0
    scale_now: float = 1.0

    def hubble_scaling(self, z: float) -> float:
        """src/hubble.c:8-19."""
        z1 = 1.0 + z
        a = 1.0 / z1
        weff = self.W0 if a == 1.0 else self.W0 + self.WA - self.WA * (a - 1.0) / math.log(a)
        return math.sqrt(self.Om * (z1 * z1 * z1) + self.Ol * math.pow(a, -3.0 * (1.0 + weff)))

    def vir_density(self, a: float) -> float:
        """Virial overdensity relative to the mean matter density (src/groupies.c:97-100)."""
        x = (self.Om / a ** 3) / self.hubble_scaling(1.0 / a - 1.0) ** 2 - 1.0
        return (18 * math.pi * math.pi + 82.0 * x - 39 * x * x) / (1.0 + x)

    @property
    def particle_rvir_dens(self) -> float:
        """Particles per (Mpc/h)^3 at the virial threshold, rounded as
        src/groupies.c:106-135 does: float scale factor, float background
        density, float threshold."""
        a = float(np.float32(self.scale_now))
        cons = np.float32(self.Om * CRITICAL_DENSITY / self.particle_mass)
        return float(np.float32(self.vir_density(a) * float(cons)))

    @property
    def rho_vir(self) -> float:
        """Virial mass density, (Msun/h)/(Mpc/h)^3."""
        return self.particle_rvir_dens * self.particle_mass


@dataclasses.dataclass
class Catalogue:
    name: str
    params: Params
    po: np.ndarray        # POTENTIAL_DTYPE, all calls end to end
    halos: np.ndarray     # HALO_DTYPE, one per call

    @property
    def n_records(self) -> int:
        return int(self.po.shape[0])

    @property
    def n_halos(self) -> int:
        return int(self.halos.shape[0])

    def __post_init__(self):
        assert self.po.dtype == POTENTIAL_DTYPE and self.po.dtype.itemsize == 48
        assert self.halos.dtype == HALO_DTYPE and self.halos.dtype.itemsize == 104

    def copy(self) -> "Catalogue":
        return Catalogue(self.name, dataclasses.replace(self.params), self.po.copy(),
                         self.halos.copy())

    def segment(self, h: int) -> np.ndarray:
        f = int(self.halos["first"][h])
        return self.po[f:f + int(self.halos["count"][h])]


# ---------------------------------------------------------------- NFW pieces

def nfw_m(x):
    """Dimensionless enclosed mass ln(1+x) - x/(1+x)."""
    x = np.asarray(x, dtype=np.float64)
    return np.log1p(x) - x / (1.0 + x)


def nfw_invert_mass(p):
    """x with nfw_m(x) = p: with y = 1 + x, -1/y = W0(-exp(-(p + 1)))."""
    from scipy.special import lambertw

    p = np.asarray(p, dtype=np.float64)
    w = lambertw(-np.exp(-(p + 1.0)), 0).real
    return -1.0 / w - 1.0


class _JeansTable:
    """sigma^2(x) / (G M_s / r_s) for an isotropic, untruncated NFW sphere:
    x (1+x)^2 * integral_x^inf m(t) / (t^3 (1+t)^2) dt, tabulated on a log grid."""

    def __init__(self, lo=1e-5, hi=1e5, n=20001):
        lx = np.linspace(math.log(lo), math.log(hi), n)
        x = np.exp(lx)
        f = nfw_m(x) / (x ** 3 * (1.0 + x) ** 2) * x  # integrand in d(ln x)
        # cumulative integral from the outer end (trapezoid in ln x) plus the
        # analytic tail beyond hi, where m ~ ln(x) - 1 and the integrand ~ m / x^5
        seg = 0.5 * (f[1:] + f[:-1]) * np.diff(lx)
        tail = (math.log(hi) - 1.0) / (4.0 * hi ** 4)
        cum = np.concatenate([np.cumsum(seg[::-1])[::-1], [0.0]]) + tail
        self.lx = lx
        self.val = x * (1.0 + x) ** 2 * cum

    def __call__(self, x):
        x = np.clip(np.asarray(x, dtype=np.float64), math.exp(self.lx[0]), math.exp(self.lx[-1]))
        return np.interp(np.log(x), self.lx, self.val)


_JEANS = None


def jeans_sigma2(x):
    global _JEANS
    if _JEANS is None:
        _JEANS = _JeansTable()
    return _JEANS(x)


def _unit_vectors(rng, n):
    u = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    s = np.sqrt(np.maximum(0.0, 1.0 - u * u))
    return np.stack([s * np.cos(phi), s * np.sin(phi), u], axis=1)


def nfw_particles(rng, n_per, conc, rvir, mvir, centre, bulk):
    """Particles of many NFW halos at once.  All per-halo arguments are arrays
    of length H (centre, bulk: H x 3).  Returns pos (N x 3, f64), vel (N x 3)."""
    n_per = np.asarray(n_per, dtype=np.int64)
    idx = np.repeat(np.arange(n_per.shape[0]), n_per)
    c = np.asarray(conc, dtype=np.float64)[idx]
    rv = np.asarray(rvir, dtype=np.float64)[idx]
    mv = np.asarray(mvir, dtype=np.float64)[idx]
    u = rng.uniform(0.0, 1.0, idx.shape[0])
    x = nfw_invert_mass(u * nfw_m(c))
    rs = rv / c
    r = x * rs
    pos = np.asarray(centre, dtype=np.float64)[idx] + r[:, None] * _unit_vectors(rng, idx.shape[0])
    ms = mv / nfw_m(c)
    sig = np.sqrt(GC * ms / rs * jeans_sigma2(x))
    vel = np.asarray(bulk, dtype=np.float64)[idx] + sig[:, None] * rng.standard_normal((idx.shape[0], 3))
    return pos, vel


def virial_radius(params: Params, mass):
    return np.cbrt(3.0 * np.asarray(mass, dtype=np.float64) / (4.0 * math.pi * params.rho_vir))


def sigma_vir(mass, rvir):
    """1-D dispersion of a singular isothermal sphere with the same V_vir."""
    return np.sqrt(GC * np.asarray(mass, dtype=np.float64) / np.asarray(rvir, dtype=np.float64) / 2.0)


def power_law_counts(rng, n, lo, hi, slope=-1.9):
    """dN/dn ~ n^slope on [lo, hi], by inverse CDF, floored to integers."""
    a = 1.0 + slope
    u = rng.uniform(0.0, 1.0, n)
    v = (lo ** a + u * (hi ** a - lo ** a)) ** (1.0 / a)
    return np.clip(np.floor(v).astype(np.int64), lo, hi)


# ------------------------------------------------------------ assembly

def _concat(parts, dtype=None):
    """np.concatenate that keeps the padded record layout (numpy 2 returns a
    packed, canonical structured dtype otherwise)."""
    dtype = dtype or parts[0].dtype
    out = np.empty(sum(p.shape[0] for p in parts), dtype=dtype)
    o = 0
    for p in parts:
        out[o:o + p.shape[0]] = p
        o += p.shape[0]
    return out


def _records(pos, vel, cen, ke=None, flags=None):
    """struct potential records with r2 about cen (src/groupies.c:220-225)."""
    n = pos.shape[0]
    po = np.zeros(n, dtype=POTENTIAL_DTYPE)
    p32 = pos.astype(np.float32)
    po["pos"][:, 0:3] = p32
    po["pos"][:, 3:6] = vel.astype(np.float32)
    c32 = np.asarray(cen, dtype=np.float32).reshape(-1, 3)
    d = (p32 - c32).astype(np.float32)
    r2 = (d[:, 0] * d[:, 0]).astype(np.float32)
    r2 = (r2 + (d[:, 1] * d[:, 1]).astype(np.float32)).astype(np.float32)
    r2 = (r2 + (d[:, 2] * d[:, 2]).astype(np.float32)).astype(np.float32)
    po["r2"] = r2
    if ke is not None:
        po["ke"] = ke
    po["flags"] = 1 if flags is None else flags
    return po


def _halo_rows(firsts, counts, centres, bulks, is_sub):
    h = np.zeros(len(counts), dtype=HALO_DTYPE)
    h["first"] = firsts
    h["count"] = counts
    h["pos"][:, 0:3] = np.asarray(centres, dtype=np.float32)
    h["pos"][:, 3:6] = np.asarray(bulks, dtype=np.float32)
    # prior values before the pass; calculate_corevel overwrites them when the
    # halo reaches the virial threshold (src/properties.c:176-177)
    h["corevel"] = np.asarray(bulks, dtype=np.float32)
    h["bulkvel"] = np.asarray(bulks, dtype=np.float32)
    h["is_sub"] = is_sub
    return h


def config1(seed=SEEDS["c1"], n_halo=10_000, n_bg=1_000) -> Catalogue:
    """One NFW halo (c = 10, r_vir = 1 Mpc/h, m = 1e9) plus uniform background
    in a 2 r_vir cube at 3 sigma_vir."""
    rng = np.random.Generator(np.random.Philox(seed))
    prm = Params(particle_mass=1e9)
    cen = np.array([[37.5, 62.5, 12.5]])
    bulk = np.array([[120.0, -80.0, 35.0]])
    mvir = n_halo * prm.particle_mass
    pos, vel = nfw_particles(rng, [n_halo], [10.0], [1.0], [mvir], cen, bulk)
    sv = float(sigma_vir(mvir, 1.0))
    bpos = cen + rng.uniform(-1.0, 1.0, (n_bg, 3))
    bvel = bulk + 3.0 * sv * rng.standard_normal((n_bg, 3))
    allp = np.concatenate([pos, bpos])
    allv = np.concatenate([vel, bvel])
    po = _records(allp, allv, cen[0])
    halos = _halo_rows([0], [po.shape[0]], cen, bulk, [0])
    return Catalogue("c1", prm, po, halos)


def config2(seed=SEEDS["c2"], n_halos=2000, lo=20, hi=200_000, contam=0.05,
            box=250.0) -> Catalogue:
    """2,000 NFW halos, n ~ n^-1.9 on [20, 200000], c log-normal (median 8,
    sigma_ln 0.3), 5% unbound contaminants per halo.  The stated law gives an
    expected total near 0.54M records; the realised total is what it is."""
    rng = np.random.Generator(np.random.Philox(seed))
    prm = Params(particle_mass=1e9)
    n = power_law_counts(rng, n_halos, lo, hi)
    conc = np.exp(rng.normal(math.log(8.0), 0.3, n_halos))
    mvir = n * prm.particle_mass
    rvir = virial_radius(prm, mvir)
    cen = rng.uniform(0.0, box, (n_halos, 3))
    bulk = rng.normal(0.0, 200.0, (n_halos, 3))
    pos, vel = nfw_particles(rng, n, conc, rvir, mvir, cen, bulk)
    nc = np.maximum(1, np.round(contam * n).astype(np.int64))
    cidx = np.repeat(np.arange(n_halos), nc)
    sv = sigma_vir(mvir, rvir)
    cpos = cen[cidx] + rvir[cidx, None] * rng.uniform(-1.0, 1.0, (cidx.shape[0], 3))
    cvel = bulk[cidx] + 3.0 * sv[cidx, None] * rng.standard_normal((cidx.shape[0], 3))
    # interleave: each call is [halo particles, contaminants]
    hoff = np.concatenate([[0], np.cumsum(n)])
    coff = np.concatenate([[0], np.cumsum(nc)])
    counts = n + nc
    firsts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    total = int(counts.sum())
    P = np.empty((total, 3))
    V = np.empty((total, 3))
    C = np.empty((total, 3))
    for h in range(n_halos):
        f = firsts[h]
        P[f:f + n[h]] = pos[hoff[h]:hoff[h + 1]]
        V[f:f + n[h]] = vel[hoff[h]:hoff[h + 1]]
        P[f + n[h]:f + counts[h]] = cpos[coff[h]:coff[h + 1]]
        V[f + n[h]:f + counts[h]] = cvel[coff[h]:coff[h + 1]]
        C[f:f + counts[h]] = cen[h]
    po = _records(P, V, C)
    halos = _halo_rows(firsts, counts, cen, bulk, np.zeros(n_halos, dtype=np.int32))
    return Catalogue("c2", prm, po, halos)


def config3(seed=SEEDS["c3"], n_host=2_000_000, n_subs=30, c_host=6.0, rvir_host=2.0,
            c_sub=10.0) -> Catalogue:
    """A 2M-particle NFW host (c = 6, r_vir = 2 Mpc/h) with 30 embedded NFW
    sub-halos (mass fractions log-uniform on [1e-4, 1e-2], positions from the
    host profile, bulk velocities from the host dispersion).  31 calls: the
    host call gathers the sub-halo particles too (src/groupies.c:235-259); each
    sub-halo call holds its own particles (the host is not added as
    potential-only: 0.3 * N_host exceeds every sub-halo, config.template.h:116)."""
    rng = np.random.Generator(np.random.Philox(seed))
    prm = Params()
    rho = Params(particle_mass=1.0).rho_vir  # mass density does not depend on m
    mvir_host = 4.0 * math.pi / 3.0 * rho * rvir_host ** 3
    prm.particle_mass = mvir_host / n_host
    cen = np.array([[125.0, 80.0, 170.0]])
    bulk = np.array([[-60.0, 150.0, 20.0]])
    hpos, hvel = nfw_particles(rng, [n_host], [c_host], [rvir_host], [mvir_host], cen, bulk)
    frac = np.exp(rng.uniform(math.log(1e-4), math.log(1e-2), n_subs))
    nsub = np.maximum(20, np.round(frac * n_host).astype(np.int64))
    msub = nsub * prm.particle_mass
    rsub = virial_radius(prm, msub)
    xs = nfw_invert_mass(rng.uniform(0.0, 1.0, n_subs) * nfw_m(c_host))
    scen = cen + (xs * rvir_host / c_host)[:, None] * _unit_vectors(rng, n_subs)
    ms_host = mvir_host / nfw_m(c_host)
    sig = np.sqrt(GC * ms_host / (rvir_host / c_host) * jeans_sigma2(xs))
    sbulk = bulk + sig[:, None] * rng.standard_normal((n_subs, 3))
    spos, svel = nfw_particles(rng, nsub, np.full(n_subs, c_sub), rsub, msub, scen, sbulk)
    soff = np.concatenate([[0], np.cumsum(nsub)])
    host_total = n_host + int(nsub.sum())
    flags_host = np.ones(host_total, dtype=np.int32)
    small = nsub < 0.03 * n_host  # src/groupies.c:229-230
    for s in range(n_subs):
        if small[s]:
            flags_host[n_host + soff[s]:n_host + soff[s + 1]] = 2
    parts = [_records(np.concatenate([hpos, spos]), np.concatenate([hvel, svel]), cen[0],
                      flags=flags_host)]
    firsts = [0]
    counts = [host_total]
    for s in range(n_subs):
        firsts.append(firsts[-1] + counts[-1])
        counts.append(int(nsub[s]))
        parts.append(_records(spos[soff[s]:soff[s + 1]], svel[soff[s]:soff[s + 1]], scen[s]))
    po = _concat(parts)
    halos = _halo_rows(firsts, counts, np.concatenate([cen, scen]), np.concatenate([bulk, sbulk]),
                       np.array([0] + [1] * n_subs, dtype=np.int32))
    return Catalogue("c3", prm, po, halos)


@dataclasses.dataclass
class Hierarchy:
    """FOF groups as the halo finder holds them before the bound pass: the
    particle copies, each halo's own particles contiguous, and the halo tree
    (struct halo / extra_halo_info links, src/halo.h:8-41)."""

    params: Params
    particles: np.ndarray   # PARTICLE_DTYPE
    tree: np.ndarray        # TREE_DTYPE
    base: np.ndarray        # the halos that get a call: num_p >= 1 (src/properties.c:524)


def hierarchy(seed=SEEDS["tree"], n_groups=200, lo=200, hi=50_000, mean_subs=1.5, depth=3,
              big_sub=0.15, empty_sub=0.03) -> Hierarchy:
    """Seeded FOF groups: an NFW host per group (n ~ n^-1.9 on [lo, hi]) with
    sub-halos to the given depth (Poisson numbers of children, each with a
    log-uniform 0.5%-50% of its parent's particles, or 45%-90% with
    probability big_sub, so that the host's subtree joins its call as
    potential-only records; some empty).  num_child_particles is each halo's
    own count plus its descendants' (src/properties.c:139-153)."""
    rng = np.random.Generator(np.random.Philox(seed))
    prm = Params()
    n_host = power_law_counts(rng, n_groups, lo, hi)
    rows = []          # [num_p, sub_of, centre(3), bulk(3)]
    children = []
    def add(num_p, parent, cen, bulk):
        rows.append([num_p, parent, cen, bulk])
        children.append([])
        if parent >= 0:
            children[parent].append(len(rows) - 1)
        return len(rows) - 1
    def grow(h, level):
        if level >= depth:
            return
        num_p, _, cen, bulk = rows[h]
        for _ in range(min(6, rng.poisson(mean_subs / (level + 1)))):
            if rng.uniform() < empty_sub:
                n = 0
            elif rng.uniform() < big_sub:
                n = int(num_p * rng.uniform(0.45, 0.9))
            else:
                n = int(num_p * math.exp(rng.uniform(math.log(0.005), math.log(0.5))))
            rv = float(virial_radius(prm, max(n, 1) * prm.particle_mass))
            rparent = float(virial_radius(prm, max(num_p, 1) * prm.particle_mass))
            off = rparent * rng.uniform(0.05, 0.8) * _unit_vectors(rng, 1)[0]
            sb = bulk + float(sigma_vir(max(num_p, 1) * prm.particle_mass, rparent)) * \
                rng.standard_normal(3)
            c = add(n, h, cen + off, sb)
            if n > 20:
                grow(c, level + 1)
    for g in range(n_groups):
        cen = rng.uniform(0.0, 100.0, 3)
        bulk = rng.normal(0.0, 200.0, 3)
        grow(add(int(n_host[g]), -1, cen, bulk), 0)
    nh = len(rows)
    num_p = np.array([r[0] for r in rows], np.int64)
    cen = np.array([r[2] for r in rows])
    bulk = np.array([r[3] for r in rows])
    mvir = np.maximum(num_p, 1) * prm.particle_mass
    pos, vel = nfw_particles(rng, num_p, np.exp(rng.normal(math.log(8.0), 0.3, nh)),
                             virial_radius(prm, mvir), mvir, cen, bulk)
    parts = np.zeros(int(num_p.sum()), PARTICLE_DTYPE)
    parts["id"] = np.arange(parts.shape[0])
    parts["pos"][:, :3] = pos
    parts["pos"][:, 3:] = vel
    tree = np.zeros(nh, TREE_DTYPE)
    tree["p_start"] = np.concatenate([[0], np.cumsum(num_p)[:-1]])
    tree["num_p"] = num_p
    tree["sub_of"] = [r[1] for r in rows]
    tree["child"] = -1
    tree["next_cochild"] = -1
    for h in range(nh):
        kids = children[h]
        if kids:
            tree["child"][h] = kids[0]
            for a, b in zip(kids[:-1], kids[1:]):
                tree["next_cochild"][a] = b
    ncp = num_p.copy()
    for h in range(nh - 1, -1, -1):  # children come after their parent
        for c in children[h]:
            ncp[h] += ncp[c]
    tree["num_child_particles"] = ncp
    tree["pos"][:, :3] = cen
    tree["pos"][:, 3:] = bulk
    return Hierarchy(prm, parts, tree, np.nonzero(num_p >= 1)[0].astype(np.int64))


@dataclasses.dataclass
class FofGroups:
    """FOF groups as the sub-FOF step receives them: each group's particles
    contiguous (struct particle, 32 B), in an arbitrary order."""

    particles: np.ndarray
    groups: np.ndarray        # FOF_GROUP_DTYPE rows: first, num_p
    params: Params


FOF_GROUP_DTYPE = np.dtype([("first", "<i8"), ("num_p", "<i8")])


def fof_groups(seed=SEEDS["fof"], n_groups=200, lo=200, hi=50_000, **kw) -> FofGroups:
    """The groups of hierarchy() as whole FOF groups (a host with its
    sub-halos, distinct in phase space), each group's particles shuffled
    (a FOF pass leaves them in no particular order)."""
    h = hierarchy(seed=seed, n_groups=n_groups, lo=lo, hi=hi, **kw)
    hosts = np.flatnonzero(h.tree["sub_of"] < 0)
    g = np.zeros(hosts.shape[0], FOF_GROUP_DTYPE)
    g["first"] = h.tree["p_start"][hosts]
    g["num_p"] = h.tree["num_child_particles"][hosts]
    rng = np.random.Generator(np.random.Philox(seed + 1))
    parts = h.particles.copy()
    for f, n in zip(g["first"], g["num_p"]):
        parts[f:f + n] = parts[f:f + n][rng.permutation(int(n))]
    return FofGroups(parts, g, h.params)


def config5(seed=SEEDS["c5"], total=16_777_216, box=250.0, n_halos=50_000, frac_halo=0.40,
            lo=20, hi=500_000) -> Catalogue:
    """The one-GPU full catalogue: 16,777,216 particles in a 250 Mpc/h box,
    m = rho_crit Om L^3 / N (src/config.c:21-23).  40% (6,710,886) sit in
    50,000 NFW halos; counts are drawn from n^-1.9 on [20, 500000] and then
    scaled to that total (clamped at 20, the remainder taken from the
    largest), because the unscaled law would need ~15.8M particles.  c is
    log-normal about 9 (M / 1e13)^-0.13 with sigma_ln 0.3.  The 60% diffuse
    particles never enter a compute_potential call, so only the halo calls
    are materialised here."""
    rng = np.random.Generator(np.random.Philox(seed))
    prm = Params(particle_mass=CRITICAL_DENSITY * box ** 3 * 0.27 / total)
    n_in = int(round(total * frac_halo))
    raw = power_law_counts(rng, n_halos, lo, hi).astype(np.float64)
    n = _scale_counts(raw, n_in, lo)
    mvir = n * prm.particle_mass
    conc = 9.0 * (mvir / 1e13) ** -0.13 * np.exp(rng.normal(0.0, 0.3, n_halos))
    rvir = virial_radius(prm, mvir)
    cen = rng.uniform(0.0, box, (n_halos, 3))
    bulk = rng.normal(0.0, 300.0 / math.sqrt(3.0), (n_halos, 3))
    pos, vel = nfw_particles(rng, n, conc, rvir, mvir, cen, bulk)
    firsts = np.concatenate([[0], np.cumsum(n)[:-1]])
    C = np.repeat(cen, n, axis=0)
    po = _records(pos, vel, C)
    halos = _halo_rows(firsts, n, cen, bulk, np.zeros(n_halos, dtype=np.int32))
    return Catalogue("c5", prm, po, halos)


def _scale_counts(raw, total, lo):
    """Integer counts max(lo, floor(s raw)) summing to `total`: s by bisection
    so that the clamped sum stays at or below the total, then the remainder
    one particle each to the largest halos (ties by index)."""
    if lo * raw.shape[0] > total:
        raise ValueError("total too small for the clamp")

    def clamped(s):
        return np.maximum(lo, np.floor(raw * s)).astype(np.int64)

    a, b = 0.0, 1.0
    while clamped(b).sum() <= total:
        b *= 2.0
    for _ in range(200):
        m = 0.5 * (a + b)
        if clamped(m).sum() <= total:
            a = m
        else:
            b = m
    n = clamped(a)
    rem = total - int(n.sum())
    order = np.argsort(-n, kind="stable")
    n[order[:rem]] += 1
    assert int(n.sum()) == total and n.min() >= lo
    return n


def make(name: str, **kw) -> Catalogue:
    return {"c1": config1, "c2": config2, "c3": config3, "c4": config1, "c5": config5}[name](**kw)


def small_batch(seed=1234, n_halos=40, lo=2, hi=3000, contam=0.05, potential_only=0.0,
                sub_fraction=0.5, particle_mass=1e9) -> Catalogue:
    """A quick mixed batch for tests: tiny to mid-size halos, some sub-halos,
    optional potential-only tails (ke = -1, src/groupies.c:226-227)."""
    rng = np.random.Generator(np.random.Philox(seed))
    prm = Params(particle_mass=particle_mass)
    n = power_law_counts(rng, n_halos, lo, hi)
    conc = np.exp(rng.normal(math.log(8.0), 0.3, n_halos))
    mvir = n * prm.particle_mass
    rvir = virial_radius(prm, mvir)
    cen = rng.uniform(0.0, 100.0, (n_halos, 3))
    bulk = rng.normal(0.0, 200.0, (n_halos, 3))
    parts, firsts, counts = [], [], []
    f = 0
    for h in range(n_halos):
        p, v = nfw_particles(rng, [n[h]], [conc[h]], [rvir[h]], [mvir[h]], cen[h:h + 1], bulk[h:h + 1])
        nc = int(round(contam * n[h]))
        if nc:
            sv = float(sigma_vir(mvir[h], rvir[h]))
            p = np.concatenate([p, cen[h] + rvir[h] * rng.uniform(-1, 1, (nc, 3))])
            v = np.concatenate([v, bulk[h] + 3 * sv * rng.standard_normal((nc, 3))])
        rec = _records(p, v, cen[h])
        npo = int(round(potential_only * rec.shape[0]))
        if npo:
            q, w = nfw_particles(rng, [npo], [conc[h]], [2 * rvir[h]], [mvir[h]], cen[h:h + 1], bulk[h:h + 1])
            rec = _concat([rec, _records(q, w, cen[h], ke=-1.0)])
        parts.append(rec)
        firsts.append(f)
        counts.append(rec.shape[0])
        f += rec.shape[0]
    po = _concat(parts)
    is_sub = (rng.uniform(0, 1, n_halos) < sub_fraction).astype(np.int32)
    halos = _halo_rows(firsts, counts, cen, bulk, is_sub)
    return Catalogue("small", prm, po, halos)
