"""ctypes access to the CPU checkers.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
bench.py's ``cpu_baseline`` leg, never by the product.

* ``Restatement`` wraps liboracle.so (restate.c): the clean-room restatement of
  src/potential.c, the POTENTIAL fast3tree build, compute_kinetic_energy,
  calculate_corevel (src/properties.c:165-214) and the bound-pass driver
  (src/properties.c:537-550), plus the FP64 direct sum.
* ``Reference`` wraps _ref/libref_potential.so, the reference's own
  src/potential.c + src/hubble.c + src/config_vars.c + src/error.cpp compiled
  from its sources by oracle/Makefile.  It reads the configuration globals
  FORCE_RES, PARTICLE_MASS, Om, Ol, W0, WA (src/config_vars.h) at call time.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent

NODE_DTYPE = np.dtype(
    {
        "names": ["lo", "hi", "first", "n", "left", "dim", "depth", "com", "m", "dmin"],
        "formats": [("<f4", (3,)), ("<f4", (3,)), "<i8", "<i8", "<i8", "<i4", "<i4",
                    ("<f4", (3,)), "<f4", "<f4"],
        "offsets": [0, 12, 24, 32, 40, 48, 52, 56, 68, 72],
        "itemsize": 80,
    }
)


class _Params(C.Structure):
    _fields_ = [("force_res", C.c_double), ("particle_mass", C.c_double),
                ("Om", C.c_double), ("Ol", C.c_double), ("W0", C.c_double),
                ("WA", C.c_double), ("scale_now", C.c_double),
                ("particle_rvir_dens", C.c_double)]


class _Stats(C.Structure):
    _fields_ = [("p2p", C.c_int64), ("p2m", C.c_int64), ("visits", C.c_int64),
                ("nodes", C.c_int64), ("leaves", C.c_int64), ("max_depth", C.c_int64),
                ("nfinite", C.c_int64)]


def build(quiet: bool = True) -> None:
    """Compile liboracle.so and, where the reference sources exist, _ref/."""
    out = subprocess.run(["make", "-C", str(HERE)], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def _record_dtype():
    """struct potential, 48 bytes (src/potential.h:7-21)."""
    return np.dtype({"names": ["pos", "r2", "pe", "ke", "flags"],
                     "formats": [("<f4", (6,)), "<f4", "<f8", "<f4", "<i4"],
                     "offsets": [0, 24, 32, 40, 44], "itemsize": 48})


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _check_records(po):
    if po.dtype.itemsize != 48 or po.dtype.names != ("pos", "r2", "pe", "ke", "flags") \
            or not po.flags.c_contiguous:
        raise TypeError("records must be contiguous 48-byte struct potential")
    return po


def _params(prm) -> _Params:
    return _Params(prm.force_res, prm.particle_mass, prm.Om, prm.Ol, prm.W0, prm.WA,
                   prm.scale_now, prm.particle_rvir_dens)


class Restatement:
    def __init__(self, path: str | os.PathLike | None = None):
        path = path or HERE / "liboracle.so"
        if not pathlib.Path(path).exists():
            build()
        self.lib = L = C.CDLL(str(path))
        L.orc_compute_potential.restype = C.c_int64
        L.orc_compute_potential.argtypes = [C.c_void_p, C.c_int64, C.POINTER(_Params), C.c_void_p,
                                            C.c_void_p, C.POINTER(_Stats), C.c_void_p, C.c_int64]
        L.orc_build_tree.restype = C.c_int64
        L.orc_build_tree.argtypes = [C.c_void_p, C.c_int64, C.POINTER(_Params), C.c_void_p,
                                     C.c_int64]
        L.orc_compute_potential_bh64.argtypes = [C.c_void_p, C.c_int64, C.POINTER(_Params)]
        L.orc_direct64.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_double, C.c_double,
                                   C.c_void_p]
        L.orc_kinetic.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_float,
                                  C.POINTER(_Params)]
        L.orc_corevel.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(_Params)]
        L.orc_bound_pass.restype = C.c_int64
        L.orc_bound_pass.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(_Params),
                                     C.c_void_p]
        L.orc_bound_pass_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_int64, C.POINTER(_Params), C.c_void_p,
                                           C.c_void_p, C.c_int]
        L.orc_bound_pass_batch_fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_int64, C.POINTER(_Params), C.c_void_p, C.c_int,
                                              C.c_void_p, C.c_void_p]
        L.orc_potential_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                          C.POINTER(_Params), C.c_void_p, C.c_void_p, C.c_int]
        L.orc_set_never_accept.argtypes = [C.c_int]
        L.orc_set_shape_reverse.argtypes = [C.c_int]
        L.orc_set_nmass_out.argtypes = [C.c_void_p]
        L.orc_hubble_scaling.restype = C.c_double
        L.orc_hubble_scaling.argtypes = [C.POINTER(_Params), C.c_double]
        L.orc_max_threads.restype = C.c_int
        L.orc_halo_props.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(_Params),
                                     C.c_void_p, C.c_void_p]
        L.orc_halo_props_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(_Params),
                                           C.c_void_p, C.c_void_p, C.c_int]
        L.orc_particle_radii.restype = C.c_int64
        L.orc_particle_radii.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_double,
                                         C.c_void_p, C.c_int64]
        L.orc_subfof.restype = C.c_int
        L.orc_subfof.argtypes = [C.c_void_p, C.c_int64, C.c_float, C.c_int, C.c_double, C.c_int64,
                                 C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_int64, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
        L.orc_subfof_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_float, C.c_int,
                                       C.c_double, C.c_int64, C.c_void_p, C.c_void_p, C.c_int]
        L.orc_next_closest.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p]
        L.orc_find_median_r.restype = C.c_float
        L.orc_find_median_r.argtypes = [C.c_void_p, C.c_int64, C.c_float, C.POINTER(C.c_uint)]
        for f in ("orc_sizeof_node", "orc_sizeof_halo", "orc_sizeof_potential",
                  "orc_sizeof_props", "orc_sizeof_tree_halo"):
            getattr(L, f).restype = C.c_int64
        assert L.orc_sizeof_node() == NODE_DTYPE.itemsize
        assert L.orc_sizeof_potential() == 48

    def compute_potential(self, po: np.ndarray, prm, counts: bool = False, tree: bool = False):
        """Barnes-Hut potential; po (POTENTIAL_DTYPE) is permuted in place.
        Returns a dict with stats and, on request, per-record counters and the tree."""
        _check_records(po)
        n = po.shape[0]
        pp = _params(prm)
        st = _Stats()
        np2p = np.zeros(n, np.int32) if counts else None
        np2m = np.zeros(n, np.int32) if counts else None
        cap = 2 * n + 1 if tree else 0
        nodes = np.zeros(cap, NODE_DTYPE) if tree else None
        nn = self.lib.orc_compute_potential(_p(po), n, C.byref(pp),
                                            _p(np2p) if counts else None,
                                            _p(np2m) if counts else None, C.byref(st),
                                            _p(nodes) if tree else None, cap)
        out = {f: getattr(st, f) for f, _ in _Stats._fields_}
        if counts:
            out["np2p"], out["np2m"] = np2p, np2m
        if tree:
            out["tree"] = nodes[:nn]
        return out

    def potential_batch(self, po: np.ndarray, halos: np.ndarray, prm, nthreads=0):
        """compute_potential over every call (OpenMP over calls); po is permuted
        within each call.  Returns per-call (p2p, p2m) counts of the walk."""
        _check_records(po)
        first = np.ascontiguousarray(halos["first"], np.int64)
        count = np.ascontiguousarray(halos["count"], np.int64)
        p2p = np.zeros(halos.shape[0], np.int64)
        p2m = np.zeros(halos.shape[0], np.int64)
        self.lib.orc_potential_batch(_p(po), _p(first), _p(count), halos.shape[0],
                                     C.byref(_params(prm)), _p(p2p), _p(p2m), nthreads)
        return p2p, p2m

    def compute_potential_mass(self, po: np.ndarray, prm):
        """compute_potential that also returns, per record in tree order, the
        P2P partner count and the number of points behind its monopoles."""
        nmass = np.zeros(po.shape[0], np.int64)
        self.lib.orc_set_nmass_out(_p(nmass))
        try:
            out = self.compute_potential(po, prm, counts=True, tree=True)
        finally:
            self.lib.orc_set_nmass_out(None)
        out["nmass"] = nmass
        return out

    def never_accept(self, on: bool) -> None:
        self.lib.orc_set_never_accept(1 if on else 0)

    def build_tree(self, po: np.ndarray, prm) -> np.ndarray:
        _check_records(po)
        n = po.shape[0]
        nodes = np.zeros(2 * n + 1, NODE_DTYPE)
        nn = self.lib.orc_build_tree(_p(po), n, C.byref(_params(prm)), _p(nodes), nodes.shape[0])
        return nodes[:nn]

    def compute_potential_bh64(self, po: np.ndarray, prm) -> None:
        _check_records(po)
        self.lib.orc_compute_potential_bh64(_p(po), po.shape[0], C.byref(_params(prm)))

    def direct64(self, po: np.ndarray, prm) -> np.ndarray:
        pos = np.ascontiguousarray(po["pos"], dtype=np.float32)
        pe = np.zeros(po.shape[0], np.float64)
        self.lib.orc_direct64(_p(pos), 6, po.shape[0], prm.force_res, prm.particle_mass, _p(pe))
        return pe

    def kinetic(self, po, vel_cen, pos_cen, scale_now, prm):
        vc = np.ascontiguousarray(vel_cen, np.float32)
        pc = np.ascontiguousarray(pos_cen, np.float32)
        self.lib.orc_kinetic(_p(po), po.shape[0], _p(vc), _p(pc), scale_now, C.byref(_params(prm)))

    def corevel(self, halo_row: np.ndarray, po: np.ndarray, prm):
        self.lib.orc_corevel(_p(halo_row), _p(po), po.shape[0], C.byref(_params(prm)))

    def bound_pass(self, po, halo_row, prm):
        _check_records(po)
        bound = np.zeros(po.shape[0], np.uint8)
        kept = self.lib.orc_bound_pass(_p(po), po.shape[0], _p(halo_row), C.byref(_params(prm)),
                                       _p(bound))
        return kept, bound[:kept]

    def bound_pass_batch(self, po, halos, prm, nthreads=0):
        """In place over a whole catalogue: OpenMP dynamic schedule over halos."""
        _check_records(po)
        first = np.ascontiguousarray(halos["first"])
        count = np.ascontiguousarray(halos["count"])
        bound = np.zeros(po.shape[0], np.uint8)
        kept = np.zeros(halos.shape[0], np.int64)
        self.lib.orc_bound_pass_batch(_p(po), _p(first), _p(count), _p(halos), halos.shape[0],
                                      C.byref(_params(prm)), _p(bound), _p(kept), nthreads)
        return kept, bound

    def bound_pass_batch_with(self, po, halos, prm, reference, nthreads=0):
        """The driver of bound_pass_batch with the compiled reference's own
        compute_potential and compute_kinetic_energy (the reference's CPU
        path); reference is a Reference whose globals are set from prm."""
        _check_records(po)
        reference.set_globals(prm)
        first = np.ascontiguousarray(halos["first"])
        count = np.ascontiguousarray(halos["count"])
        kept = np.zeros(halos.shape[0], np.int64)
        cp = C.cast(reference.lib.compute_potential, C.c_void_p)
        ke = C.cast(reference.lib.compute_kinetic_energy, C.c_void_p)
        self.lib.orc_bound_pass_batch_fn(_p(po), _p(first), _p(count), _p(halos), halos.shape[0],
                                         C.byref(_params(prm)), _p(kept), nthreads, cp, ke)
        return kept

    def hubble_scaling(self, prm, z):
        return self.lib.orc_hubble_scaling(C.byref(_params(prm)), z)

    def subfof(self, parts: np.ndarray, thresh: float = 0.7, top: bool = True,
               scaling: float = 1.0, exact_ll: int = 0, min_p: int = 10) -> dict:
        """The sub-FOF step of one group (orc_subfof): the particles
        normalised and in tree order, the sample's next-closest distances,
        the linking length, each particle's component label (smallest tree
        position of its component, -1 without a neighbour), and the
        reference's own output: the particles sorted by its union-find roots
        ("final") and its sub-FOFs of at least min_p particles in root order
        ("fofs", rows of first and num_p)."""
        assert parts.dtype.itemsize == 32
        pp = np.ascontiguousarray(parts).copy()
        n = pp.shape[0]
        sample = np.zeros(max(n, 10000), np.float32)
        label = np.zeros(max(1, n), np.int64)
        final = np.zeros(max(1, n), pp.dtype)
        cap = n // 2 + 2
        ff, fn = np.zeros(cap, np.int64), np.zeros(cap, np.int64)
        nf = C.c_int64()
        tr = C.c_float()
        ns = C.c_int64()
        rc = self.lib.orc_subfof(_p(pp), n, thresh, int(top), scaling, exact_ll, C.byref(tr),
                                 _p(sample), C.byref(ns), _p(label), _p(final), min_p, _p(ff),
                                 _p(fn), cap, C.byref(nf))
        if rc:
            raise ValueError("non-finite coordinate")
        k = min(nf.value, cap)
        return {"tree_order": pp, "target_r": np.float32(tr.value),
                "sample_r": sample[:ns.value], "label": label[:n], "final": final[:n],
                "fofs": np.stack([ff[:k], fn[:k]], 1)}

    def subfof_batch(self, parts: np.ndarray, groups: np.ndarray, thresh: float = 0.7,
                     top: bool = True, scaling: float = 1.0, exact_ll: int = 0,
                     nthreads: int = 0):
        """orc_subfof over every group (FOF_GROUP_DTYPE rows), in place on
        parts; returns (target_r per group, label per particle)."""
        assert parts.dtype.itemsize == 32 and parts.flags.c_contiguous
        g = np.ascontiguousarray(groups)
        tr = np.zeros(max(1, g.shape[0]), np.float32)
        label = np.zeros(max(1, parts.shape[0]), np.int64)
        self.lib.orc_subfof_batch(_p(parts), _p(g), g.shape[0], thresh, int(top), scaling,
                                  exact_ll, _p(tr), _p(label), nthreads)
        return tr[:g.shape[0]], label[:parts.shape[0]]

    def next_closest(self, pts: np.ndarray, queries: np.ndarray) -> np.ndarray:
        pp = np.ascontiguousarray(pts).copy()
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 6)
        out = np.zeros(q.shape[0], np.float32)
        self.lib.orc_next_closest(_p(pp), pp.shape[0], _p(q), q.shape[0], _p(out))
        return out

    def find_median_r(self, rad, frac: float, seed: int) -> np.float32:
        r = np.ascontiguousarray(rad, np.float32).copy()
        sd = C.c_uint(seed)
        return np.float32(self.lib.orc_find_median_r(_p(r), r.shape[0], frac, C.byref(sd)))

    def particle_radii(self, particles: np.ndarray, tree: np.ndarray, base: int, ratio: float,
                       cap: int) -> np.ndarray:
        """calc_particle_radii (src/groupies.c:212-268) for halo `base`: its
        records (48-byte layout).  particles: 32-byte struct particle copies;
        tree: rows of the product's hp_tree_halo layout (72 bytes)."""
        assert particles.dtype.itemsize == 32 and particles.flags.c_contiguous
        assert tree.dtype.itemsize == self.lib.orc_sizeof_tree_halo() and tree.flags.c_contiguous
        po = np.zeros(max(1, cap), _record_dtype())
        n = self.lib.orc_particle_radii(_p(particles), _p(tree), tree.shape[0], base, ratio,
                                        _p(po), cap)
        if n < 0:
            raise ValueError("records do not fit, or a sibling list returns to its first child")
        return po[:n]

    def halo_props(self, po: np.ndarray, kept: int, pos, prm, thresh, out: np.ndarray) -> None:
        """Property pass (src/properties.c:384-510, bound = 0 then 1) of one
        call's kept records, sorted by r2 with pe and ke; out: one record of
        the product's PROPS_DTYPE layout (120 bytes)."""
        _check_records(po)
        assert out.dtype.itemsize == self.lib.orc_sizeof_props() and out.flags.c_contiguous
        ps = np.ascontiguousarray(pos, np.float32)
        th = np.ascontiguousarray(thresh, np.float64)
        self.lib.orc_halo_props(_p(po), kept, _p(ps), C.byref(_params(prm)), _p(th), _p(out))

    def halo_props_batch(self, po: np.ndarray, halos: np.ndarray, prm, thresh, out: np.ndarray,
                         nthreads: int = 0) -> None:
        """halo_props for every call of a batch after its bound pass: halos
        carry first, kept and the centre (pos[3..5] = corevel); OpenMP
        dynamic over calls."""
        _check_records(po)
        assert out.dtype.itemsize == self.lib.orc_sizeof_props() and out.flags.c_contiguous
        assert halos.dtype.itemsize == self.lib.orc_sizeof_halo() and halos.flags.c_contiguous
        th = np.ascontiguousarray(thresh, np.float64)
        self.lib.orc_halo_props_batch(_p(po), _p(halos), halos.shape[0], C.byref(_params(prm)),
                                      _p(th), _p(out), nthreads)

    def max_threads(self):
        return self.lib.orc_max_threads()


class Reference:
    """The compiled reference (src/potential.c et al.)."""

    def __init__(self, path: str | os.PathLike | None = None):
        path = pathlib.Path(path or HERE / "_ref" / "libref_potential.so")
        if not path.exists():
            raise FileNotFoundError(f"{path} missing: run make -C oracle where the reference exists")
        self.lib = L = C.CDLL(str(path))
        L.compute_potential.argtypes = [C.c_void_p, C.c_int64]
        L.compute_kinetic_energy.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                             C.c_float]

    @staticmethod
    def available(path=None) -> bool:
        return pathlib.Path(path or HERE / "_ref" / "libref_potential.so").exists()

    def set_globals(self, prm) -> None:
        for name, val in (("FORCE_RES", prm.force_res), ("PARTICLE_MASS", prm.particle_mass),
                          ("Om", prm.Om), ("Ol", prm.Ol), ("W0", prm.W0), ("WA", prm.WA),
                          ("SCALE_NOW", prm.scale_now)):
            C.c_double.in_dll(self.lib, name).value = val

    def compute_potential(self, po: np.ndarray, prm) -> None:
        _check_records(po)
        self.set_globals(prm)
        self.lib.compute_potential(_p(po), po.shape[0])

    def compute_kinetic_energy(self, po, vel_cen, pos_cen, scale_now, prm) -> None:
        self.set_globals(prm)
        vc = np.ascontiguousarray(vel_cen, np.float32)
        pc = np.ascontiguousarray(pos_cen, np.float32)
        self.lib.compute_kinetic_energy(_p(po), po.shape[0], _p(vc), _p(pc), scale_now)


class ReferenceProps:
    """The reference's own halo property pass and bound-pass steps:
    src/properties.c (inside src/groupies.c) with src/nfw.c, src/jacobi.c,
    src/hubble.c, src/config_vars.c, src/potential.c and src/error.cpp,
    compiled by oracle/Makefile and called through the harness
    oracle/ref_props.c."""

    FIELDS = ("m", "mgrav", "r", "vmax", "rvmax", "vrms", "Xoff", "Voff", "alt_m",
              "halfmass_radius", "J", "energy", "kin_to_pot", "spin", "bullock_spin",
              "rs", "klypin_rs", "b_to_a", "c_to_a", "A", "b_to_a2", "c_to_a2", "A2",
              "m_pe_b", "m_pe_d", "chi2", "inertia_tensor", "inertia_tensor2")
    DTYPE = np.dtype([(f, "<f4", (4,)) if f == "alt_m" else
                      (f, "<f4", (6,)) if f.startswith("inertia") else
                      (f, "<f4", (3,)) if f in ("J", "A", "A2") else (f, "<f4") for f in FIELDS])

    def __init__(self, path: str | os.PathLike | None = None):
        path = pathlib.Path(path or HERE / "_ref" / "libref_props.so")
        if not path.exists():
            raise FileNotFoundError(f"{path} missing: run make -C oracle where the reference exists")
        self.lib = L = C.CDLL(str(path))
        L.ref_halo_props.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ref_mass_definition.argtypes = [C.c_void_p]
        L.ref_halo_props_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_int64, C.c_void_p, C.c_void_p, C.c_int]
        L.ref_bound_pass.restype = C.c_int64
        L.ref_bound_pass.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
        L.ref_particle_radii.restype = C.c_int64
        L.ref_particle_radii.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_double,
                                         C.c_void_p]
        L.ref_particle_radii_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                               C.c_void_p, C.c_int64, C.c_double, C.c_void_p,
                                               C.c_void_p, C.c_int]
        L.ref_subfof.restype = C.c_int64
        L.ref_subfof.argtypes = [C.c_void_p, C.c_int64, C.c_float, C.c_int] + [C.c_void_p] * 7 + \
            [C.c_int64] + [C.c_void_p] * 3
        L.ref_next_closest.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p]
        L.ref_subfof_batch.restype = C.c_int64
        L.ref_subfof_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_float,
                                       C.c_int, C.c_int]
        L.ref_find_median_r.restype = C.c_float
        L.ref_find_median_r.argtypes = [C.c_void_p, C.c_int64, C.c_float, C.c_uint]

    @staticmethod
    def available(path=None) -> bool:
        return pathlib.Path(path or HERE / "_ref" / "libref_props.so").exists()

    def set_globals(self, prm, mass_defs=None) -> None:
        for name, val in (("FORCE_RES", prm.force_res), ("PARTICLE_MASS", prm.particle_mass),
                          ("Om", prm.Om), ("Ol", prm.Ol), ("W0", prm.W0), ("WA", prm.WA),
                          ("SCALE_NOW", prm.scale_now)):
            C.c_double.in_dll(self.lib, name).value = val
        if mass_defs is not None:
            # the char * globals point at buffers this object keeps alive
            self._defs = [C.create_string_buffer(d.encode()) for d in mass_defs]
            for name, d in zip(("MASS_DEFINITION", "MASS_DEFINITION2", "MASS_DEFINITION3",
                                "MASS_DEFINITION4", "MASS_DEFINITION5"), self._defs):
                C.c_void_p.in_dll(self.lib, name).value = C.addressof(d)

    BOUND_HALO = np.dtype({"names": ["pos", "corevel", "bulkvel", "min_vel_err",
                                     "min_bulkvel_err", "n_core", "is_sub"],
                           "formats": [("<f4", (6,)), ("<f4", (3,)), ("<f4", (3,)), "<f4",
                                       "<f4", "<i8", "<i4"],
                           "offsets": [0, 24, 36, 48, 52, 56, 64], "itemsize": 72})

    def particle_radii(self, particles: np.ndarray, tree: np.ndarray, base: int, ratio: float,
                       cap: int) -> np.ndarray:
        """The reference's own calc_particle_radii(h, h, h->pos, 0, 0, 0)
        for halo `base` (cap: room for the records it writes)."""
        assert particles.dtype.itemsize == 32 and particles.flags.c_contiguous
        assert tree.dtype.itemsize == 72 and tree.flags.c_contiguous
        po = np.zeros(max(1, cap), _record_dtype())
        n = self.lib.ref_particle_radii(_p(particles), _p(tree), tree.shape[0], base, ratio,
                                        _p(po))
        assert n <= cap
        return po[:n]

    def particle_radii_batch(self, particles, tree, base, offset, ratio, out, nthreads=0):
        """particle_radii for every halo of `base`, its records at
        out[offset[i]:], OpenMP dynamic over halos; returns the counts."""
        b = np.ascontiguousarray(base, np.int64)
        off = np.ascontiguousarray(offset, np.int64)
        cnt = np.zeros(b.shape[0], np.int64)
        self.lib.ref_particle_radii_batch(_p(particles), _p(tree), tree.shape[0], _p(b), _p(off),
                                          b.shape[0], ratio, _p(out), _p(cnt), nthreads)
        return cnt

    def bound_pass(self, po: np.ndarray, halo_row, prm):
        """The reference's own bound-pass steps (src/properties.c:537-550:
        compute_potential, potential-only strip, qsort with dist_compare,
        calculate_corevel, compute_kinetic_energy) on one call's records, in
        place.  Returns (kept, halo fields as BOUND_HALO)."""
        _check_records(po)
        self.set_globals(prm)
        bh = np.zeros(1, self.BOUND_HALO)
        for f in ("pos", "corevel", "bulkvel", "is_sub"):
            bh[f] = halo_row[f]
        kept = self.lib.ref_bound_pass(_p(po), po.shape[0], _p(bh))
        return kept, bh[0]

    def set_subfof_config(self, min_halo_particles=10, exact_ll_calc=0,
                          initial_metric_scaling=1.0) -> None:
        """MIN_HALO_PARTICLES, EXACT_LL_CALC, INITIAL_METRIC_SCALING
        (src/config.template.h:96,114,118)."""
        C.c_int64.in_dll(self.lib, "MIN_HALO_PARTICLES").value = min_halo_particles
        C.c_int64.in_dll(self.lib, "EXACT_LL_CALC").value = exact_ll_calc
        C.c_double.in_dll(self.lib, "INITIAL_METRIC_SCALING").value = initial_metric_scaling

    def subfof(self, parts: np.ndarray, thresh: float = 0.7, top: bool = True) -> dict:
        """The reference's sub-FOF step on one group (ref_subfof): tree
        order after norm_sd and the rebuild, the sample's next-closest
        distances, the linking length, and _find_subfofs_better2's own
        result (its particle array and sub-FOFs as (first, num_p)); also the
        sub-FOFs of _find_subfofs_at_r at that length on the tree order."""
        assert parts.dtype.itemsize == 32 and parts.flags.c_contiguous
        n = parts.shape[0]
        m = max(1, n)
        tree = np.zeros(m, parts.dtype)
        final = np.zeros(m, parts.dtype)
        sample = np.zeros(max(n, 10000), np.float32)
        ns = C.c_int64()
        tr = C.c_float()
        cap = n // 2 + 2
        ff, fn, af, an = (np.zeros(cap, np.int64) for _ in range(4))
        nat = C.c_int64()
        nf = self.lib.ref_subfof(_p(parts), n, thresh, int(top), C.byref(tr), _p(tree), _p(sample),
                                 C.byref(ns), _p(final), _p(ff), _p(fn), cap, _p(af), _p(an),
                                 C.byref(nat))
        return {"target_r": np.float32(tr.value), "tree_order": tree[:n],
                "sample_r": sample[:ns.value], "final": final[:n],
                "fofs": np.stack([ff[:nf], fn[:nf]], 1),
                "fofs_at_r": np.stack([af[:nat.value], an[:nat.value]], 1)}

    def subfof_batch(self, parts: np.ndarray, groups: np.ndarray, thresh: float = 0.7,
                     top: bool = True, nthreads: int = 0) -> int:
        """_find_subfofs_better2 on every group (first, num_p), in place,
        OpenMP dynamic over groups; returns the number of sub-FOFs."""
        assert parts.dtype.itemsize == 32 and parts.flags.c_contiguous
        f = np.ascontiguousarray(groups["first"], np.int64)
        n = np.ascontiguousarray(groups["num_p"], np.int64)
        return self.lib.ref_subfof_batch(_p(parts), _p(f), _p(n), f.shape[0], thresh, int(top),
                                         nthreads)

    def next_closest(self, pts: np.ndarray, queries: np.ndarray) -> np.ndarray:
        """fast3tree_find_next_closest_distance (src/fast3tree.c:992-1030)
        of each query on the 6D GROUPIES tree of pts (rebuilt in place)."""
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 6)
        out = np.zeros(q.shape[0], np.float32)
        self.lib.ref_next_closest(_p(pts), pts.shape[0], _p(q), q.shape[0], _p(out))
        return out

    def find_median_r(self, rad, frac: float, seed: int) -> np.float32:
        r = np.ascontiguousarray(rad, np.float32)
        return np.float32(self.lib.ref_find_median_r(_p(r), r.shape[0], frac, seed))

    def mass_definition(self, prm, mass_defs) -> np.ndarray:
        """particle_thresh_dens[0..4], particle_rvir_dens and
        particle_rvir_dens_z0."""
        self.set_globals(prm, mass_defs)
        out = np.zeros(7, np.float64)
        self.lib.ref_mass_definition(_p(out))
        return out

    def halo_props(self, po: np.ndarray, kept: int, pos, prm, thresh) -> np.ndarray:
        _check_records(po)
        self.set_globals(prm)
        ps = np.ascontiguousarray(pos, np.float32)
        th = np.ascontiguousarray(thresh, np.float64)
        out = np.zeros(1, self.DTYPE)
        self.lib.ref_halo_props(_p(po), kept, _p(ps), _p(th), _p(out))
        return out[0]

    def halo_props_batch(self, po: np.ndarray, halos: np.ndarray, prm, thresh,
                         nthreads: int = 0) -> np.ndarray:
        """ref_halo_props for every call (first, kept, pos of each halo row),
        OpenMP dynamic over calls."""
        _check_records(po)
        self.set_globals(prm)
        first = np.ascontiguousarray(halos["first"], np.int64)
        kept = np.ascontiguousarray(halos["kept"], np.int64)
        pos = np.ascontiguousarray(halos["pos"], np.float32)
        th = np.ascontiguousarray(thresh, np.float64)
        out = np.zeros(halos.shape[0], self.DTYPE)
        self.lib.ref_halo_props_batch(_p(po), _p(first), _p(kept), _p(pos), halos.shape[0],
                                      _p(th), _p(out), nthreads)
        return out
