"""Python access to libhalopot (the MI355X potential + bound pass).

Mirrors the reference's operator interface for this path:
  * ``compute_potential(po)`` -- src/potential.h:24, on an array of
    ``struct potential`` records (``catalog.POTENTIAL_DTYPE``); pe is written in
    place and the records are permuted into tree order, as the reference's
    compute_potential leaves them.
  * ``compute_kinetic_energy(po, vel_cen, pos_cen, scale_now)`` --
    src/potential.h:22-23.
  * ``set_config(params)`` -- the configuration globals the path reads
    (FORCE_RES, PARTICLE_MASS, Om, Ol, W0, WA, SCALE_NOW).
plus the batched path (``Plan``, whose ``props`` runs the halo property pass
of src/properties.c:384-510), ``gather_calls`` (the records of each call
from the halo hierarchy, src/groupies.c:212-268), ``mass_definitions`` and
``bound_pass_host``.

There is no CPU fallback: if libhalopot.so is missing or no gfx950 device is
usable, every call raises.
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

from . import catalog  # noqa: F401  (re-exported)
from .catalog import HALO_DTYPE, POTENTIAL_DTYPE, PROPS_DTYPE, Params  # noqa: F401

PKG = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG / "libhalopot.so"

ACC_COMPENSATED, ACC_NAIVE, ACC_F64 = 0, 1, 2
WALK_PAIRS, WALK_NODES = 0, 1
METHOD_TREE, METHOD_DIRECT, METHOD_AUTO = 0, 1, 2

NODE_OUT_DTYPE = np.dtype(
    {
        "names": ["lo", "hi", "com", "mass", "dmin2", "dmin", "first", "count", "left", "skip",
                  "parent", "pad"],
        "formats": [("<f4", (3,)), ("<f4", (3,)), ("<f4", (3,)), "<f4", "<f4", "<f4", "<i4",
                    "<i4", "<i4", "<i4", "<i4", "<i4"],
        "offsets": [0, 12, 24, 36, 40, 44, 48, 52, 56, 60, 64, 68],
        "itemsize": 72,
    }
)


class HalopotError(RuntimeError):
    pass


class Options(C.Structure):
    _fields_ = [("accumulate", C.c_int32), ("method", C.c_int32), ("use_graph", C.c_int32),
                ("timing", C.c_int32), ("count_work", C.c_int32), ("per_call", C.c_int32),
                ("walk", C.c_int32), ("reserved", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("ms_total", C.c_float), ("ms_build", C.c_float), ("ms_multipole", C.c_float),
                ("ms_walk", C.c_float), ("ms_direct", C.c_float), ("ms_unbind", C.c_float),
                ("nodes", C.c_int64), ("leaves", C.c_int64), ("levels", C.c_int64),
                ("max_depth", C.c_int64), ("walk_launches", C.c_int64),
                ("walk_groups", C.c_int64), ("visits", C.c_int64), ("p2p", C.c_int64),
                ("p2m", C.c_int64), ("records", C.c_int64), ("halos", C.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class MassDefs(C.Structure):
    _fields_ = [("thresh_dens", C.c_double * 5), ("rvir_dens", C.c_double),
                ("rvir_dens_z0", C.c_double)]

    def as_array(self) -> np.ndarray:
        """particle_thresh_dens[0..4], particle_rvir_dens, particle_rvir_dens_z0."""
        return np.array(list(self.thresh_dens) + [self.rvir_dens, self.rvir_dens_z0], np.float64)


TUNE_BUCKETS = 4
TUNE_BUCKET_MAX = (1024, 8192, 65536)  # calls of at most this many records; the last: more


def tune_bucket(count: int) -> int:
    """The call-size bucket of a call of `count` records."""
    for b, m in enumerate(TUNE_BUCKET_MAX):
        if count <= m:
            return b
    return len(TUNE_BUCKET_MAX)


class Tuning(C.Structure):
    _fields_ = [("big_threshold", C.c_int32 * TUNE_BUCKETS), ("wave_max", C.c_int32 * TUNE_BUCKETS),
                ("direct_max", C.c_int32), ("ms_potential", C.c_float)]

    def as_dict(self):
        return {"big_threshold": list(self.big_threshold), "wave_max": list(self.wave_max),
                "direct_max": self.direct_max, "ms_potential": self.ms_potential}


_LIB = None

EXPORTS = ("compute_potential", "compute_kinetic_energy", "hp_plan_create", "hp_plan_run",
           "hp_plan_potential", "hp_plan_stats", "hp_plan_destroy", "hp_plan_tune",
           "hp_plan_set_tuning", "hp_compute_potential_noalloc",
           "hp_mass_definitions", "hp_plan_props", "hp_gather_calls", "hp_find_subfofs",
           "hp_bound_pass_host",
           "hp_debug_tree", "hp_last_error", "hp_set_error_handler")
GLOBALS = ("FORCE_RES", "PARTICLE_MASS", "Om", "Ol", "W0", "WA", "SCALE_NOW")


def lib() -> C.CDLL:
    """Loads libhalopot.so (built in-tree by build.py); raises if absent."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise HalopotError(f"{LIB_PATH} is missing: build it (python {PKG / 'build.py'})")
        L = C.CDLL(str(LIB_PATH))
        L.compute_potential.argtypes = [C.c_void_p, C.c_int64]
        L.compute_potential.restype = None
        L.compute_kinetic_energy.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                             C.c_float]
        L.compute_kinetic_energy.restype = None
        L.hp_plan_create.argtypes = [C.POINTER(C.c_void_p), C.c_void_p, C.c_int64,
                                     C.POINTER(Options)]
        L.hp_plan_run.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.hp_plan_potential.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.hp_plan_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
        L.hp_plan_destroy.argtypes = [C.c_void_p]
        L.hp_plan_tune.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(Tuning)]
        L.hp_plan_set_tuning.argtypes = [C.c_void_p, C.POINTER(Tuning)]
        L.hp_compute_potential_noalloc.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
        L.hp_mass_definitions.argtypes = [C.POINTER(C.c_char_p), C.POINTER(MassDefs)]
        L.hp_gather_calls.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p,
                                      C.c_int64, C.c_double, C.c_void_p, C.c_int64, C.c_void_p,
                                      C.POINTER(C.c_int64)]
        L.hp_find_subfofs.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                      C.POINTER(SubfofConfig), C.c_void_p, C.c_int64,
                                      C.POINTER(C.c_int64)]
        L.hp_plan_props.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(MassDefs),
                                    C.c_void_p, C.c_void_p]
        L.hp_bound_pass_host.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                         C.c_void_p, C.POINTER(Options)]
        L.hp_debug_tree.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64]
        L.hp_debug_tree.restype = C.c_int64
        L.hp_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


# hp_error_handler: void (*)(int code, const char *file, int line)
ERROR_HANDLER = C.CFUNCTYPE(None, C.c_int, C.c_char_p, C.c_int)
_HANDLER_REF = None


def set_error_handler(fn) -> None:
    """hp_set_error_handler: fn(code, file, line) is called on a fatal error
    inside the void reference entry points (compute_potential,
    compute_kinetic_energy) when the program does not link the reference's
    rockstar_abort_impl (src/error.h:15).  None removes the handler."""
    global _HANDLER_REF
    L = lib()
    L.hp_set_error_handler.argtypes = [ERROR_HANDLER]
    L.hp_set_error_handler.restype = None
    cb = ERROR_HANDLER(lambda code, f, line: fn(code, f.decode(), line)) if fn else ERROR_HANDLER()
    L.hp_set_error_handler(cb)
    _HANDLER_REF = cb  # the C side keeps the pointer: keep the thunk alive


def last_error() -> str:
    return lib().hp_last_error().decode()


def _check(code: int) -> None:
    if code != 0:
        raise HalopotError(f"halopot error {code}: {lib().hp_last_error().decode()}")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def set_config(prm: Params) -> None:
    """Sets the configuration globals that libhalopot reads at call time."""
    L = lib()
    for name, val in (("FORCE_RES", prm.force_res), ("PARTICLE_MASS", prm.particle_mass),
                      ("Om", prm.Om), ("Ol", prm.Ol), ("W0", prm.W0), ("WA", prm.WA),
                      ("SCALE_NOW", prm.scale_now)):
        C.c_double.in_dll(L, name).value = val


def get_config() -> dict:
    L = lib()
    return {n: C.c_double.in_dll(L, n).value for n in GLOBALS}


def _records(po: np.ndarray) -> np.ndarray:
    if po.dtype != POTENTIAL_DTYPE or not po.flags.c_contiguous:
        raise TypeError("po must be a C-contiguous array of POTENTIAL_DTYPE records")
    return po


def compute_potential(po: np.ndarray) -> None:
    """src/potential.c:183-192 on the GPU; pe written in place, records
    permuted as the reference permutes them."""
    _records(po)
    lib().compute_potential(_ptr(po), po.shape[0])


NOALLOC_MAX_RECORDS = 32768


def compute_potential_noalloc(d_po: int, n: int, stream: int | None = None) -> None:
    """hp_compute_potential_noalloc: compute_potential on n device-resident
    records at d_po, with no device allocation (at most NOALLOC_MAX_RECORDS
    records; overlapping calls on one GPU race)."""
    _check(lib().hp_compute_potential_noalloc(C.c_void_p(d_po), int(n),
                                              C.c_void_p(stream) if stream else None))


def compute_kinetic_energy(po: np.ndarray, vel_cen, pos_cen, scale_now: float) -> None:
    """src/potential.c:194-211 on the GPU; ke written in place where ke >= 0."""
    _records(po)
    vc = np.ascontiguousarray(vel_cen, np.float32)
    pc = np.ascontiguousarray(pos_cen, np.float32)
    lib().compute_kinetic_energy(_ptr(po), po.shape[0], _ptr(vc), _ptr(pc), float(scale_now))


def options(accumulate=ACC_COMPENSATED, method=METHOD_TREE, use_graph=0, timing=0,
            count_work=0, per_call=0, walk=WALK_PAIRS) -> Options:
    o = Options()
    o.accumulate, o.method, o.use_graph, o.timing, o.count_work, o.per_call, o.walk = (
        accumulate, method, use_graph, timing, count_work, per_call, walk)
    return o


def gather_calls(tree: np.ndarray, base, d_particles: int, nparticles: int, d_po: int,
                 cap_records: int, include_host_ratio: float = 0.3):
    """hp_gather_calls: calc_particle_radii (src/groupies.c:212-268) for the
    halos `base` of the hierarchy `tree` (TREE_DTYPE), from the particle
    copies at device address d_particles into records at d_po.  Returns the
    call descriptors (HALO_DTYPE) and the record count."""
    from .catalog import TREE_DTYPE

    if tree.dtype != TREE_DTYPE or not tree.flags.c_contiguous:
        raise TypeError("tree must be a C-contiguous TREE_DTYPE array")
    b = np.ascontiguousarray(base, np.int64)
    calls = np.zeros(b.shape[0], HALO_DTYPE)
    n = C.c_int64()
    _check(lib().hp_gather_calls(_ptr(tree), tree.shape[0], _ptr(b), b.shape[0],
                                 C.c_void_p(d_particles), nparticles, float(include_host_ratio),
                                 C.c_void_p(d_po), cap_records, _ptr(calls), C.byref(n)))
    return calls, n.value


class SubfofConfig(C.Structure):
    """hp_subfof_config: FOF_FRACTION, INITIAL_METRIC_SCALING,
    MIN_HALO_PARTICLES, EXACT_LL_CALC (src/config.template.h)."""

    _fields_ = [("fof_fraction", C.c_double), ("initial_metric_scaling", C.c_double),
                ("min_halo_particles", C.c_int64), ("exact_ll_calc", C.c_int32),
                ("ordered_search", C.c_int32)]


FOF_GROUP_DTYPE = np.dtype({"names": ["first", "num_p", "top", "target_r", "first_sub", "num_sub"],
                            "formats": ["<i8", "<i8", "<i4", "<f4", "<i8", "<i8"],
                            "offsets": [0, 8, 16, 20, 24, 32], "itemsize": 40})
SUBFOF_DTYPE = np.dtype([("first", "<i8"), ("num_p", "<i8")])


def find_subfofs(d_particles: int, nparticles: int, groups: np.ndarray, fof_fraction=0.7,
                 min_halo_particles=10, exact_ll_calc=0, initial_metric_scaling=1.0,
                 ordered_search=0):
    """hp_find_subfofs: _find_subfofs_better2 (src/groupies.c:287-314) for
    every group (rows with first, num_p, top) of the particles at device
    address d_particles, in place.  Returns (groups with target_r, first_sub
    and num_sub filled, the sub-FOFs as SUBFOF_DTYPE rows)."""
    g = np.zeros(len(groups), FOF_GROUP_DTYPE)
    for f in ("first", "num_p"):
        g[f] = groups[f]
    g["top"] = groups["top"] if "top" in (groups.dtype.names or ()) else 1
    # a sub-FOF holds at least min_halo_particles (and at least 2) particles
    cap = int(np.sum(g["num_p"]) // max(2, min_halo_particles) + 1)
    subs = np.empty(cap, SUBFOF_DTYPE)
    n = C.c_int64()
    cfg = SubfofConfig(fof_fraction, initial_metric_scaling, min_halo_particles, exact_ll_calc,
                       ordered_search)
    _check(lib().hp_find_subfofs(C.c_void_p(d_particles), nparticles, _ptr(g), g.shape[0],
                                 C.byref(cfg), _ptr(subs), cap, C.byref(n)))
    return g, subs[:n.value]


def mass_definitions(defs=None) -> MassDefs:
    """calc_mass_definition (src/groupies.c:124-142) with the current
    configuration globals: particle_thresh_dens[0..4] for the five mass
    definitions (None: the halo finder's defaults vir, 200b, 200c, 500c,
    2500c), particle_rvir_dens and particle_rvir_dens_z0."""
    arr = (C.c_char_p * 5)()
    for k in range(5):
        arr[k] = defs[k].encode() if defs is not None and defs[k] is not None else None
    out = MassDefs()
    _check(lib().hp_mass_definitions(arr, C.byref(out)))
    return out


def bound_pass_host(po: np.ndarray, halos: np.ndarray, opts: Options | None = None,
                    want_bound: bool = True):
    """Batched bound pass on host arrays (copied in and out).  Returns the
    bound flags (one byte per record) or None."""
    _records(po)
    if halos.dtype != HALO_DTYPE or not halos.flags.c_contiguous:
        raise TypeError("halos must be a C-contiguous HALO_DTYPE array")
    bound = np.zeros(po.shape[0], np.uint8) if want_bound else None
    o = opts or options()
    _check(lib().hp_bound_pass_host(_ptr(po), po.shape[0], _ptr(halos), halos.shape[0],
                                    _ptr(bound) if want_bound else None, C.byref(o)))
    return bound


def debug_tree(po: np.ndarray):
    """GPU build of one call's tree: (nodes, perm) where perm[i] is the record
    at tree position i."""
    _records(po)
    n = po.shape[0]
    perm = np.zeros(n, np.int32)
    cap = 2 * n + 2
    nodes = np.zeros(cap, NODE_OUT_DTYPE)
    nn = lib().hp_debug_tree(_ptr(po), n, _ptr(perm), _ptr(nodes), cap)
    if nn < 0:
        _check(int(nn))
    return nodes[:nn], perm


class Plan:
    """A planned batch (hp_plan_create) run on device-resident data.

    Records and halos are passed as raw device pointers to the C ABI (for
    example halopot.hipmem.DeviceArray addresses)."""

    def __init__(self, halos: np.ndarray, opts: Options | None = None):
        if halos.dtype != HALO_DTYPE or not halos.flags.c_contiguous:
            raise TypeError("halos must be a C-contiguous HALO_DTYPE array")
        self.opts = opts or options()
        self._h = C.c_void_p()
        _check(lib().hp_plan_create(C.byref(self._h), _ptr(halos), halos.shape[0],
                                    C.byref(self.opts)))

    def tune(self, d_po: int, reps: int = 3) -> dict:
        """hp_plan_tune: times the potential pass (pe overwritten) for the
        result-neutral launch settings, per call-size bucket, and keeps the
        fastest.  Returns the per-bucket choice (lists indexed by bucket)."""
        t = Tuning()
        _check(lib().hp_plan_tune(self._h, C.c_void_p(d_po), int(reps), C.byref(t)))
        return t.as_dict()

    def set_tuning(self, big_threshold, wave_max, direct_max: int | None = None) -> None:
        """hp_plan_set_tuning: one of the tuner's settings, applied untimed.
        big_threshold and wave_max: one value for every bucket, or one per
        bucket."""
        def per_bucket(v):
            v = list(v) if hasattr(v, "__len__") else [v] * TUNE_BUCKETS
            if len(v) != TUNE_BUCKETS:
                raise ValueError(f"{TUNE_BUCKETS} values, one per bucket")
            return v
        t = Tuning()
        t.big_threshold[:] = per_bucket(big_threshold)
        t.wave_max[:] = per_bucket(wave_max)
        t.direct_max = 10 if direct_max is None else direct_max
        _check(lib().hp_plan_set_tuning(self._h, C.byref(t)))

    def run(self, d_po: int, d_halos: int, d_bound: int | None = None,
            stream: int | None = None) -> None:
        """hp_plan_run; stream: a hipStream_t whose queued work the pass
        follows (and which waits for the pass), or None."""
        _check(lib().hp_plan_run(self._h, C.c_void_p(d_po), C.c_void_p(d_halos),
                                 C.c_void_p(d_bound) if d_bound else None,
                                 C.c_void_p(stream) if stream else None))

    def potential(self, d_po: int, stream: int | None = None) -> None:
        _check(lib().hp_plan_potential(self._h, C.c_void_p(d_po),
                                       C.c_void_p(stream) if stream else None))

    def props(self, d_po: int, d_halos: int, d_props: int, defs: MassDefs | None = None,
              stream: int | None = None) -> None:
        """hp_plan_props: the halo property pass (src/properties.c:384-510)
        on the results of the last run(); d_props receives one PROPS_DTYPE
        record per call."""
        md = defs if defs is not None else mass_definitions()
        _check(lib().hp_plan_props(self._h, C.c_void_p(d_po), C.c_void_p(d_halos), C.byref(md),
                                   C.c_void_p(d_props), C.c_void_p(stream) if stream else None))

    def stats(self) -> dict:
        s = Stats()
        _check(lib().hp_plan_stats(self._h, C.byref(s)))
        return s.as_dict()

    def close(self) -> None:
        if self._h:
            lib().hp_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
