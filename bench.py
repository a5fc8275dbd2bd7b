#!/usr/bin/env python3
"""Throughput of the batched potential + bound pass on one MI355X.

A step is one hp_plan_run over the whole synthetic catalogue: the
Barnes-Hut potential of every call (src/potential.c:183-192) and the bound
pass after it (src/properties.c:539-550: drop potential-only records, sort by
r2, core / bulk velocity, kinetic energy, bound flag).  Records and halo
descriptors are resident in HBM before the timed region; before each step
the pristine input is restored with a device-to-device copy (the pass
reorders records in place), waited for, and kept out of the step's clock.

Default workload: the one-GPU catalogue, 16,777,216 particles in a
250 Mpc/h box of which 40% (6,710,886) sit in 50,000 NFW halos with counts
from n^-1.9 on [20, 500000] (seeded); every halo is one compute_potential
call, all of them in one batched plan replayed from hipGraphs.  Beside it,
the same catalogue through the plain per-halo GPU path (the drop-in
compute_potential / compute_kinetic_energy called per halo from a pool of
host threads, as the reference's OpenMP driver calls them, and a per-call
plan on one stream), through the host-buffer entry point (PCIe included),
and through the reference's CPU path (its own compute_potential /
compute_kinetic_energy from oracle/_ref, OpenMP over halos) on a bounded
sample of host time.  After the timed steps, the halo property pass
(hp_plan_props) is timed on the last step's results, with the reference's
own property code beside it (reported under "halo_props", not in "value").
Outside the timed region, pe of the benched arithmetic is compared with the
FP64 evaluation of the same trees ("accuracy"); the run fails above 1e-4.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "mpi-rockstar_amd"))

METRIC = "halo-particle potentials/s, batched vs per-halo GPU, ≤1e-4 rel err vs FP64"
UNIT = "halo-particle potentials/s"
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector, published (packed / dual-issue FP32)
FP32_NONPACKED_TFLOPS = 78.6  # the same without packed FP32, which this build does not emit
HBM_PEAK_GBS = 8000.0
ACCURACY_BAR = 1e-4        # the metric's relative-error bar vs FP64

WORKLOADS = {
    "c2": "batching: 2000 NFW halos, n ~ n^-1.9 on [20, 200000], c ~ lognormal(8, 0.3), "
          "+5% unbound contaminants, seed 0x5EED0002",
    "c5": "one-GPU catalogue: 16,777,216 particles / 250 Mpc/h, 40% in 50,000 NFW halos "
          "(n^-1.9 on [20, 500000] scaled to 6,710,886), seed 0x5EED0005",
    "c3": "deep tree: 2,000,000-particle NFW host (c=6, r_vir=2) + 30 sub-halos, 31 calls",
    "c1": "one NFW halo of 10,000 + 1,000 background particles",
}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_threads() -> int:
    """Host threads for the CPU legs: OMP_NUM_THREADS when set (the GPU box
    sets it to its CPU share for one GPU, 16), else the cores this process
    may run on, at most 16."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        return int(env)
    try:
        return min(16, len(os.sched_getaffinity(0)))
    except AttributeError:
        return min(16, os.cpu_count() or 1)


def cpu_model() -> dict:
    """The host CPU as lscpu names it (from /proc/cpuinfo), and its counts."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return {"model": model, "logical_cpus": os.cpu_count(), "cpus_allowed": allowed}


def walk_traffic(workload: str):
    """HBM bytes per k_walk launch for this workload, from the committed
    FETCH_SIZE / WRITE_SIZE counter runs (scaled by a copy of known size),
    or None when no such run exists."""
    f = ROOT / "profiles" / f"hbm_traffic.{workload}_catalogue.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith(f"traffic k_walk {workload}:"):
            return float(line.split(":")[1])
    return None


def run_cpu_baseline(cat, seconds: float):
    """The reference's CPU path on the host cores: oracle/_ref's compiled
    compute_potential + compute_kinetic_energy inside the OpenMP-over-halos
    bound pass, repeated over the whole catalogue until `seconds` elapse."""
    import numpy as np

    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    O = oracle.Restatement()
    nt = cpu_threads()
    if oracle.Reference.available():
        R = oracle.Reference()
        kind = "reference"

        def one():
            po = cat.po.copy()
            h = cat.halos.copy()
            O.bound_pass_batch_with(po, h, cat.params, R, nthreads=nt)
    else:
        kind = "port"

        def one():
            po = cat.po.copy()
            h = cat.halos.copy()
            O.bound_pass_batch(po, h, cat.params, nthreads=nt)
    passes = 0
    busy = 0.0
    while busy < seconds or passes == 0:
        t0 = time.perf_counter()
        one()
        busy += time.perf_counter() - t0
        passes += 1
    out = {
        "value": passes * cat.n_records / busy,
        "unit": UNIT,
        "cores": nt,
        "kind": kind,
        "sample": f"whole catalogue x {passes} pass(es), {busy:.1f} s wall on {nt} threads "
                  f"(OpenMP dynamic over halos, as src/mpi_main.cpp:1377)",
        "host": cpu_model(),
    }
    # the same on one thread: a bounded sample of calls, every k-th call so
    # that the sample keeps the catalogue's size distribution
    k = max(1, int(round(nt * 1.5)))
    sel = np.arange(0, cat.n_halos, k)
    counts = cat.halos["count"][sel]
    sub_h = cat.halos[sel].copy()
    sub_h["first"] = np.concatenate([[0], np.cumsum(counts)[:-1]])
    idx = np.concatenate([np.arange(f, f + c) for f, c in zip(cat.halos["first"][sel], counts)])
    sub_po = cat.po[idx]
    t0 = time.perf_counter()
    if kind == "reference":
        O.bound_pass_batch_with(sub_po.copy(), sub_h.copy(), cat.params, R, nthreads=1)
    else:
        O.bound_pass_batch(sub_po.copy(), sub_h.copy(), cat.params, nthreads=1)
    e1 = time.perf_counter() - t0
    out["one_thread"] = {"value": idx.shape[0] / e1, "unit": UNIT, "cores": 1,
                         "sample": f"every {k}th call ({sel.shape[0]} calls, {idx.shape[0]} "
                                   f"records), one pass, {e1:.1f} s"}
    return out


def run_drop_in(cat, seconds: float):
    """The plain per-halo GPU path as the reference's driver calls it: the
    drop-in compute_potential and compute_kinetic_energy (libhalopot's own
    entry points, host records in and out) once per halo, from a pool of
    host threads taking halos dynamically (src/mpi_main.cpp:1369-1402), each
    thread with its own engine and stream.  Only the two library calls run
    per halo (the caller's strip, sort and corevel scan are CPU work of the
    reference, not of this path)."""
    import threading

    import numpy as np

    import halopot

    nt = cpu_threads()
    L = halopot.lib()
    po = np.ascontiguousarray(cat.po.copy())
    first = cat.halos["first"].astype(np.int64)
    count = cat.halos["count"].astype(np.int64)
    centre = np.ascontiguousarray(cat.halos["pos"][:, :3], np.float32)
    vel = np.ascontiguousarray(cat.halos["bulkvel"], np.float32)
    base = po.ctypes.data
    vbase, cbase = vel.ctypes.data, centre.ctypes.data  # float[3] per halo
    rec = po.dtype.itemsize
    scale = float(cat.params.scale_now)
    lock = threading.Lock()
    nxt = [0]

    # one pool for every pass, as the reference's OpenMP threads: each keeps
    # its engine (and stream) from pass to pass
    start, done = threading.Barrier(nt + 1), threading.Barrier(nt + 1)
    stop, errors = [False], []

    def worker():
        while True:
            start.wait()
            if stop[0]:
                return
            try:
                while True:
                    with lock:
                        h = nxt[0]
                        nxt[0] += 1
                    if h >= cat.n_halos:
                        break
                    n = int(count[h])
                    if n <= 0:
                        continue
                    ptr = base + int(first[h]) * rec
                    if n > 1:
                        L.compute_potential(ptr, n)
                    L.compute_kinetic_energy(ptr, n, vbase + 12 * h, cbase + 12 * h, scale)
            except Exception as exc:  # reported by the pass, after every thread is back
                errors.append(exc)
            done.wait()

    ts = [threading.Thread(target=worker, daemon=True) for _ in range(nt)]
    for t in ts:
        t.start()

    def one():
        nxt[0] = 0
        start.wait()
        done.wait()
        if errors:
            raise errors[0]

    one()  # warm: every thread's engine and scratch
    passes, busy = 0, 0.0
    while busy < seconds or passes == 0:
        t0 = time.perf_counter()
        one()
        busy += time.perf_counter() - t0
        passes += 1
    stop[0] = True
    start.wait()
    for t in ts:
        t.join()
    return {"value": passes * cat.n_records / busy, "unit": UNIT,
            "ms_per_pass": 1e3 * busy / passes, "passes": passes, "threads": nt,
            "calls": "compute_potential + compute_kinetic_energy per halo, host buffers"}


def run_gather(args):
    """Record assembly (hp_gather_calls, src/groupies.c:212-268) on seeded
    FOF groups with sub-halo hierarchies: every halo with particles is a call
    whose records are its subtree (and its host's, potential-only, where the
    host rule holds).  Timed per call of the entry point (synchronous, host
    planning and table upload included); the reference's own
    calc_particle_radii beside it (oracle/_ref, OpenMP over halos)."""
    import numpy as np

    import halopot
    from halopot import catalog
    from halopot.hipmem import DeviceArray, synchronize

    H = catalog.hierarchy(n_groups=args.gather_groups)
    t = H.tree
    ncp = t["num_child_particles"]
    par = t["sub_of"][H.base]
    joined = (par > -1) & (ncp[np.maximum(par, 0)] * 0.3 < ncp[H.base])
    counts = np.where(joined, ncp[np.maximum(par, 0)], ncp[H.base])
    cap = int(counts.sum())
    d_parts = DeviceArray.from_numpy(H.particles)
    d_po = DeviceArray.zeros(cap * 48)
    halopot.gather_calls(t, H.base, d_parts.address, H.particles.shape[0], d_po.address, cap)
    reps = max(5, args.steps)
    t0 = time.perf_counter()
    for _ in range(reps):
        _, n = halopot.gather_calls(t, H.base, d_parts.address, H.particles.shape[0],
                                    d_po.address, cap)
    synchronize()
    e = (time.perf_counter() - t0) / reps
    d_parts.free()
    d_po.free()
    res = {"ms_per_pass": 1e3 * e, "value": n / e, "unit": "records/s", "records": int(n),
           "calls": int(H.base.shape[0]), "particles": int(H.particles.shape[0]),
           "bytes_per_record": 80, "hbm_gbs_wall": 80.0 * n / e / 1e9,
           "workload": f"{args.gather_groups} FOF groups (n ~ n^-1.9 on [200, 50000]) with "
                       f"sub-halos to depth 3, seed 0x5EED0006"}
    log(f"record assembly: {1e3 * e:.3f} ms per pass, {n} records, {n / e:.4g} records/s")
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle

        if oracle.ReferenceProps.available():
            R = oracle.ReferenceProps()
            off = np.concatenate([[0], np.cumsum(counts)[:-1]])
            out = np.zeros(cap, catalog.POTENTIAL_DTYPE)
            nt = cpu_threads()
            passes, busy = 0, 0.0
            while busy < 2.0 or passes == 0:
                t0 = time.perf_counter()
                R.particle_radii_batch(H.particles, t, H.base, off, 0.3, out, nthreads=nt)
                busy += time.perf_counter() - t0
                passes += 1
            res["cpu_baseline"] = {"value": n * passes / busy, "unit": "records/s", "cores": nt,
                                   "kind": "reference", "ms_per_pass": 1e3 * busy / passes,
                                   "sample": f"whole set x {passes} pass(es), {busy:.1f} s on "
                                             f"{nt} threads"}
            log(f"record assembly on the CPU (reference, {nt} threads): "
                f"{1e3 * busy / passes:.1f} ms per pass")
    return res


def run_subfof(args):
    """The phase-space sub-FOF step (hp_find_subfofs, replacing
    _find_subfofs_better2, src/groupies.c:287-314) over seeded FOF groups
    (hosts with sub-halos distinct in phase space).  Each pass starts from
    the pristine particles (restored device to device, untimed) because the
    step normalises and reorders them in place.  Timed per call of the entry
    point (synchronous: host sample draws, table upload, the level-by-level
    tree build with its per-level count readback, results back); the
    reference's own _find_subfofs_better2 beside it (oracle/_ref, OpenMP over
    groups)."""
    import numpy as np

    import halopot
    from halopot import catalog
    from halopot.hipmem import DeviceArray, synchronize

    fg = catalog.fof_groups(n_groups=args.subfof_groups)
    n = int(fg.particles.shape[0])
    d0 = DeviceArray.from_numpy(fg.particles)
    d1 = DeviceArray.zeros(n * 32)
    nbytes = n * 32
    import ctypes as C

    from halopot.hipmem import hip

    def restore():  # device to device (hipMemcpyDeviceToDevice = 3), complete before use
        assert hip().hipMemcpy(C.c_void_p(d1.address), C.c_void_p(d0.address), nbytes, 3) == 0
        assert hip().hipDeviceSynchronize() == 0

    restore()
    g, subs = halopot.find_subfofs(d1.address, n, fg.groups)  # warm
    reps = max(5, args.steps // 2)
    busy = 0.0
    for _ in range(reps):
        restore()
        synchronize()
        t0 = time.perf_counter()
        g, subs = halopot.find_subfofs(d1.address, n, fg.groups)
        busy += time.perf_counter() - t0
    e = busy / reps
    d0.free()
    d1.free()
    res = {"ms_per_pass": 1e3 * e, "value": n / e, "unit": "particles/s through the sub-FOF step",
           "particles": n, "groups": int(fg.groups.shape[0]), "subfofs": int(subs.shape[0]),
           "largest_group": int(fg.groups["num_p"].max()),
           "workload": f"{args.subfof_groups} FOF groups (hosts n ~ n^-1.9 on [200, 50000] with "
                       f"sub-halos to depth 3), seed 0x5EED0007"}
    log(f"sub-FOF step: {1e3 * e:.3f} ms per pass, {n} particles in {fg.groups.shape[0]} groups, "
        f"{subs.shape[0]} sub-FOFs, {n / e:.4g} particles/s")
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle

        if oracle.ReferenceProps.available():
            R = oracle.ReferenceProps()
            R.set_subfof_config()
            nt = cpu_threads()
            passes, busy = 0, 0.0
            while busy < 2.0 or passes == 0:
                parts = fg.particles.copy()
                t0 = time.perf_counter()
                ns = R.subfof_batch(parts, fg.groups, nthreads=nt)
                busy += time.perf_counter() - t0
                passes += 1
            res["cpu_baseline"] = {"value": n * passes / busy, "unit": res["unit"], "cores": nt,
                                   "kind": "reference", "ms_per_pass": 1e3 * busy / passes,
                                   "sample": f"whole set x {passes} pass(es), {busy:.1f} s on "
                                             f"{nt} threads (OpenMP dynamic over groups)"}
            res["cpu_subfofs"] = int(ns)
            log(f"sub-FOF step on the CPU (reference, {nt} threads): "
                f"{1e3 * busy / passes:.1f} ms per pass, {ns} sub-FOFs")
    return res


def run_props(plan, cat, d_po, d_h, args):
    """The halo property pass (hp_plan_props, src/properties.c:384-510) on
    the last step's results: its own time per pass over the kept records and,
    beside it, the reference's own property code (oracle/_ref, OpenMP over
    halos) on the same records for a bounded sample of host time."""
    import numpy as np

    import halopot
    from halopot import catalog
    from halopot.hipmem import DeviceArray, synchronize

    md = halopot.mass_definitions()
    d_p = DeviceArray.zeros(cat.n_halos * catalog.PROPS_DTYPE.itemsize)
    plan.props(d_po.address, d_h.address, d_p.address, md)  # warm
    reps = max(5, args.steps)
    t = time.perf_counter()
    for _ in range(reps):
        plan.props(d_po.address, d_h.address, d_p.address, md)  # synchronous
    synchronize()
    e = time.perf_counter() - t
    h_g = d_h.download(np.empty_like(cat.halos))
    kept = int(h_g["kept"].sum())
    res = {"ms_per_pass": 1e3 * e / reps, "value": kept * reps / e,
           "unit": "kept records/s through the property pass", "kept_records": kept,
           "calls": cat.n_halos, "passes": reps,
           "scope": "src/properties.c:384-510 for bound = 0 and 1: masses, vmax, offsets, "
                    "J, energy, spins, calc_shape (both windows), calc_scale_radius, "
                    "pseudo-evolution masses"}
    log(f"halo property pass: {res['ms_per_pass']:.3f} ms/pass, {res['value']:.4g} kept records/s")
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle

        po_g = d_po.download(np.empty_like(cat.po))
        nt = cpu_threads()
        if oracle.ReferenceProps.available():
            R = oracle.ReferenceProps()
            kind = "reference"

            def one():
                R.halo_props_batch(po_g, h_g, cat.params, md.as_array(), nthreads=nt)
        else:
            O = oracle.Restatement()
            kind = "port"
            buf = np.zeros(cat.n_halos, catalog.PROPS_DTYPE)

            def one():
                O.halo_props_batch(po_g, h_g, cat.params, md.as_array(), buf, nthreads=nt)
        passes, busy = 0, 0.0
        while busy < min(4.0, args.cpu_seconds) or passes == 0:
            t = time.perf_counter()
            one()
            busy += time.perf_counter() - t
            passes += 1
        res["cpu_baseline"] = {"value": kept * passes / busy, "unit": res["unit"], "cores": nt,
                               "kind": kind, "ms_per_pass": 1e3 * busy / passes,
                               "sample": f"whole catalogue x {passes} pass(es), {busy:.1f} s "
                                         f"on {nt} threads (OpenMP dynamic over halos)"}
        res["gpu_vs_cpu"] = res["value"] / res["cpu_baseline"]["value"]
        log(f"halo property pass on the CPU ({kind}, {nt} threads): "
            f"{res['cpu_baseline']['ms_per_pass']:.1f} ms/pass")
        if kind == "reference":
            # the restatement of the same pass (the oracle's C code, pinned
            # bit for bit to the reference's)
            O = oracle.Restatement()
            buf = np.zeros(cat.n_halos, catalog.PROPS_DTYPE)
            passes, busy = 0, 0.0
            while busy < 2.0 or passes == 0:
                t = time.perf_counter()
                O.halo_props_batch(po_g, h_g, cat.params, md.as_array(), buf, nthreads=nt)
                busy += time.perf_counter() - t
                passes += 1
            res["cpu_port_same_fields"] = {"ms_per_pass": 1e3 * busy / passes, "cores": nt,
                                           "kind": "port", "passes": passes}
            log(f"halo property pass, CPU restatement: {1e3 * busy / passes:.1f} ms/pass")
    d_p.free()
    return res


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c5", choices=sorted(WORKLOADS))
    ap.add_argument("--accumulate", default="compensated", choices=["compensated", "naive", "f64"])
    ap.add_argument("--no-graph", action="store_true", help="eager launches instead of hipGraph replay")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (hp_plan_tune)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-per-halo", action="store_true", help="skip the per-halo GPU path")
    ap.add_argument("--no-host", action="store_true", help="skip the host-buffer (PCIe) leg")
    ap.add_argument("--no-props", action="store_true", help="skip the halo property pass leg")
    ap.add_argument("--no-gather", action="store_true", help="skip the record assembly leg")
    ap.add_argument("--gather-groups", type=int, default=4000)
    ap.add_argument("--no-subfof", action="store_true", help="skip the sub-FOF leg")
    ap.add_argument("--subfof-groups", type=int, default=4000)
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for "
              f"--gpus {args.gpus}); the calls would be dealt to the GPUs whole, each "
              f"GPU running this one-GPU step on its share",
              file=sys.stderr)
        return 2

    import numpy as np

    import halopot
    from halopot import catalog
    from halopot.hipmem import DeviceArray, synchronize

    t = time.perf_counter()
    cat = catalog.make(args.workload)
    log(f"workload {args.workload}: {cat.n_records} records in {cat.n_halos} calls "
        f"({time.perf_counter() - t:.1f} s to generate)")
    halopot.set_config(cat.params)
    acc = {"compensated": halopot.ACC_COMPENSATED, "naive": halopot.ACC_NAIVE,
           "f64": halopot.ACC_F64}[args.accumulate]

    d_src = DeviceArray.from_numpy(cat.po)
    d_hsrc = DeviceArray.from_numpy(cat.halos)
    d_po = DeviceArray(cat.po.nbytes)
    d_h = DeviceArray(cat.halos.nbytes)
    d_b = DeviceArray.zeros(cat.n_records)
    hip = halopot.hipmem.hip()

    def restore():  # a device-to-device copy can return before it is done: wait for it
        assert hip.hipMemcpy(d_po.ptr, d_src.ptr, cat.po.nbytes, halopot.hipmem.D2D) == 0
        assert hip.hipMemcpy(d_h.ptr, d_hsrc.ptr, cat.halos.nbytes, halopot.hipmem.D2D) == 0
        assert hip.hipDeviceSynchronize() == 0

    # work counters of the reference's walk decisions (untimed)
    plan_c = halopot.Plan(cat.halos, halopot.options(accumulate=acc, count_work=1))
    restore()
    plan_c.potential(d_po.address)
    wst = plan_c.stats()
    plan_c.close()
    flops = 11 * wst["p2p"] + 10 * wst["p2m"]
    log(f"walk work: {wst['p2p']} P2P pairs, {wst['p2m']} P2M terms, {wst['visits']} "
        f"wavefront node visits, {wst['nodes']} nodes, depth {wst['max_depth']}")

    use_graph = 0 if args.no_graph else 1
    plan = halopot.Plan(cat.halos, halopot.options(accumulate=acc, timing=1, use_graph=use_graph))
    tuning = None
    if not args.no_tune:
        restore()
        tuning = plan.tune(d_po.address, reps=2)
        log(f"autotuner, per call-size bucket (<= 1024, 8192, 65536, more records): big "
            f"threshold {tuning['big_threshold']}, wavefront subtrees up to {tuning['wave_max']} "
            f"points ({tuning['ms_potential']:.3f} ms per potential pass)")
    for i in range(args.warmup):
        restore()
        plan.run(d_po.address, d_h.address, d_b.address)
    synchronize()
    log(f"warmup done ({args.warmup} steps)")
    walk_ms, phase = [], {"build": 0.0, "multipole": 0.0, "walk": 0.0, "unbind": 0.0}
    # The pass sorts the records in place, so every step starts from a fresh
    # copy of the input; the copy is finished before the step's clock starts
    # and is not part of the timed work.
    elapsed = 0.0
    for i in range(args.steps):
        restore()
        synchronize()
        t0 = time.perf_counter()
        plan.run(d_po.address, d_h.address, d_b.address)
        synchronize()
        elapsed += time.perf_counter() - t0
        st = plan.stats()
        walk_ms.append(st["ms_walk"])
        for k in phase:
            phase[k] += st["ms_" + k]
    props = None if args.no_props else run_props(plan, cat, d_po, d_h, args)
    # The walk's launch duration for the roofline (after the property pass,
    # which reads the last step's results): a recorded step walks the calls
    # in two launches on two branches (the smaller calls' beside the split
    # levels' multipoles), so its walk phase is not one launch.  The same
    # plan's eager potential pass launches the walk once over every call,
    # timed by HIP events on the plan's stream around that launch.
    walk_eager = []
    for i in range(5):
        restore()
        plan.potential(d_po.address)
        synchronize()
        walk_eager.append(plan.stats()["ms_walk"])
    plan.close()

    # accuracy of the benched path (untimed): pe of the same plan's trees and
    # decisions evaluated in FP64 (HP_ACC_F64) against the benched arithmetic
    pes = {}
    for a in (acc, halopot.ACC_F64):
        pl = halopot.Plan(cat.halos, halopot.options(accumulate=a))
        if tuning is not None:
            pl.set_tuning(tuning["big_threshold"], tuning["wave_max"], tuning["direct_max"])
        restore()
        pl.potential(d_po.address)
        synchronize()
        pes[a] = d_po.download(np.empty_like(cat.po))["pe"]
        pl.close()
    live = pes[halopot.ACC_F64] != 0
    rerr = np.abs(pes[acc][live] - pes[halopot.ACC_F64][live]) / np.abs(pes[halopot.ACC_F64][live])
    accuracy = {"pe_vs_fp64_same_trees": {"median": float(np.median(rerr)),
                                          "p99": float(np.quantile(rerr, 0.99)),
                                          "max": float(rerr.max()), "records": int(live.sum())},
                "bar": ACCURACY_BAR, "ok": bool(rerr.max() <= ACCURACY_BAR)}
    log(f"accuracy: pe vs FP64 of the same trees, median {np.median(rerr):.3g}, max "
        f"{rerr.max():.3g} (bar {ACCURACY_BAR:g})")
    ms_step = 1e3 * elapsed / args.steps
    value = cat.n_records * args.steps / elapsed
    log(f"batched: {ms_step:.3f} ms/step, {value:.4g} records/s; phases per step (ms): " +
        ", ".join(f"{k} {v / args.steps:.3f}" for k, v in phase.items()))

    walk_avg = sum(walk_eager) / len(walk_eager)
    achieved = flops / (walk_avg * 1e-3) / 1e12
    roofline = {
        "bound": "valu",
        "achieved": achieved,
        "peak": FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / FP32_PEAK_TFLOPS,
        "frac_nonpacked": achieved / FP32_NONPACKED_TFLOPS,
        "peak_nonpacked": FP32_NONPACKED_TFLOPS,
        "traffic": walk_traffic(args.workload),
        "kernel": "k_walk (Barnes-Hut group walk)",
        "bound_detail": "FP32 VALU + v_rsq issue (no dense contraction: MFMA unused); peak is "
                        "the published FP32 vector figure (packed FP32), peak_nonpacked the "
                        "rate without packed FP32, which this build does not emit",
        "flops_per_launch": flops,
        "flop_convention": "11 per P2P pair, 10 per P2M term, counts of the reference's walk",
        "walk_ms_avg": walk_avg,
        "walk_ms_source": "5 eager potential passes of the benched plan: one k_walk launch over "
                          "every call, HIP events on the plan's stream",
        "walk_phase_ms_step": sum(walk_ms) / len(walk_ms),
    }

    out = {
        "metric": METRIC,
        "value": value,
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload + ": " + WORKLOADS[args.workload],
                   "records": cat.n_records, "calls": cat.n_halos,
                   "accumulate": args.accumulate, "graph": bool(use_graph),
                   "tuning": tuning,
                   "parallelism": "one GPU"},
        "roofline": roofline,
        "accuracy": accuracy,
    }

    if props is not None:
        out["halo_props"] = props
    if not args.no_gather:
        out["record_assembly"] = run_gather(args)
    if not args.no_subfof:
        out["subfof"] = run_subfof(args)

    if not args.no_per_halo:
        plan_p = halopot.Plan(cat.halos, halopot.options(accumulate=acc, per_call=1))
        restore()
        plan_p.run(d_po.address, d_h.address, d_b.address)  # warm
        synchronize()
        reps = 0
        tp = time.perf_counter()
        while reps == 0 or (time.perf_counter() - tp < 2.0 and reps < args.steps):
            restore()
            plan_p.run(d_po.address, d_h.address, d_b.address)
            reps += 1
        synchronize()
        ep = time.perf_counter() - tp
        plan_p.close()
        vp = cat.n_records * reps / ep
        out["per_halo_serial"] = {"value": vp, "unit": UNIT, "ms_per_pass": 1e3 * ep / reps,
                                  "passes": reps,
                                  "path": "per-call plan: every call its own launches, one "
                                          "after another on one stream (device buffers)"}
        log(f"per-halo GPU path, serial: {1e3 * ep / reps:.1f} ms/pass, {vp:.4g} records/s")
        di = run_drop_in(cat, min(6.0, args.cpu_seconds))
        out["per_halo_drop_in"] = di
        out["batched_vs_per_halo"] = value / di["value"]
        log(f"per-halo GPU path, drop-in from {di['threads']} host threads: "
            f"{di['ms_per_pass']:.1f} ms/pass, {di['value']:.4g} records/s; "
            f"batched/per-halo = {value / di['value']:.1f}x")

    if not args.no_host:
        # the same pass from pageable host buffers: H2D, plan, run, D2H
        po_h = cat.po.copy()
        h_h = cat.halos.copy()
        th = time.perf_counter()
        halopot.bound_pass_host(po_h, h_h, halopot.options(accumulate=acc))
        eh = time.perf_counter() - th
        out["host_buffers"] = {"value": cat.n_records / eh, "unit": UNIT, "ms_per_pass": 1e3 * eh,
                               "includes": "plan creation, device scratch, H2D + D2H over PCIe"}
        log(f"host-buffer pass (PCIe included): {1e3 * eh:.1f} ms, {cat.n_records / eh:.4g} records/s")

    if not args.no_cpu:
        cb = run_cpu_baseline(cat, args.cpu_seconds)
        out["cpu_baseline"] = cb
        log(f"cpu baseline ({cb['kind']}, {cb['cores']} threads): {cb['value']:.4g} records/s")

    for d in (d_src, d_hsrc, d_po, d_h, d_b):
        d.free()
    print(json.dumps(out), flush=True)
    if not accuracy["ok"]:
        log(f"FAILED: pe differs from FP64 of the same trees by {accuracy['pe_vs_fp64_same_trees']['max']:.3g} "
            f"> {ACCURACY_BAR:g}")
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
