"""Writes the golden vectors of tests/golden/ from the compiled reference.

The reference (src/potential.c with src/fast3tree.c, src/hubble.c,
src/config_vars.c, src/error.cpp) is built by oracle/Makefile into
oracle/_ref/libref_potential.so from its sources under /root/reference; this
script runs its compute_potential and compute_kinetic_energy on seeded
synthetic inputs and stores inputs and outputs as .npz (no pickles):

  c1_reference.npz     config 1 (one NFW halo + background, 11,000 records):
                       pe of every record, ke about a given centre
  micro_reference.npz  small and degenerate calls (N = 1..24, coincident
                       points, non-finite positions, two particles)
  batch_reference.npz  a 30-call batch of small NFW halos, pe per record

pe is stored in input order (the reference permutes the records; the flags
field carries the input index through the call).

Run:  python tests/golden/make_golden.py   (needs oracle/_ref)
"""
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT / "mpi-rockstar_amd"))
sys.path.insert(0, str(ROOT / "oracle"))

from halopot import catalog  # noqa: E402
import oracle  # noqa: E402


def ref_pe(R, po, prm):
    q = po.copy()
    q["flags"] = np.arange(q.shape[0], dtype=np.int32)
    R.compute_potential(q, prm)
    pe = np.empty(q.shape[0])
    pe[q["flags"]] = q["pe"]
    return pe


def micro_cases():
    rng = np.random.Generator(np.random.Philox(0x5EED00AA))
    cases = []
    for n in range(1, 25):  # across the leaf size (10) and the degenerate limit (20)
        p = np.zeros(n, catalog.POTENTIAL_DTYPE)
        p["pos"][:, :3] = rng.uniform(10.0, 10.5, (n, 3))
        cases.append(p)
    two = np.zeros(2, catalog.POTENTIAL_DTYPE)
    two["pos"][0, :3] = [5.0, 5.0, 5.0]
    two["pos"][1, :3] = [5.0, 5.0, 5.002]  # closer than FORCE_RES
    cases.append(two)
    for n in (12, 20, 21, 35):  # coincident points: pairs at FORCE_RES, or none
        p = np.zeros(n, catalog.POTENTIAL_DTYPE)
        p["pos"][:, :3] = 7.25
        cases.append(p)
    mix = np.zeros(40, catalog.POTENTIAL_DTYPE)  # a coincident clump inside a cloud
    mix["pos"][:, :3] = rng.uniform(0.0, 1.0, (40, 3))
    mix["pos"][5:30, :3] = 0.5
    cases.append(mix)
    nf = np.zeros(60, catalog.POTENTIAL_DTYPE)  # non-finite positions are dropped
    nf["pos"][:, :3] = rng.uniform(0.0, 1.0, (60, 3))
    nf["pos"][7, 0] = np.nan
    nf["pos"][31, 2] = np.inf
    nf["pos"][44, 1] = -np.inf
    cases.append(nf)
    return cases


def main():
    oracle.build()
    R = oracle.Reference()

    cat = catalog.config1()
    prm = cat.params
    pe = ref_pe(R, cat.po, prm)
    vc = np.array([110.0, -75.0, 30.0], np.float32)
    pc = cat.halos["pos"][0, :3].astype(np.float32)
    k = cat.po.copy()
    k["ke"][::9] = -1.0
    R.compute_kinetic_energy(k, vc, pc, 1.0, prm)
    np.savez_compressed(HERE / "c1_reference.npz", pos=cat.po["pos"], r2=cat.po["r2"], pe=pe,
                        ke_in=np.where(np.arange(cat.n_records) % 9 == 0, -1.0, 0.0).astype(np.float32),
                        ke=k["ke"], vel_cen=vc, pos_cen=pc,
                        params=np.array([prm.force_res, prm.particle_mass, prm.Om, prm.Ol,
                                         prm.W0, prm.WA, prm.scale_now]))

    mprm = catalog.Params(particle_mass=1e9)
    cases = micro_cases()
    pos = np.concatenate([c["pos"] for c in cases])
    off = np.cumsum([0] + [c.shape[0] for c in cases])
    mpe = np.concatenate([ref_pe(R, c, mprm) for c in cases])
    np.savez_compressed(HERE / "micro_reference.npz", pos=pos, offsets=off, pe=mpe,
                        params=np.array([mprm.force_res, mprm.particle_mass]))

    b = catalog.small_batch(seed=2024, n_halos=30, lo=11, hi=2500, contam=0.05)
    bpe = np.concatenate([ref_pe(R, b.segment(h), b.params) for h in range(b.n_halos)])
    np.savez_compressed(HERE / "batch_reference.npz", pos=b.po["pos"], first=b.halos["first"],
                        count=b.halos["count"], pe=bpe,
                        params=np.array([b.params.force_res, b.params.particle_mass]))
    for f in sorted(HERE.glob("*.npz")):
        print(f.name, f.stat().st_size)


if __name__ == "__main__":
    main()
