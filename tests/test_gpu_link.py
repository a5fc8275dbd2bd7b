"""The drop-in as a maintainer would link it: the reference's own caller
(src/groupies.c with src/properties.c; the bound-pass steps of
src/properties.c:537-550 sequenced by oracle/ref_props.c) built without
potential.o and linked against libhalopot.so (oracle/_ref/
libref_props_halopot.so, oracle/Makefile), so that the reference's strip,
qsort, calculate_corevel and the kinetic-energy centre call the GPU
compute_potential and compute_kinetic_energy.  Compared call by call with the
same harness built with the reference's own potential.o (all CPU):

  * kept counts, n_core and the record order (ties in r2 included):
    identical -- the GPU compute_potential permutes the records as the
    reference's does;
  * corevel, bulkvel and ke: identical bits (the same records in the same
    order reach the same reference code, and the GPU kinetic energy is the
    reference's expression);
  * pe: within 1e-5 relative (FP32 pair terms, compensated sums).
"""
import pathlib

import numpy as np
import pytest

import halopot
from halopot import catalog
import oracle

pytestmark = pytest.mark.gpu

LINKED = pathlib.Path(oracle.__file__).resolve().parent / "_ref" / "libref_props_halopot.so"


@pytest.fixture(scope="module")
def linked():
    if not LINKED.exists():
        pytest.skip("oracle/_ref/libref_props_halopot.so not built (reference sources absent)")
    return oracle.ReferenceProps(LINKED)


@pytest.mark.parametrize("which", ["c2", "potential_only"])
def test_reference_caller_on_libhalopot(linked, which):
    if not oracle.ReferenceProps.available():
        pytest.skip("compiled reference (oracle/_ref) not present")
    cpu = oracle.ReferenceProps()
    if which == "c2":
        cat = catalog.config2()
        calls = range(cat.n_halos)
    else:
        cat = catalog.small_batch(seed=61, n_halos=30, lo=12, hi=6000, potential_only=0.2,
                                  sub_fraction=0.5)
        calls = range(cat.n_halos)
    prm = cat.params
    # the globals the GPU entry points read, wherever the loader bound them
    halopot.set_config(prm)
    n_ties = 0
    for h in calls:
        f, c = int(cat.halos["first"][h]), int(cat.halos["count"][h])
        recs = cat.po[f:f + c].copy()
        recs["flags"] = np.arange(c, dtype=np.int32)
        if which == "potential_only":
            recs["r2"] = np.round(recs["r2"], 2).astype(np.float32)  # many ties
        a, b = recs.copy(), recs.copy()
        ka, ha = linked.bound_pass(a, cat.halos[h], prm)
        kb, hb = cpu.bound_pass(b, cat.halos[h], prm)
        assert ka == kb, h
        if np.unique(b["r2"][:kb]).size < kb:
            n_ties += 1
        np.testing.assert_array_equal(a["flags"][:ka], b["flags"][:kb], err_msg=f"call {h}")
        for fld in ("n_core", "corevel", "bulkvel", "min_vel_err", "min_bulkvel_err", "pos"):
            np.testing.assert_array_equal(ha[fld], hb[fld], err_msg=f"call {h} {fld}")
        np.testing.assert_array_equal(a["ke"][:ka], b["ke"][:kb], err_msg=f"call {h}")
        if ka > 1:
            r = np.abs(a["pe"][:ka] - b["pe"][:kb]) / np.abs(b["pe"][:kb])
            assert r.max() < 1e-5, (h, r.max())
    assert n_ties > 0
