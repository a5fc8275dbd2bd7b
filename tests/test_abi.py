"""The C ABI of libhalopot, checked without a GPU: the library loads, exports
every function and global that include/halopot.h declares (and nothing of
the oracle), its records have the reference's layout, and with no device it
fails loudly instead of computing on the CPU."""
import ctypes as C
import pathlib
import re
import subprocess

import numpy as np
import pytest

import halopot
from halopot import catalog

ROOT = pathlib.Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "halopot.h"


def _declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = "\n".join(l for l in text.splitlines() if not l.lstrip().startswith("typedef"))
    funcs = set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w ]*?[\s*]+(\w+)\s*\(", text, flags=re.M))
    funcs -= {"if", "while", "for", "return", "sizeof"}
    globs = set()
    for m in re.finditer(r"extern\s+double\s+([^;]+);", text):
        globs |= {g.strip() for g in m.group(1).split(",")}
    return funcs, globs


@pytest.fixture(scope="module")
def lib():
    if not halopot.LIB_PATH.exists():
        from halopot import build

        build.build()
    return halopot.lib()


def test_header_declares_the_reference_signatures():
    funcs, globs = _declared()
    assert {"compute_potential", "compute_kinetic_energy"} <= funcs
    assert {"hp_plan_create", "hp_plan_run", "hp_plan_potential", "hp_plan_destroy",
            "hp_plan_stats", "hp_bound_pass_host", "hp_debug_tree", "hp_last_error",
            "hp_set_error_handler"} <= funcs
    assert {"FORCE_RES", "PARTICLE_MASS", "Om", "Ol", "W0", "WA", "SCALE_NOW"} <= globs


def test_library_exports_every_declared_symbol(lib):
    funcs, globs = _declared()
    out = subprocess.run(["nm", "-D", "--defined-only", str(halopot.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = (funcs | globs) - exported
    assert not missing, missing
    for f in funcs:
        assert hasattr(lib, f)
    # only the C ABI is visible: no kernel or helper symbols, nothing of the oracle
    public = {s for s in exported if not s.startswith("_")}
    assert public <= funcs | globs, public - funcs - globs
    assert not any(s.startswith("orc_") for s in exported)


def test_product_does_not_use_the_oracle():
    pkg = ROOT / "mpi-rockstar_amd"
    for f in list(pkg.rglob("*.py")) + list(pkg.rglob("*.hip")) + list(pkg.rglob("*.h")):
        text = f.read_text()
        assert "import oracle" not in text and "liboracle" not in text, f
        assert "oracle/" not in text, f


def test_record_layouts():
    assert catalog.POTENTIAL_DTYPE.itemsize == 48  # src/potential.h:7-21
    assert catalog.POTENTIAL_DTYPE.fields["pe"][1] == 32
    assert catalog.HALO_DTYPE.itemsize == 104
    assert C.sizeof(halopot.Options) == 32
    assert C.sizeof(halopot.Stats) == 6 * 4 + 11 * 8
    assert halopot.NODE_OUT_DTYPE.itemsize == 72


def test_globals_read_and_written(lib):
    old = halopot.get_config()
    try:
        halopot.set_config(catalog.Params(force_res=0.004, particle_mass=3e9, scale_now=0.5))
        cfg = halopot.get_config()
        assert cfg["FORCE_RES"] == 0.004 and cfg["PARTICLE_MASS"] == 3e9
        assert cfg["SCALE_NOW"] == 0.5
    finally:
        for k, v in old.items():
            C.c_double.in_dll(lib, k).value = v


def _has_gpu():
    try:
        return halopot.hipmem.device_count() > 0
    except Exception:
        return False


def test_no_device_fails_loudly(lib):
    if _has_gpu():
        pytest.skip("a GPU is present")
    halos = np.zeros(1, catalog.HALO_DTYPE)
    halos["count"] = 4
    with pytest.raises(halopot.HalopotError, match="no HIP device|gfx950"):
        halopot.Plan(halos)


@pytest.mark.gpu
def test_bad_plans_are_rejected(lib):
    halos = np.zeros(2, catalog.HALO_DTYPE)
    halos["first"] = [0, 5]
    halos["count"] = [4, 4]  # gap between the calls
    with pytest.raises(halopot.HalopotError, match="tile the record array"):
        halopot.Plan(halos)
