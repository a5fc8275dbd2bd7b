"""The closed form the GPU build, the non-finite removal and the bound pass's
potential-only strip use for the reference's in-place loops
(src/fast3tree.c:762-773, 875-887; src/properties.c:539-544): the library's
own hp::inplace_dest, compiled for the host, against the swap loop on
300,000 random partitions of up to 70 points.  No GPU code runs."""
import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent


def test_inplace_dest_matches_the_swap_loop(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not pathlib.Path(hipcc).exists():
        pytest.skip("hipcc not available")
    exe = tmp_path / "inplace_dest_check"
    subprocess.run([hipcc, "-O2", "-std=c++17", "--offload-arch=gfx950",
                    f"-I{ROOT / 'mpi-rockstar_amd' / 'csrc'}", f"-I{ROOT / 'include'}",
                    str(ROOT / "tests" / "native" / "inplace_dest_check.hip"), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "match the swap loop" in out.stdout
