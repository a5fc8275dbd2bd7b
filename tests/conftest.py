"""Shared test setup: import paths, the gpu marker, oracle fixtures."""
import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
for p in (ROOT / "mpi-rockstar_amd", ROOT / "oracle", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs through libhalopot)")


@pytest.fixture(scope="session")
def restatement():
    import oracle

    return oracle.Restatement()


@pytest.fixture(scope="session")
def reference():
    import oracle

    if not oracle.Reference.available():
        pytest.skip("compiled reference (oracle/_ref) not present")
    return oracle.Reference()


def with_ids(po):
    """Copy of po whose flags field holds the record index (compute_potential
    leaves flags alone and the restatement permutes records)."""
    q = po.copy()
    q["flags"] = np.arange(q.shape[0], dtype=np.int32)
    return q


def by_id(po, field="pe"):
    out = np.empty(po.shape[0], dtype=po[field].dtype)
    out[po["flags"]] = po[field]
    return out
