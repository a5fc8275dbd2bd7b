"""The error conventions of the C ABI, checked without any GPU work.

The reference's fatal errors go exit() -> rockstar_abort_impl(code, file,
line) -> the installed handler (src/error.h:11-20, src/error.cpp:14-22).
libhalopot's void entry points (compute_potential, compute_kinetic_energy)
report the same way: rockstar_abort_impl when the program links it, else
the handler given to hp_set_error_handler.  The int entry points return an
HP_E* code and leave the message in hp_last_error.  Every argument that
would overflow the library's 32-bit node ids is rejected before any device
work, so these run on a host without a GPU too."""
import numpy as np
import pytest

import halopot
from halopot import hipmem
from halopot import catalog

HP_EINVAL, HP_ENODEV = -1, -2
MAX_RECORDS = (2**31 - 1 - (1 << 22) - 1024) // 5   # kMaxRecords in hp_lib.hip


@pytest.fixture
def handler():
    calls = []
    halopot.set_error_handler(lambda code, f, line: calls.append((code, f, line)))
    yield calls
    halopot.set_error_handler(None)


def test_handler_gets_code_file_line_from_compute_potential(handler):
    po = np.zeros(4, catalog.POTENTIAL_DTYPE)
    # a record count beyond the node-id range: rejected before po is read
    halopot.lib().compute_potential(po.ctypes.data, MAX_RECORDS + 1)
    assert len(handler) == 1
    code, f, line = handler[0]
    assert code == HP_EINVAL and f.endswith("hp_lib.hip") and line > 0
    assert "too large" in halopot.last_error()


def test_handler_gets_code_from_compute_kinetic_energy(handler):
    po = np.zeros(4, catalog.POTENTIAL_DTYPE)
    v = np.zeros(3, np.float32)
    halopot.lib().compute_kinetic_energy(po.ctypes.data, MAX_RECORDS + 1, v.ctypes.data,
                                         v.ctypes.data, 1.0)
    assert [c for c, _, _ in handler] == [HP_EINVAL]
    assert "compute_kinetic_energy" in halopot.last_error()


def test_handler_without_device(handler):
    if hipmem.device_count() > 0:
        pytest.skip("a GPU is present: the no-device route cannot be taken")
    po = np.zeros(5, catalog.POTENTIAL_DTYPE)
    halopot.compute_potential(po)
    assert [c for c, _, _ in handler] == [HP_ENODEV]
    assert "no HIP device" in halopot.last_error()
    assert np.all(po["pe"] == 0)  # nothing was computed on the CPU


def test_oversized_batch_rejected_before_the_device():
    halos = np.zeros(1, catalog.HALO_DTYPE)
    halos["count"] = MAX_RECORDS + 1
    with pytest.raises(halopot.HalopotError, match="too large"):
        halopot.Plan(halos)
    halos["count"] = 10
    halos2 = np.zeros(2, catalog.HALO_DTYPE)
    halos2["count"] = [MAX_RECORDS, 1]
    halos2["first"] = [0, MAX_RECORDS]
    with pytest.raises(halopot.HalopotError, match="too large"):
        halopot.Plan(halos2)


def test_untiled_calls_rejected_before_the_device():
    halos = np.zeros(2, catalog.HALO_DTYPE)
    halos["first"] = [0, 5]
    halos["count"] = [4, 4]
    with pytest.raises(halopot.HalopotError, match="tile the record array"):
        halopot.Plan(halos)
