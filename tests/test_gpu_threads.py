"""compute_potential from many host threads at once, as the reference's
OpenMP loop over halos calls it (src/mpi_main.cpp:1369-1402): every thread
has its own stream and device scratch, so concurrent calls must give the
same bits as the same calls made one after another."""
import threading

import numpy as np
import pytest

import halopot
from halopot import catalog
from conftest import with_ids

pytestmark = pytest.mark.gpu


def test_concurrent_calls_match_serial():
    cat = catalog.small_batch(seed=77, n_halos=48, lo=5, hi=4000)
    halopot.set_config(cat.params)
    segs = [with_ids(cat.segment(h)) for h in range(cat.n_halos)]
    want = []
    for s in segs:
        g = s.copy()
        halopot.compute_potential(g)
        want.append(g["pe"].copy())

    got = [None] * len(segs)
    errors = []

    def worker(t, nthreads):
        try:
            for rep in range(3):
                for h in range(t, len(segs), nthreads):
                    g = segs[h].copy()
                    halopot.compute_potential(g)
                    if rep == 0:
                        got[h] = g["pe"].copy()
                    elif not np.array_equal(g["pe"], got[h]):
                        errors.append(f"thread {t} call {h} repeat {rep} differs")
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(f"thread {t}: {e!r}")

    threads = [threading.Thread(target=worker, args=(t, 8)) for t in range(8)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not any(th.is_alive() for th in threads), "a worker did not finish"
    assert not errors, errors
    for h in range(len(segs)):
        np.testing.assert_array_equal(got[h], want[h], err_msg=f"call {h}")
