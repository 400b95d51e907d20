"""Host logic of the batched tracer's queue order (no device needed)."""

import numpy as np

from blackhorizon_amd import batch
from blackhorizon_amd.kerr import KerrSpacetime


def _job(escape, rtol, max_step=2.0):
    return batch.TraceJob(KerrSpacetime(spin=0.5), None, escape, rtol=rtol, max_step=max_step)


def test_longest_first_ranks_jobs_and_keeps_rays_in_order():
    jobs = [_job(40.0, 1e-6), _job(2000.0, 1e-6), _job(40.0, 1e-10), _job(40.0, 1e-6),
            _job(2000.0, 1e-6, max_step=20.0)]
    counts = [3, 2, 4, 1, 2]
    offsets = np.concatenate([[0], np.cumsum(counts)])
    order = batch.longest_first(jobs, offsets)
    assert sorted(order.tolist()) == list(range(offsets[-1]))
    # estimates: 20 + 95, 1000 + 95, 20 + 600, 20 + 95, 100 + 95
    # -> job 1, job 2, job 4, then the tied jobs 0 and 3 in the given order
    want = [3, 4] + [5, 6, 7, 8] + [10, 11] + [0, 1, 2] + [9]
    assert order.tolist() == want


def test_longest_first_empty_and_single():
    assert batch.longest_first([], np.array([0])).shape == (0,)
    one = batch.longest_first([_job(50.0, 1e-8)], np.array([0, 5]))
    assert one.tolist() == [0, 1, 2, 3, 4]
