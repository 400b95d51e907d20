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


def test_longest_first_ranks_orbit_rays_innermost_first():
    """Rays integrated to an affine end are ranked one by one by their
    expected steps lambda_end / min(max_step, r0^1.5 rtol^0.2), across jobs;
    ties keep the given order; a job traced to escape keeps its job-level
    estimate."""
    import torch

    st = KerrSpacetime(spin=0.9)

    def orbit_job(radii, as_tensor=False):
        s = np.zeros((len(radii), 8))
        s[:, 1] = radii
        if as_tensor:
            s = torch.from_numpy(s)
        return batch.TraceJob(st, s, 1000.0, rtol=1e-10, lambda_end=1.0e4)

    jobs = [orbit_job([30.0, 6.0, 40.0]), orbit_job([6.5, 50.0], as_tensor=True),
            _job(40.0, 1e-6)]
    offsets = np.array([0, 3, 5, 6])
    est = batch.expected_steps(jobs[0])
    # r = 30: h = 30^1.5 1e-2 = 1.64; r = 6: 0.147; r = 40: capped at 2
    np.testing.assert_allclose(est, [1.0e4 / (30.0 ** 1.5 * 1e-2), 1.0e4 / (6.0 ** 1.5 * 1e-2),
                                     5000.0])
    assert batch.expected_steps(jobs[2]) == 20.0 + 6.0 * 1e-6 ** -0.2
    # 6 (68k), 6.5 (60k), 30 (6.1k), then the ties 40 and 50 (5000), photons
    assert batch.longest_first(jobs, offsets).tolist() == [1, 3, 0, 2, 4, 5]
