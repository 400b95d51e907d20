"""Host logic of the trajectory store's sizing (no device needed): the page
count of rays taking runs of pages, the constants shared with the C header,
and the first-call step estimate."""

import re

import numpy as np
import pytest
import torch

from blackhorizon_amd import adjoint
from blackhorizon_amd.adjoint import Recorded, TraceSettings

HEADER = "include/bh_trace.h"


def _runs(steps):
    """Pages one ray requests: runs of 1, 2, 4, ... 16 pages of 32 steps until
    they cover its steps (the record kernels' rule, step by step)."""
    need, got, run = -(-int(steps) // 32), 0, 1
    while got < need:
        got += run
        run = min(2 * run, 16)
    return got


def test_constants_match_the_header(request):
    text = open(request.config.rootpath / HEADER).read()
    page = int(re.search(r"#define BH_TRAJ_PAGE_STEPS (\d+)", text).group(1))
    run = int(re.search(r"#define BH_TRAJ_RUN_MAX (\d+)", text).group(1))
    assert (page, run) == (adjoint.TRAJ_PAGE_STEPS, adjoint.TRAJ_RUN_MAX)


@pytest.mark.parametrize("steps", [0, 1, 31, 32, 33, 64, 65, 96, 97, 224, 225, 480, 481, 992,
                                   993, 1504, 1505, 5000, 76023])
def test_pages_of_one_ray(steps):
    assert Recorded._pages(np.array([steps]), 1) == max(_runs(steps), 1)
    assert Recorded._pages(torch.tensor([steps]), 1) == max(_runs(steps), 1)


def test_pages_of_many_rays_and_of_one_estimate():
    rng = np.random.default_rng(4)
    steps = rng.integers(0, 30000, 2000)
    want = sum(_runs(s) for s in steps)
    assert Recorded._pages(steps, 2000) == want
    assert Recorded._pages(torch.from_numpy(steps), 2000) == want
    assert Recorded._pages(steps.astype(np.float64), 2000) == want
    # one expected count for all rays: n times its runs
    assert Recorded._pages(261.0, 7) == 7 * _runs(261)
    assert Recorded._pages(np.zeros(5), 5) == 1  # at least one page in all
    assert Recorded._pages(np.zeros(0), 0) == 1


def test_first_call_estimates():
    cfg = TraceSettings(escape_radius=45.0, rtol=1e-8, max_step=2.0)
    est = Recorded._expected_steps(None, torch.zeros((3, 8), dtype=torch.float64), cfg)
    assert np.ndim(est) == 0
    assert est == pytest.approx(45.0 / 2.0 + 3.0 * 1e-8 ** -0.2)
    capped = TraceSettings(escape_radius=45.0, rtol=1e-8, max_steps=50)
    assert Recorded._expected_steps(None, torch.zeros((3, 8)), capped) == 50.0
    # rays to an affine end: one estimate per ray, from its initial radius
    s0 = torch.zeros((3, 8), dtype=torch.float64)
    s0[:, 1] = torch.tensor([6.0, 30.0, 100.0])
    orb = TraceSettings(escape_radius=1000.0, rtol=1e-10, max_step=2.0, lambda_end=1e4,
                        max_steps=10_000_000)
    est = Recorded._expected_steps(None, s0, orb)
    want = [1e4 / min(2.0, r ** 1.5 * 1e-10 ** 0.2) for r in (6.0, 30.0, 100.0)]
    np.testing.assert_allclose(est, want)


def test_store_budget_rule(monkeypatch):
    class Props:
        total_memory = 288 * 10 ** 9

    monkeypatch.setattr(torch.cuda, "get_device_properties", lambda dev: Props())
    per_page = adjoint.TRAJ_PAGE_STEPS * adjoint.TRAJ_VALUES * 8 + 8
    limit = min(adjoint.TRAJ_BUDGET_BYTES, Props.total_memory // 3)
    assert Recorded._store_fits(limit // per_page, "cuda")
    assert not Recorded._store_fits(limit // per_page + 1, "cuda")


def test_pool_hints_keep_the_most_recent(monkeypatch):
    monkeypatch.setattr(adjoint, "_POOL_HINTS", {})
    for k in range(adjoint._POOL_HINTS_MAX + 10):
        adjoint._remember_pages(("cfg", k), k)
    adjoint._remember_pages(("cfg", 10), 99)  # stored again: now the newest
    adjoint._remember_pages(("cfg", "new"), 1)
    assert len(adjoint._POOL_HINTS) == adjoint._POOL_HINTS_MAX
    assert ("cfg", 10) in adjoint._POOL_HINTS and adjoint._POOL_HINTS[("cfg", 10)] == 99
    assert ("cfg", 11) not in adjoint._POOL_HINTS  # the oldest left
    assert ("cfg", adjoint._POOL_HINTS_MAX + 9) in adjoint._POOL_HINTS
