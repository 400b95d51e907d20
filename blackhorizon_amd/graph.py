"""HIP-graph capture of trace launches.

A plan's ``launch()`` (``tracer.TracePlan``, ``batch.BatchPlan``) only
enqueues kernels on the current stream -- the library's zeroing is done by
small kernels, never memset nodes, and nothing allocates or synchronises --
so the whole launch sequence can be recorded once as a HIP graph and
replayed on the same device buffers with a single graph launch.  The
reference has no counterpart (CuPy issues ~800 launches per controller
iteration, tracer.py:104-120); replay removes the per-launch host cost that
dominates plans made of many small traces.
"""

from __future__ import annotations

import torch

from . import _lib


class TraceGraph:
    """``plan.launch()`` captured in a HIP graph.

    ``replay()`` re-runs the captured launches on the current stream; the
    plan's outputs then hold that replay's results (read them with
    ``plan.result()`` / ``plan.results()`` as after an eager launch).
    Inputs may be rewritten in place between replays (same buffers).
    """

    def __init__(self, plan, warmup: bool = True):
        _lib.require_device()
        self.plan = plan
        cur = torch.cuda.current_stream()
        if warmup:
            # first launches initialise lazily loaded code objects outside
            # the capture
            plan.launch()
            torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(cur)
        with torch.cuda.graph(self.graph, stream=side):
            plan.launch()
        cur.wait_stream(side)

    def replay(self) -> None:
        """Enqueue one replay of the captured launches."""
        self.graph.replay()
