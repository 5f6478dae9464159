"""Execution-shape autotuner for the batched gradient.

A plan can run a call's shots in passes of G shots, do each time step as one
fused launch or two in-place launches, and fix the strip height of the
column-marching kernels (``pyfwi_plan_set_tuning``).  The best choice depends
on the grid, the number of shots and how much of the working set the 256 MB
on-package cache holds, so it is measured: each candidate runs a forward and
a reference-mode gradient of the actual shots, truncated to ``nt_trial``
steps, on the caller's device buffers, and the fastest is kept.  Results do
not depend on the choice.
"""
import copy
import time

from . import _native
from .wave_propagation import checkpoint_steps

__all__ = ["default_candidates", "autotune_gradient"]


def default_candidates(nshots):
    """(shots per pass, fused, strip rows) triples; 0 rows = automatic."""
    passes = sorted({g for g in (nshots, 10, 8, 5, 4, 2) if 1 <= g <= nshots}, reverse=True)
    out = []
    for g in passes:
        for fused in (False, True):
            out.append((g, fused, 0))
    out += [(passes[0], False, 16), (passes[0], False, 48)]
    return out


def _truncate(shots, nt):
    out = []
    for sh in shots:
        s = copy.copy(sh)
        s.nt = min(sh.nt, nt)
        s.n_samples = min(sh.n_samples, -(-s.nt // max(1, sh.dt_scale)))
        out.append(s)
    return out


def autotune_gradient(plan, shots, src, seis, res, grad, synchronize, candidates=None,
                      nt_trial=200, log=None):
    """Time each candidate on a truncated forward + reference gradient and
    configure ``plan`` with the fastest.  Device buffers as for
    pyfwi_forward / pyfwi_gradient (MEM_DEVICE); ``synchronize()`` waits for
    the device.  Returns (best, {candidate: seconds})."""
    trial = _truncate(shots, nt_trial)
    nt = max(s.nt for s in trial)
    chp = checkpoint_steps(nt, 0)
    candidates = candidates or default_candidates(len(shots))
    times = {}
    for cand in candidates:
        g, fused, rows = cand
        plan.set_tuning(g, fused, rows)
        best = float("inf")
        for rep in range(2):  # first run warms up (allocation, caches)
            synchronize()
            t0 = time.perf_counter()
            plan.forward(trial, src, seis, chp, 0, _native.MEM_DEVICE)
            plan.gradient(trial, src, res, grad, _native.MODE_PYFWI, 0, 0, _native.MEM_DEVICE)
            synchronize()
            if rep:
                best = min(best, time.perf_counter() - t0)
        times[cand] = best
        if log:
            log(f"tune: shots/pass={g} fused={int(fused)} rows={rows}: {best * 1e3:.1f} ms")
    best = min(times, key=times.get)
    plan.set_tuning(*best)
    return best, times
