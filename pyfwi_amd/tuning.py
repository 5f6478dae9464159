"""Execution-shape autotuner for the batched gradient.

A plan can run a call's shots in passes of G shots, do each forward or
time-reversed step as one fused launch or two in-place launches, run the
kept-state forward two steps a launch, do each reference-adjoint step as one
launch or two, and fix the strip height of the column-marching kernels
(``pyfwi_plan_set_tuning``).  The best choice depends
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

__all__ = ["default_candidates", "truncate", "autotune", "autotune_gradient"]


def default_candidates(nshots, max_pass=None, fused=(0, _native.TUNE_FUSED_STEPS),
                       rows=(12, 24)):
    """(shots per pass, fused bits, strip rows) triples; 0 rows = automatic.
    `fused`: the bit sets of ``Plan.set_tuning`` to try (TUNE_FUSED_STEPS,
    TUNE_FWD_TWO, TUNE_ADJ_FUSED; True stands for TUNE_FUSED_STEPS); the fixed strip
    heights `rows` are tried with the largest pass and each bit set."""
    top = nshots if max_pass is None else max(1, min(nshots, max_pass))
    passes = sorted({g for g in (top, 10, 8, 5, 4, 2) if 1 <= g <= top}, reverse=True)
    bits = [int(f) for f in fused]
    out = [(g, f, 0) for g in passes for f in bits]
    out += [(passes[0], f, r) for f in bits for r in rows]
    return out


def truncate(shots, nt):
    """Copies of the shot table entries limited to `nt` time steps."""
    out = []
    for sh in shots:
        s = copy.copy(sh)
        s.nt = min(sh.nt, nt)
        s.n_samples = min(sh.n_samples, -(-s.nt // max(1, sh.dt_scale)))
        out.append(s)
    return out


def autotune(plan, run, synchronize, candidates, log=None):
    """Time ``run()`` (a truncated workload on the caller's buffers) for each
    candidate tuning and configure ``plan`` with the fastest.  Returns
    (best, {candidate: seconds})."""
    times = {}
    for cand in candidates:
        plan.set_tuning(*cand)
        best = float("inf")
        for rep in range(2):  # the first run warms up (allocation, caches)
            synchronize()
            t0 = time.perf_counter()
            run()
            synchronize()
            if rep:
                best = min(best, time.perf_counter() - t0)
        times[cand] = best
        if log:
            log(f"tune: shots/pass={cand[0]} fused={int(cand[1])} rows={cand[2]}: "
                f"{best * 1e3:.1f} ms")
    best = min(times, key=times.get)
    plan.set_tuning(*best)
    return best, times


def autotune_gradient(plan, shots, src, seis, res, grad, synchronize, candidates=None,
                      nt_trial=200, log=None):
    """Autotune the forward + reference gradient pair (time reversal between
    the checkpoints of chpr = 0)."""
    trial = truncate(shots, nt_trial)
    chp = checkpoint_steps(max(s.nt for s in trial), 0)

    def run():
        plan.forward(trial, src, seis, chp, 0, _native.MEM_DEVICE)
        plan.gradient(trial, src, res, grad, _native.MODE_PYFWI, 0, 0, _native.MEM_DEVICE)

    return autotune(plan, run, synchronize, candidates or default_candidates(len(shots)), log)
