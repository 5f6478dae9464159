"""Benchmark: fwd+adj gradient Gcell-updates/s on one MI355X, batched vs per-shot.

Workload (default): uneven_1000x2000_64shots -- nz=1000, nx=2000, dh=4 m
(1040 x 2040 cells with the 20-cell absorbing layer), 64 surface shots, each
with its own source position, nt (1564..3987), receiver aperture (201..1591)
and Ricker frequency (10..25 Hz), order-4 stencil, FP32.  One step = one full
gradient of the L2 misfit, as pyfwi_misfit_gradient computes it: the shots
packed into passes whose kept states fill the free HBM, batched forward
modelling with every forward state kept (the reference's chpr = 100),
residual and misfit against observed data on the device, and the
reference-semantics adjoint (adjoint sweeps + imaging condition) of each pass.
`--workload plain_medium` runs the 500 x 1000-cell, 20-shot case instead;
`--mode reverse` keeps no states and rebuilds them by time reversal
(chpr = 0).  Inputs are resident in HBM before the timed region.

A gradient cell-update is one (cell incl. absorbing layer, time step, shot)
taken through forward plus adjoint; value = shots * cells * nt / seconds.
Forward modelling alone is reported beside it.

Prints ONE JSON line.  --gpus above 1 is rejected: this build runs on one GPU.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from pyfwi_amd import _native  # noqa: E402
from pyfwi_amd.wave_propagation import checkpoint_steps  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of one launch (FP32 words read + written once per cell):
# (bytes per cell of the whole grid and shot, extra bytes per cell of the
# interior where the imaging condition runs and shot, bytes per imaging cell
# and shot group).  The gradient accumulators are shared by the shots of a
# block, which add their terms up first (shots_per_block); the exact adjoint
# images every cell, the reference adjoint the interior.
KERNEL_BYTES = {
    "fwd_vel": (32, 0, 0),       # R txx tzz txz vx vz b, W vx vz
    "fwd_str": (40, 0, 0),       # R vx vz txx tzz txz lam mu, W txx tzz txz
    "rev_str": (40, 0, 0),
    "rev_vel": (32, 0, 0),
    "adj_str": (36, 8, 16),      # all: R avx avz b 3 atau, W 3 atau; interior: R fwd vx vz; group: RW g_mu g_lam
    "adj_vel_grad": (36, 16, 8),  # all: R 3 atau lam mu 2 av, W 2 av; interior: R fwd 3 tau, b; group: RW g_rho
    "exact_str": (44, 0, 16),     # R 3 atau lam mu 2 av 2 fwd v, W 2 av; group: RW g_mu g_lam
    "exact_vel": (56, 0, 8),      # R 2 av b 3 fwd tau 3 atau, W 3 atau 2 av; group: RW g_b
    "adj_step": (52, 20, 24),     # all: R 5 adjoint lam mu b, W 5 adjoint; interior: R 5 fwd; group: RW 6 acc
    "fwd_step": (52, 0, 0),       # R 5 fields lam mu b, W 5 fields
    "fwd_two": (72, 0, 0),        # two steps: R 5 fields lam mu b, W 2 x 5 fields
    "rev_step": (52, 0, 0),
}
def shots_per_block(order, shots):
    """Shots per block of the stencil kernels, as the library picks them
    (march_spb in pyfwi_amd/csrc/pyfwi_kernels.hip)."""
    n = int(round(shots))
    return 5 if order == 4 and n == shots and n % 5 == 0 and n % 4 != 0 else 4


def launch_bytes(kernel, cells, interior, shots, order=4):
    a, b, c = KERNEL_BYTES[kernel]
    groups = math.ceil(shots / shots_per_block(order, shots))
    imaged = cells if kernel.startswith("exact") else interior
    return (a * cells + b * interior) * shots + c * imaged * groups


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_source_sha1():
    """SHA-1 of the kernel sources, which alone set the HBM bytes a launch
    moves (the plan code around them does not)."""
    import hashlib
    files = [os.path.join(ROOT, "pyfwi_amd", "csrc", n)
             for n in ("pyfwi_kernels.hip", "pyfwi_kernels.h")]
    return hashlib.sha1(b"".join(open(f, "rb").read() for f in files)).hexdigest()


def _traffic_records(workload_name):
    """The committed rocprofv3 FETCH_SIZE / WRITE_SIZE summaries
    (profiles/hbm_traffic*.json) of this workload taken with the current
    kernel sources."""
    import glob
    out = []
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "hbm_traffic*.json"))):
        try:
            d = json.load(open(path))
        except (OSError, ValueError):
            continue
        if d.get("workload") == workload_name and \
                d.get("kernel_source_sha1") == kernel_source_sha1():
            out.append(d)
    return out


def measured_traffic(workload_name, kernel, shots=None):
    """HBM bytes per dispatch from the committed counter summary, when it was
    taken with the current kernel sources (and, given `shots`, with
    dispatches of that many shots)."""
    for d in _traffic_records(workload_name):
        if shots is not None and d.get("shots_per_dispatch") != shots:
            continue
        got = d.get("bytes_per_dispatch", {}).get(kernel)
        if got is not None:
            return got
    return None


def build_problem(workload, nt=None, shots=None, order=0):
    from pyfwi_amd import workloads
    from pyfwi_amd.wave_propagation import WavePropagator
    kw = {}
    if nt:
        kw["nt"] = nt
    if shots:
        kw["ns"] = shots
    w = getattr(workloads, workload)(**kw)
    if order:
        w.inpa["sdo"] = order
        w.name += f"_o{order}"
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=w.components,
                       n_well_rec=w.meta.get("n_well_rec", 0), chpr=w.chpr)
    return w, W


def device_setup(w, W, torch):
    """Upload model and tables once; return device buffers and shot tables."""
    from pyfwi_amd import _native
    from pyfwi_amd.fwi_tools import expand_model
    dev = torch.device("cuda:0")
    plan = W._plan
    plan.set_stream(torch.cuda.current_stream().cuda_stream)
    m = {k: expand_model(v, int(W.tnz), int(W.tnx), W.npml) for k, v in w.model.items()}
    W.pml_preparation(m["vp"].max())
    W.elastic_buffers(m)
    ids = list(range(int(W.ns)))
    shots = W._shots(ids)
    src = torch.from_numpy(W._source_table()).to(dev)
    nrec = len(ids) * 5 * int(W.nt_ext) * int(W.nr)
    seis = torch.zeros(nrec, dtype=torch.float32, device=dev)
    obs = torch.zeros(nrec, dtype=torch.float32, device=dev)
    res = torch.zeros(nrec, dtype=torch.float32, device=dev)
    grad = torch.zeros(6 * int(W.tnz) * int(W.tnx), dtype=torch.float32, device=dev)
    # observed data: the same acquisition over a +5% vp anomaly
    from pyfwi_amd.workloads import gaussian_anomaly
    mo = gaussian_anomaly(w.model, w.model_shape[0] / 2, w.model_shape[1] / 2, 40.0, 0.05)
    mo = {k: expand_model(v, int(W.tnz), int(W.tnx), W.npml) for k, v in mo.items()}
    W.elastic_buffers(mo)
    plan.forward(shots, src, obs, W.chp, 0, _native.MEM_DEVICE)
    W.elastic_buffers(m)
    torch.cuda.synchronize()
    from pyfwi_amd.wave_propagation import checkpoint_steps
    return dict(plan=plan, shots=shots, src=src, seis=seis, obs=obs, res=res, grad=grad,
                chp_all=checkpoint_steps(W.nt, 100))


def kept_step(W, b, shots=None):
    """One misfit + gradient with every forward state kept in HBM per pass
    (the reference's chpr = 100), residual on the GPU."""
    from pyfwi_amd import _native
    shots = b["shots"] if shots is None else shots
    return b["plan"].misfit_gradient(shots, b["src"], b["obs"], b["seis"], b["grad"],
                                     b["chp_all"], 0, 0, _native.MEM_DEVICE)


def gradient_step(W, b, shots=None, seis=None, obs=None, res=None):
    """One fwd+adj gradient of the given shots (all by default)."""
    from pyfwi_amd import _native
    plan = b["plan"]
    shots = b["shots"] if shots is None else shots
    seis = b["seis"] if seis is None else seis
    obs = b["obs"] if obs is None else obs
    res = b["res"] if res is None else res
    plan.forward(shots, b["src"], seis, W.chp, 0, _native.MEM_DEVICE)
    res.copy_(seis)
    res.sub_(obs)
    plan.gradient(shots, b["src"], res, b["grad"], _native.MODE_PYFWI, 0, 0, _native.MEM_DEVICE)


def exact_step(W, b, shots=None):
    """One forward and exact-adjoint gradient of the given shots (all by
    default), as the autograd path runs it: the forward keeps checkpoints
    every ceil(sqrt(nt)) steps when all shots' fit in HBM, else the gradient
    recomputes them pass by pass; each segment is recomputed once."""
    from pyfwi_amd import _native
    plan = b["plan"]
    shots = b["shots"] if shots is None else shots
    plan.forward(shots, b["src"], b["seis"], None, _native.CKPT_AUTO, _native.MEM_DEVICE)
    b["res"].copy_(b["seis"])
    b["res"].sub_(b["obs"])
    plan.gradient(shots, b["src"], b["res"], b["grad"], _native.MODE_EXACT, 0, 0,
                  _native.MEM_DEVICE)


def exact_kept_step(W, b, shots=None):
    """One forward keeping every state (the reference's chpr = 100) and the
    exact-adjoint gradient reading those states: no segment is recomputed."""
    from pyfwi_amd import _native
    plan = b["plan"]
    shots = b["shots"] if shots is None else shots
    plan.forward(shots, b["src"], b["seis"], b["chp_all"], 0, _native.MEM_DEVICE)
    b["res"].copy_(b["seis"])
    b["res"].sub_(b["obs"])
    plan.gradient(shots, b["src"], b["res"], b["grad"], _native.MODE_EXACT, 0, 0,
                  _native.MEM_DEVICE)


def per_shot_rate(W, b, torch, nsub, step):
    """The plain per-shot path: the same kernels on one shot at a time (one
    shot per pass, the reference's schedule), on the first `nsub` shots."""
    cells = int(W.tnz) * int(W.tnx)
    plan = b["plan"]
    plan.set_tuning(1, False, 0)
    plan.set_graphs(False)  # the reference's schedule: every launch from the host
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step(W, b, b["shots"][:nsub])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return nsub * cells * W.nt / dt / 1e9, dt


def forward_rate(W, b, torch, tuning):
    """Forward modelling alone (records, no kept states) of all shots in one
    pass with the gradient's kernel shape: Gcell-updates/s."""
    cells = int(W.tnz) * int(W.tnx)
    plan = b["plan"]
    plan.set_tuning(0, tuning[1], tuning[2])
    no_chp = np.zeros(0, np.int32)
    plan.forward(b["shots"], b["src"], b["seis"], no_chp, 0, _native.MEM_DEVICE)  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.forward(b["shots"], b["src"], b["seis"], no_chp, 0, _native.MEM_DEVICE)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    plan.set_tuning(*tuning)
    return int(W.ns) * cells * W.nt / dt / 1e9, dt


def kernel_profile(W, b, torch, step):
    """In-sequence launch durations (total ms, launches) of the stencil kernels
    over one gradient step, from HIP events on the plan's stream.

    A recorded event stalls the stream for several microseconds, part of which
    falls inside a bracket around one launch.  So one run brackets every launch
    only to split each time loop between its kernels, and a second run
    brackets only the forward and backward time loops, inside which the
    kernels run as in an untimed call; each kernel gets its share of its
    loop's time."""
    from pyfwi_amd import _native
    plan = b["plan"]
    plan.set_timing(_native.TIMING_LAUNCHES)
    step(W, b)
    torch.cuda.synchronize()
    per = {}
    for k, name in enumerate(_native.KERNEL_NAMES):
        ms, n = plan.kernel_timing(k) if name else (0.0, 0)
        if n:
            per[name] = (ms, n)
    plan.set_timing(_native.TIMING_LOOPS)
    step(W, b)
    torch.cuda.synchronize()
    loops = {"fwd": plan.kernel_timing(_native.K_FWD_LOOP)[0],
             "bwd": plan.kernel_timing(_native.K_BWD_LOOP)[0]}
    plan.set_timing(0)
    groups = {"fwd": ("fwd_vel", "fwd_str", "fwd_step", "fwd_two", "fwd_sparse"),
              "bwd": ("rev_str", "rev_vel", "rev_step", "adj_str", "adj_vel_grad", "adj_step",
                      "bwd_sparse")}
    if "exact_str" in per:
        # the exact adjoint's backward sweep interleaves the segment
        # recomputation and brackets no loop: bracketed launch times as they are
        return per
    out = {}
    for loop, kinds in groups.items():
        ks = [k for k in kinds if k in per]
        if not ks:
            continue
        # the stall a bracket adds is about the same for every launch: take it
        # out of each kind (per launch), then share the loop's time
        stall = max(0.0, (sum(per[k][0] for k in ks) - loops[loop])
                    / sum(per[k][1] for k in ks))
        net = {k: max(per[k][0] - per[k][1] * stall, 0.0) for k in ks}
        total = sum(net.values()) or 1.0
        for k in ks:
            if not k.endswith("_sparse"):  # the small gather / injection launches
                out[k] = (loops[loop] * net[k] / total, per[k][1])
    return out


def _set_omp_threads(n):
    import ctypes
    try:
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))
        return True
    except OSError:
        return False


def cpu_baseline(w, steps_cap, threads):
    """The CPU oracle (C, OpenMP over rows) on one shot of the same workload,
    forward + reference adjoint, for `steps_cap` time steps."""
    from oracle import oracle as O
    inpa = dict(w.inpa)
    nt = steps_cap
    inpa["t"] = (nt - 1) * inpa["dt"] + inpa["dt"] / 2
    src_loc = np.c_[w.src.i, w.src.j].astype(np.float64) * inpa["dh"]
    st = O.Setup(inpa, src_loc[:1], w.rec_loc, w.model_shape, w.src.w, w.src.component,
                 chpr=0, components=w.components, n_well_rec=w.meta.get("n_well_rec", 0))
    O.lib()
    _set_omp_threads(threads)
    t0 = time.perf_counter()
    seis, states = O.forward(st, w.model, np.float32)
    res = O.section(seis, 4)
    O.raw_gradient(st, w.model, res, states, np.float32)
    dt = time.perf_counter() - t0
    cells = st.tnz * st.tnx
    name, ncpu = cpu_model()
    return {"value": round(cells * nt / dt / 1e9, 5), "unit": "Gcell-updates/s",
            "cores": threads, "kind": "port", "cpu": name, "host_cpus": ncpu,
            "sample": f"1 shot x {nt} steps of the same grid ({st.tnz}x{st.tnx}), forward + "
                      f"reference adjoint, C oracle, {threads} OpenMP thread(s), {dt:.1f} s"}


def copy_calibration(torch, nbytes=1 << 30):
    """A plain device copy of known size, for scaling HBM byte counters."""
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"bench: copy calibration {nbytes} bytes read + {nbytes} bytes written, 3 launches")


def copy_rate(torch, nbytes=1 << 30, reps=5):
    """Bytes moved per second (read + write) by a plain device-to-device copy
    of `nbytes`, timed with events on torch's stream: what a streaming kernel
    reaches on this HBM, to set the stencil kernels' rates against."""
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    rate = 2.0 * nbytes * reps / (e0.elapsed_time(e1) / 1e3)
    del a, b
    torch.cuda.empty_cache()
    return rate


def two_step_pairs(live):
    """Steps that start a two-step forward launch (PYFWI_TUNE_FWD_TWO), as the
    plan pairs them: from t = 0, steps t and t+1 share a launch when the same
    shots are live at both, else step t runs alone."""
    starts, t = [], 0
    while t < len(live):
        if t + 1 < len(live) and live[t + 1] == live[t]:
            starts.append(t)
            t += 2
        else:
            t += 1
    return np.asarray(starts, np.int64)


def launch_live_shots(kind, pass_nts, two=False):
    """Per launch of kernel `kind` over passes of uneven shots, the number of
    shots it advances.  pass_nts: per pass, the shots' nt in the order the
    plan runs them (decreasing).  A forward launch of step t advances the
    shots with t < nt_s, an adjoint launch of step t (nt-1 .. 1) those with
    1 <= t < nt_s: each launch takes only the live shots (the plan's
    time-major kept states).  two: the forward runs two steps a launch where
    it can (two_step_pairs); "fwd_two" counts those launches and the
    forward's one-step kinds the steps left alone."""
    adjoint = kind.startswith("adj") or kind.startswith("exact")
    out = []
    for nts in pass_nts:
        nts = np.sort(np.asarray(nts, np.int64))[::-1]
        t = np.arange(1 if adjoint else 0, int(nts[0]))
        live = (nts[None, :] > t[:, None]).sum(1)
        if two and not adjoint:
            starts = two_step_pairs(live)
            paired = np.zeros(len(t), bool)
            paired[starts] = paired[starts + 1] = True
            live = live[starts] if kind == "fwd_two" else live[~paired]
        elif kind == "fwd_two":
            live = live[:0]
        out.append(live)
    return np.concatenate(out) if out else np.zeros(0, np.int64)


def pass_kernel_bytes(kind, pass_nts, cells, interior, order=4, two=False):
    """Algorithmic bytes and launch count of kernel `kind` over passes of
    uneven shots (launch_live_shots); the accumulators move once per block's
    group with a live shot, shots per block following the live count as the
    library picks them."""
    a, b, c = KERNEL_BYTES[kind]
    live = launch_live_shots(kind, pass_nts, two)
    spb = np.where((order == 4) & (live % 5 == 0) & (live % 4 != 0), 5, 4)
    groups = -(-live // spb)
    imaged = cells if kind.startswith("exact") else interior
    total = float(((a * cells + b * interior) * live).sum() + (c * imaged * groups).sum())
    return total, len(live)


def cpu_model():
    """CPU model name and logical CPU count of this host."""
    name = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                name = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return name, os.cpu_count() or 1


def cpu_baseline_shot(inpa, src_loc, rec_loc, model, wavelet, component, components, steps,
                      threads):
    """The CPU oracle (C, OpenMP over rows) on one shot -- source at
    ``src_loc`` (x, z in metres), receivers ``rec_loc`` -- forward plus the
    reference adjoint (chpr = 0) for ``steps`` time steps."""
    from oracle import oracle as O
    inpa = dict(inpa)
    inpa["t"] = (steps - 1) * inpa["dt"] + inpa["dt"] / 2
    st = O.Setup(inpa, np.asarray([src_loc], np.float64), rec_loc, model["vp"].shape, wavelet,
                 component, chpr=0, components=components)
    O.lib()
    _set_omp_threads(threads)
    t0 = time.perf_counter()
    seis, states = O.forward(st, model, np.float32)
    res = O.section(seis, 4)
    O.raw_gradient(st, model, res, states, np.float32)
    dt = time.perf_counter() - t0
    cells = st.tnz * st.tnx
    name, ncpu = cpu_model()
    return {"value": round(cells * steps / dt / 1e9, 5), "unit": "Gcell-updates/s",
            "cores": threads, "kind": "port", "cpu": name, "host_cpus": ncpu,
            "sample": f"1 shot x {steps} steps of the same grid ({st.tnz}x{st.tnx}), forward + "
                      f"reference adjoint (chpr=0), C oracle, {threads} OpenMP thread(s), "
                      f"{dt:.1f} s"}


def bench_uneven(args, torch):
    """Uneven shot batch (uneven_batch: 1000x2000 cells, 64 shots with their
    own nt, aperture and Ricker frequency): one misfit + reference gradient
    of all shots, every state kept, passes sized to the free HBM."""
    from pyfwi_amd import workloads
    from pyfwi_amd.batch import UnevenBatch
    wl = workloads.uneven_batch(ns=args.shots or 64)
    B = UnevenBatch(wl.inpa, wl.specs, wl.model_shape, components=wl.components)
    plan = B.plan
    plan.set_stream(torch.cuda.current_stream().cuda_stream)
    dev = torch.device("cuda:0")
    src = torch.from_numpy(B.src_table).to(dev)
    obs = torch.zeros(B.n_records, dtype=torch.float32, device=dev)
    seis = torch.zeros_like(obs)
    grad = torch.zeros(6 * B.tnz * B.tnx, dtype=torch.float32, device=dev)
    units = B.cell_updates()
    log(f"bench: {wl.name}: {B.ns} shots, nt {min(B.nt)}..{max(B.nt)}, receivers "
        f"{min(B.nr)}..{max(B.nr)}, {units / 1e9:.1f} G cell-updates per gradient")
    B.set_model(wl.model_obs)
    B.forward(seis=obs, mem=_native.MEM_DEVICE, src_table=src)
    B.set_model(wl.model)
    torch.cuda.synchronize()
    chp = checkpoint_steps(B.nt_max, 100)
    mask = 3  # vx, vz records (components = 2)
    graphs = args.graphs if args.graphs >= 0 else int(B.tnz * B.tnx < 1_000_000)
    plan.set_graphs(bool(graphs))
    bits = (_native.TUNE_ADJ_FUSED if args.adj_fused else 0) | \
        (_native.TUNE_FUSED_STEPS if args.fused_fwd else 0) | \
        (_native.TUNE_FWD_TWO if args.fused_fwd == 2 else 0)
    plan.set_tuning(0, bits, args.rows)

    def run(shots):
        return plan.misfit_gradient(shots, src, obs, seis, grad, chp, mask, 0,
                                    _native.MEM_DEVICE)

    def step(_W=None, _b=None):
        return run(B.shots())

    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"bench: warmup {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    misfit = None
    for i in range(args.steps):
        ts = time.perf_counter()
        misfit = step()
        torch.cuda.synchronize()
        log(f"bench: step {i + 1}/{args.steps}: {time.perf_counter() - ts:.2f} s")
    elapsed = (time.perf_counter() - t0) / args.steps
    value = units / elapsed / 1e9
    log(f"bench: {elapsed * 1e3:.0f} ms per gradient, {value:.2f} Gcell/s")
    passes = plan.last_passes()
    if args.minimal:
        print(json.dumps({"metric": "fwd+adj gradient Gcell-updates/s, batched vs per-shot, "
                                    "% HBM roofline", "value": round(value, 3),
                          "unit": "Gcell-updates/s", "n_gpus": 1, "steps": args.steps,
                          "warmup": args.warmup, "ms_per_step": round(elapsed * 1e3, 1),
                          "config": {"workload": wl.name, "passes": passes},
                          "note": "--minimal: timed steps only"}), flush=True)
        return
    # the passes take consecutive runs of the plan's shot order
    order_ids = plan.last_order()
    pass_nts, k = [], 0
    for n in passes:
        pass_nts.append([int(B.nt[j]) for j in order_ids[k:k + n]])
        k += n
    log(f"bench: passes {passes}, nt {pass_nts}")

    # dominant kernel: its launches' algorithmic bytes over their durations
    cells = B.tnz * B.tnx
    sdo, npml = int(B.base.sdo), B.npml
    interior = (B.tnz - 2 * (npml + sdo) + 1) * (B.tnx - 2 * (npml + sdo) + 1)
    order = 2 * sdo
    prof = kernel_profile(None, {"plan": plan}, torch, step)
    kernels = {}
    for kname, (ms, n) in prof.items():
        nbytes, nl = pass_kernel_bytes(kname, pass_nts, cells, interior, order,
                                       two=args.fused_fwd == 2)
        if nl != n:
            log(f"bench: {kname}: {n} launches timed, {nl} in the byte model")
        kernels[kname] = {"avg_us": round(ms / n * 1e3, 2), "launches": n,
                          "bytes_per_launch": round(nbytes / nl),
                          "GBps": round(nbytes / (ms / 1e3) / 1e9, 1)}
    dom = max(prof, key=lambda k: prof[k][0])
    log(f"bench: kernels {kernels}")
    achieved = kernels[dom]["GBps"]
    # launches of the dominant kernel by the number of shots they advance
    launch_mix = {}
    for live in launch_live_shots(dom, pass_nts, two=args.fused_fwd == 2):
        launch_mix[int(live)] = launch_mix.get(int(live), 0) + 1

    # forward modelling alone: all shots, records only
    no_chp = np.zeros(0, np.int32)
    plan.forward(B.shots(), src, seis, no_chp, 0, _native.MEM_DEVICE)
    torch.cuda.synchronize()
    ts = time.perf_counter()
    plan.forward(B.shots(), src, seis, no_chp, 0, _native.MEM_DEVICE)
    torch.cuda.synchronize()
    fwd_rate = units / (time.perf_counter() - ts) / 1e9
    log(f"bench: forward-only {fwd_rate:.2f} Gcell/s")
    copy_gbs = copy_rate(torch) / 1e9
    log(f"bench: device copy {copy_gbs:.0f} GB/s")

    ps = None
    if args.per_shot_sample > 0:
        # the same kernels (the same launch shapes) on one shot at a time (the
        # reference's schedule), on shots spread over the range of nt
        order_ids = np.argsort(B.nt)
        pick = order_ids[np.linspace(0, B.ns - 1, args.per_shot_sample).round().astype(int)]
        sub = B.shots([int(k) for k in pick])
        sub_units = cells * sum(int(B.nt[int(k)]) for k in pick)
        plan.set_tuning(1, bits, 0)
        run(sub[:1])  # warm
        torch.cuda.synchronize()
        ts = time.perf_counter()
        run(sub)
        torch.cuda.synchronize()
        ps_time = time.perf_counter() - ts
        plan.set_tuning(0, bits, args.rows)
        ps_rate = sub_units / ps_time / 1e9
        log(f"bench: per-shot path {ps_rate:.2f} Gcell/s on shots {list(map(int, pick))}")
        ps = {"value": round(ps_rate, 3), "unit": "Gcell-updates/s",
              "shots_sampled": [int(k) for k in pick],
              "nt_sampled": [int(B.nt[int(k)]) for k in pick],
              "batched_over_per_shot": round(value / ps_rate, 3)}

    cpu = cpu1 = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", 0) or os.cpu_count() or 1)
        log(f"bench: cpu baseline, {threads} threads, {args.cpu_steps} steps")
        sp, P0 = wl.specs[0], B.props[0]
        inpa = dict(wl.inpa, rec_dis=float(sp.rec_loc[1, 0] - sp.rec_loc[0, 0]))
        args_cpu = (inpa, sp.src_loc, sp.rec_loc, wl.model, P0.src.w, P0.src.component,
                    wl.components)
        cpu = cpu_baseline_shot(*args_cpu, args.cpu_steps, threads)
        log(f"bench: cpu baseline, 1 thread, {args.cpu_steps_1t} steps")
        cpu1 = cpu_baseline_shot(*args_cpu, args.cpu_steps_1t, 1)

    print(json.dumps({
        "metric": "fwd+adj gradient Gcell-updates/s, batched vs per-shot, % HBM roofline",
        "value": round(value, 3), "unit": "Gcell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed * 1e3, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded random-layered model, per-shot Ricker sources)",
        "config": {"workload": wl.name, "nz": int(wl.model_shape[0]),
                   "nx": int(wl.model_shape[1]), "npml": npml, "cells": cells,
                   "order": order, "shots": B.ns, "nt_range": [min(B.nt), max(B.nt)],
                   "receivers_range": [min(B.nr), max(B.nr)],
                   "fdom_range": [round(min(s.fdom for s in wl.specs), 2),
                                  round(max(s.fdom for s in wl.specs), 2)],
                   "gradient": "reference adjoint, every forward state kept in HBM (chpr=100), "
                               "shots packed first-fit by decreasing nt into passes whose "
                               "kept states fill the free HBM; L2 misfit of vx, vz on the GPU",
                   "passes": passes, "cell_updates": units, "graphs": bool(graphs),
                   "forward_step": {0: "two launches", 1: "one launch",
                                    2: "two steps a launch"}[args.fused_fwd],
                   "parallelism": "single GPU, shots batched per launch"},
        "kernel_timing": "per-step time of each time loop (HIP events around the loops) split "
                         "between its kernels by event-bracketed launches",
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(wl.name, dom),
                     "bytes_per_launch": kernels[dom]["bytes_per_launch"],
                     "launches_by_live_shots": {str(k): v for k, v in sorted(launch_mix.items())},
                     "avg_launch_us": kernels[dom]["avg_us"],
                     "device_copy_GBps": round(copy_gbs, 1),
                     "achieved_over_copy": round(achieved / copy_gbs, 4)},
        "gradient_roofline_frac": round(value * 148.0 / HBM_PEAK_GBS, 4),
        "forward_only": {"value": round(fwd_rate, 3), "unit": "Gcell-updates/s",
                         "roofline_frac": round(fwd_rate * 52.0 / HBM_PEAK_GBS, 4)},
        "misfit": misfit,
        "per_shot": ps,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "cpu_baseline_1thread": cpu1}), flush=True)


def _inversion_setup(args, torch):
    """The workload with observed data of its baseline (and, for the
    time-lapse pair, monitor) model, and the initial model."""
    from pyfwi_amd import workloads
    from pyfwi_amd.wave_propagation import WavePropagator
    w = getattr(workloads, args.workload)(**({"ns": args.shots} if args.shots else {}))
    if args.nt:
        w.inpa["t"] = workloads.t_for_nt(args.nt, w.inpa["dt"])
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=w.components,
                       chpr=w.chpr)
    t0 = time.perf_counter()
    obs = {"base": W.forward_modeling(w.model_obs)}
    if "monitor" not in w.meta:
        # a -8 % vp patch in the middle of the model
        nz, nx = w.model_shape
        mon = {k: np.array(v) for k, v in w.model_obs.items()}
        mon["vp"][nz // 2 - nz // 10:nz // 2 + nz // 10, nx // 2 - nx // 10:nx // 2 + nx // 10] \
            *= np.float32(0.92)
        w.meta["monitor"] = mon
    if args.mode == "timelapse":
        obs["monitor"] = W.forward_modeling(w.meta["monitor"])
    W._plan.close()
    log(f"bench: observed data of {len(obs)} model(s), {time.perf_counter() - t0:.1f} s")
    m0 = w.meta.get("initial", w.model)
    return w, obs, m0


KMS = 1000.0  # the inversions' unknowns: vp, vs in km/s and rho in g/cm^3


def km_s_parameters():
    """param_functions (the reference's hook for other parameterisations,
    fwi.py:52-66) whose unknowns are vp, vs and rho in units of 1000 (km/s,
    g/cm^3): of order one, so that the L-BFGS first step (1/|g|) is of the
    model's scale.  Models given to and returned by the drivers are in these
    units; the physics and the misfit are unchanged and the gradient follows
    the chain rule."""
    from pyfwi_amd import fwi_tools as T

    def model_to_dv(m, args):
        return {k: np.float32(v * KMS) for k, v in m.items()}

    def grad_from_dv(g, args, m):
        return {k: v * KMS for k, v in g.items()}
    return {"dict2vec": T.vel_dict2vec, "vec2dict": T.vec2vel_dict,
            "model_to_dv": model_to_dv, "grad_from_dv": grad_from_dv}


def _in_km_s(m):
    return {k: np.float32(v / KMS) for k, v in m.items()}


class _Budget(Exception):
    """The run's time budget is spent; the results so far are reported."""


def _instrument(fwi, m0, obs, budget, t_start):
    """Evaluate the misfit gradient at m0 once and scale the wavelet and the
    observed data (the problem is linear in the source amplitude: misfit and
    gradient scale with its square) so that |g| = 0.03 in the km/s unknowns,
    which puts the L-BFGS first step at about a hundredth of the model's
    norm; the scaled first evaluation is then served to the optimiser.  Every
    later evaluation is timed and logged; past `budget` seconds (from
    t_start) the run stops.  Returns (evaluations [(seconds, misfit)],
    scale)."""
    x0 = np.asarray(fwi.dict2vec(m0))
    t0 = time.perf_counter()
    rms, g = fwi.fprime(np.float32(x0), None)
    t_first = time.perf_counter() - t0
    norm = float(np.linalg.norm(g))
    a = float(np.sqrt(0.03 / norm)) if norm > 0 else 1.0
    fwi.src.w = np.float32(fwi.src.w * a)
    for d in obs:
        for k in d:
            d[k] *= np.float32(a)
    fwi.__dict__.pop("_obs_cache", None)
    first = (np.float32(x0), rms * a * a, g * (a * a))
    evals = [(t_first, float(first[1]))]
    log(f"bench: first gradient {t_first:.1f} s (misfit {rms:.4e}, |g| {norm:.3e}); scaled "
        f"by {a:.3e}")
    inner = fwi.fprime

    base = obs[0]

    def fprime(m, freq):
        if freq is None and fwi.d_obs is base and np.array_equal(np.float32(m), first[0]):
            return first[1], first[2]
        if time.perf_counter() - t_start > budget:
            raise _Budget()
        log(f"bench: evaluation {len(evals) + 1} starts")
        ts = time.perf_counter()
        out = inner(m, freq)
        evals.append((time.perf_counter() - ts, float(out[0])))
        log(f"bench: evaluation {len(evals)}: misfit {out[0]:.6e}, {evals[-1][0]:.1f} s")
        return out
    fwi.fprime = fprime
    return evals, a


def bench_lbfgs(args, torch):
    """L-BFGS-B inversion (the FWI driver with the reference's settings) of
    the workload's observed data from its initial model, vp only, one band
    with no lowpass, in km/s unknowns: every misfit evaluation is one
    pyfwi_misfit_gradient call of all shots (reference adjoint, the
    workload's checkpoint list).  Stops after `--time-budget` seconds."""
    from pyfwi_amd.fwi import FWI
    t_start = time.perf_counter()
    w, obs, m0 = _inversion_setup(args, torch)
    fwi = FWI(obs["base"], w.inpa, w.src, w.rec_loc, w.model_shape, components=w.components,
              chpr=w.chpr, param_functions=km_s_parameters())
    fwi.verbose = False
    m0 = _in_km_s(m0)
    evals, a = _instrument(fwi, m0, [fwi.d_obs], args.time_budget, t_start)
    iters = []
    from pyfwi_amd import fwi as fwi_mod
    run = fwi_mod._run_lbfgs

    def counting(f, fp, x0, n):
        from scipy.optimize import fmin_l_bfgs_b
        x, fopt, _ = fmin_l_bfgs_b(f, x0, fp, m=10, factr=1e7, pgtol=1e-8, bounds=None,
                                   maxfun=15000, maxiter=n, maxls=10,
                                   callback=lambda xk: iters.append(len(evals)))
        return x, fopt
    fwi_mod.OPTIMIZERS["lbfgs"] = counting
    stopped = None
    t0 = time.perf_counter()
    try:
        m_est, hist = fwi(m0, method="lbfgs", iter=[args.iters], freqs=[None], n_params=1,
                          k_0=1, k_end=2)
        err1 = float(np.linalg.norm(m_est["vp"] * KMS - w.model_obs["vp"]))
    except _Budget:
        stopped, err1 = "time budget", None
    finally:
        fwi_mod.OPTIMIZERS["lbfgs"] = run
    total = time.perf_counter() - t0
    cells = (w.model_shape[0] + 2 * w.inpa["npml"]) * (w.model_shape[1] + 2 * w.inpa["npml"])
    units = cells * len(w.src.i) * int(fwi.nt)
    t_mean = float(np.mean([e[0] for e in evals]))
    print(json.dumps({
        "metric": "L-BFGS inversion: misfit evaluations (forward + L2 misfit + reference "
                  "gradient) of the full workload", "workload": w.name,
        "shots": int(len(w.src.i)), "nt": int(fwi.nt), "cells": cells,
        "iterations_asked": args.iters, "iterations_done": len(iters),
        "evaluations_at_iteration_ends": iters, "evaluations": len(evals),
        "stopped": stopped, "seconds": round(total, 1),
        "seconds_per_evaluation": round(t_mean, 2),
        "gradient_gcell_updates_per_s": round(units / t_mean / 1e9, 3),
        "misfit_history": [e[1] for e in evals],
        "vp_error_l2": {"initial": float(np.linalg.norm(m0["vp"] * KMS - w.model_obs["vp"])),
                        "final": err1},
        "scale": a, "unknowns": "vp in km/s (param_functions)",
        "checkpoints": f"chpr={w.chpr}"}), flush=True)


def bench_timelapse(args, torch):
    """Time-lapse pair (baseline and monitor observed data, the monitor with
    vp -8 % in a patch) inverted with the cascaded strategy: baseline from
    the initial model, monitor from the baseline result; `--iters` L-BFGS
    iterations each, km/s unknowns."""
    from pyfwi_amd.tl_fwi import TimeLapse
    t_start = time.perf_counter()
    w, obs, m0 = _inversion_setup(args, torch)
    tl = TimeLapse(obs["base"], obs["monitor"], w.inpa, w.src, w.rec_loc, w.model_shape,
                   components=w.components, chpr=w.chpr, param_functions=km_s_parameters())
    tl.verbose = tl.inv_obj.verbose = False
    inv = tl.inv_obj
    m0 = _in_km_s(m0)
    inv.d_obs = tl.b_dobs
    evals, a = _instrument(inv, m0, [tl.b_dobs, tl.m_dobs], args.time_budget, t_start)
    tl.src.w = inv.src.w
    t0 = time.perf_counter()
    out, hist, stopped = {}, {}, None
    try:
        out, hist = tl(m0, iter=[args.iters], freqs=[None], tl_method="cc", n_params=1, k_0=1,
                       k_end=2)
    except _Budget:
        stopped = "time budget"
    total = time.perf_counter() - t0
    line = {"metric": "time-lapse pair, cascaded strategy", "workload": w.name,
            "shots": int(len(w.src.i)), "iterations_per_inversion": args.iters,
            "evaluations": len(evals), "stopped": stopped, "seconds": round(total, 1),
            "seconds_per_evaluation": round(float(np.mean([e[0] for e in evals])), 2),
            "misfit_history": [e[1] for e in evals], "scale": a,
            "unknowns": "vp in km/s (param_functions)"}
    if out:
        dm = out["dm"]["vp"] * KMS
        patch = w.meta["monitor"]["vp"] != w.model_obs["vp"]
        line.update(dvp_patch_mean=float(dm[patch].mean()),
                    dvp_outside_abs_mean=float(np.abs(dm[~patch]).mean()),
                    hist={k: np.asarray(v).ravel().tolist() for k, v in hist.items()},
                    pt=out.get("pt"))
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="uneven_batch",
                    help="uneven_batch (1000x2000, 64 uneven shots; the default), plain_medium, "
                         "large_step, differentiation, plain_small, crosswell_medium, "
                         "wells_surface_medium")
    ap.add_argument("--nt", type=int, default=0, help="override time steps (profiling runs)")
    ap.add_argument("--order", type=int, default=0, choices=[0, 4, 8],
                    help="stencil order (0: the workload's own, 4)")
    ap.add_argument("--shots", type=int, default=0,
                    help="override the number of shots (e.g. one GPU's share of large_step)")
    ap.add_argument("--per-shot-sample", type=int, default=0,
                    help="shots run one per pass for the per-shot rate (0: the workload's "
                         "default, 4 for uneven_batch, 2 otherwise; -1 skips)")
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="time steps of the CPU baseline (0: 600 for uneven_batch, 2000 else)")
    ap.add_argument("--cpu-steps-1t", type=int, default=0,
                    help="time steps of the 1-thread CPU baseline (0: 80 / 150)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--rows", type=int, default=0,
                    help="uneven_batch: fixed strip height of the stencil kernels (0: automatic)")
    ap.add_argument("--graphs", type=int, default=-1, choices=[-1, 0, 1],
                    help="replay the time loops as captured hipGraphs (1), launch eagerly (0); "
                         "-1: graphs for grids under 1 M cells (launch-bound), eager otherwise")
    ap.add_argument("--minimal", action="store_true",
                    help="uneven_batch: the timed steps only (no kernel timing, forward-only, "
                         "per-shot or CPU legs), for counter runs")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--adj-fused", action="store_true",
                    help="reference adjoint in one launch a step instead of two "
                         "(stresses, then velocities)")
    ap.add_argument("--fused-fwd", type=int, default=1, choices=[0, 1, 2],
                    help="uneven_batch: each forward step in one launch (1, measured "
                         "1.4-2.4 %% faster over the gradient), two (0), or two steps "
                         "a launch (2)")
    ap.add_argument("--no-tune", action="store_true", help="skip the execution-shape autotuner")
    ap.add_argument("--tune", default="", help="fixed tuning 'shots_per_pass,fused,rows'")
    ap.add_argument("--iters", type=int, default=5,
                    help="L-BFGS iterations of --mode lbfgs / timelapse")
    ap.add_argument("--time-budget", type=float, default=1e9,
                    help="--mode lbfgs / timelapse: stop (and report) after this many seconds")
    ap.add_argument("--mode", default="kept",
                    choices=["kept", "reverse", "exact", "exact-kept", "lbfgs", "timelapse"],
                    help="kept: every forward state kept in HBM per pass (chpr=100), misfit and "
                         "gradient in one call; reverse: chpr=0, states rebuilt by time reversal; "
                         "exact: the exact adjoint from checkpoints sized to the free HBM; "
                         "exact-kept: the exact adjoint reading every forward state, all shots' "
                         "states kept in HBM; lbfgs / timelapse: the inversion drivers on the "
                         "workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(2)
    torch.cuda.set_device(0)

    log("bench: building problem")
    if args.copy_calibration:
        copy_calibration(torch)
        if args.steps == 0:
            return
    big = args.workload == "uneven_batch"
    if args.per_shot_sample == 0:
        args.per_shot_sample = 4 if big else 2
    args.cpu_steps = args.cpu_steps or (600 if big else 2000)
    args.cpu_steps_1t = args.cpu_steps_1t or (80 if big else 150)
    if big:
        return bench_uneven(args, torch)
    if args.mode == "lbfgs":
        return bench_lbfgs(args, torch)
    if args.mode == "timelapse":
        return bench_timelapse(args, torch)
    w, W = build_problem(args.workload, args.nt or None, args.shots or None, args.order)
    b = device_setup(w, W, torch)
    cells = int(W.tnz) * int(W.tnx)
    sdo, npml = int(W.sdo), int(W.npml)
    interior = (int(W.tnz) - 2 * (npml + sdo) + 1) * (int(W.tnx) - 2 * (npml + sdo) + 1)
    units = int(W.ns) * cells * W.nt
    log(f"bench: {w.name}: {int(W.ns)} shots x {W.nt} steps x {cells} cells")

    step = {"kept": kept_step, "reverse": gradient_step, "exact": exact_step,
            "exact-kept": exact_kept_step}[args.mode]
    state_bytes = 5 * 4 * (int(W.tnz) + 16) * (((int(W.tnx) + 8 + 63) // 64) * 64)
    free, _total = torch.cuda.mem_get_info()
    max_pass = int(0.8 * free // (state_bytes * W.nt)) if args.mode == "kept" else None
    if args.mode == "kept" and max_pass < 1:
        print(json.dumps({"error": "kept states of one shot do not fit in memory"}))
        sys.exit(2)
    tuning = (max_pass or 0, False, 0)
    if args.tune:
        g, f, r = (int(x) for x in args.tune.split(","))
        tuning = (g, bool(f), r)
        b["plan"].set_tuning(*tuning)
    elif not args.no_tune and not args.mode.startswith("exact"):
        from pyfwi_amd import tuning as tn
        if args.mode == "kept":
            trial = tn.truncate(b["shots"], 120)
            chp_t = checkpoint_steps(trial[0].nt, 100)

            def run():
                b["plan"].misfit_gradient(trial, b["src"], b["obs"], b["seis"], b["grad"],
                                          chp_t, 0, 0, _native.MEM_DEVICE)
            cands = tn.default_candidates(int(W.ns), max_pass,
                                          fused=(0, _native.TUNE_FUSED_STEPS,
                                                 _native.TUNE_FUSED_STEPS | _native.TUNE_FWD_TWO,
                                                 _native.TUNE_ADJ_FUSED))
            tuning, _ = tn.autotune(b["plan"], run, torch.cuda.synchronize, cands, log=log)
        else:
            tuning, _ = tn.autotune_gradient(b["plan"], b["shots"], b["src"], b["seis"], b["res"],
                                             b["grad"], torch.cuda.synchronize, log=log)
    else:
        b["plan"].set_tuning(*tuning)
    if args.adj_fused:
        tuning = (tuning[0], int(tuning[1]) | _native.TUNE_ADJ_FUSED, tuning[2])
        b["plan"].set_tuning(*tuning)
    graphs = args.graphs if args.graphs >= 0 else int(cells < 1_000_000)
    b["plan"].set_graphs(bool(graphs))
    log(f"bench: mode={args.mode} tuning shots/pass={tuning[0]} fused={int(tuning[1])} "
        f"rows={tuning[2]} graphs={graphs}")

    for i in range(args.warmup):
        step(W, b)
        torch.cuda.synchronize()
        log(f"bench: warmup {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    misfit = None
    for i in range(args.steps):
        misfit = step(W, b)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    value = units / (elapsed / args.steps) / 1e9
    log(f"bench: timed {args.steps} steps, {ms_per_step:.1f} ms/step, {value:.2f} Gcell/s")
    if args.minimal:
        passes, ck = b["plan"].last_passes(with_interval=True)
        print(json.dumps({"metric": "fwd+adj gradient Gcell-updates/s", "value": round(value, 3),
                          "unit": "Gcell-updates/s", "steps": args.steps,
                          "ms_per_step": round(ms_per_step, 1), "mode": args.mode,
                          "config": {"workload": w.name, "shots": int(W.ns), "nt": int(W.nt),
                                     "cells": cells, "passes": passes, "ckpt_every": ck},
                          "note": "--minimal: timed steps only"}), flush=True)
        return

    prof = kernel_profile(W, b, torch, step)
    dom = max(prof, key=lambda k: prof[k][0])
    ms, n = prof[dom]
    avg_s = ms / n / 1e3
    # passes of `tuning[0]` shots, the last one possibly smaller; every pass
    # makes the same number of launches, so a launch's mean bytes is the mean
    # over the passes
    per = tuning[0] or int(W.ns)
    passes = [min(per, int(W.ns) - g0) for g0 in range(0, int(W.ns), per)]
    shots_per_launch = int(W.ns) / len(passes)
    order = int(W.sdo) * 2

    def mean_bytes(kind):
        return sum(launch_bytes(kind, cells, interior, n, order) for n in passes) / len(passes)
    per_launch_bytes = mean_bytes(dom)
    achieved = per_launch_bytes / avg_s / 1e9
    kernels = {k: {"avg_us": round(v[0] / v[1] * 1e3, 2), "launches": v[1],
                   "GBps": round(mean_bytes(k) / (v[0] / v[1] / 1e3) / 1e9, 1)}
               for k, v in prof.items()}
    log(f"bench: kernels {kernels}")

    fwd_rate, _ = forward_rate(W, b, torch, tuning)
    log(f"bench: forward-only {fwd_rate:.2f} Gcell/s")

    ps = None
    if args.per_shot_sample > 0:
        ps_rate, ps_time = per_shot_rate(W, b, torch, args.per_shot_sample, step)
        b["plan"].set_tuning(*tuning)
        b["plan"].set_graphs(bool(graphs))
        log(f"bench: per-shot path {ps_rate:.2f} Gcell/s")
        ps = {"value": round(ps_rate, 3), "shots_sampled": args.per_shot_sample,
              "batched_over_per_shot": round(value / ps_rate, 2)}

    cpu = cpu1 = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", 0) or os.cpu_count() or 1)
        log(f"bench: cpu baseline, {threads} threads")
        cpu = cpu_baseline(w, args.cpu_steps, threads)
        log(f"bench: cpu baseline, 1 thread")
        cpu1 = cpu_baseline(w, args.cpu_steps_1t, 1)

    grad_bytes = 148.0
    line = {
        "metric": "fwd+adj gradient Gcell-updates/s, batched vs per-shot, % HBM roofline",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random-layered model, Ricker sources)",
        "config": {"workload": w.name, "nz": int(W.nz), "nx": int(W.nx), "npml": int(W.npml),
                   "cells": cells, "shots": int(W.ns), "nt": int(W.nt), "receivers": int(W.nr),
                   "order": int(W.sdo) * 2,
                   "gradient": {
                       "kept": "reference adjoint, chpr=100: every forward state kept in HBM "
                               "per pass; L2 misfit on the GPU",
                       "reverse": "reference adjoint, chpr=0: states rebuilt by time reversal",
                       "exact": "exact adjoint, two-level checkpoints sized to the free HBM, "
                                "segments recomputed",
                       "exact-kept": "exact adjoint reading every kept forward state"}[args.mode],
                   "parallelism": "single GPU, shots batched per launch",
                   "tuning": {"shots_per_pass": tuning[0],
                              "fused_steps": bool(int(tuning[1]) & _native.TUNE_FUSED_STEPS),
                              "two_steps_a_launch": bool(int(tuning[1]) & _native.TUNE_FWD_TWO),
                              "adjoint_one_launch": bool(int(tuning[1]) & _native.TUNE_ADJ_FUSED),
                              "strip_rows": tuning[2], "graphs": bool(graphs)}},
        "kernel_timing": ("event-bracketed launches (each holds a few microseconds of the "
                          "event's stall)" if args.mode.startswith("exact") else
                          "per-step time of each time loop (HIP events around the loops) split "
                          "between its kernels by event-bracketed launches"),
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": (measured_traffic(w.name, dom, passes[0])
                                 if len(set(passes)) == 1 else None),
                     "bytes_per_launch": per_launch_bytes, "avg_launch_us": round(avg_s * 1e6, 2),
                     "shots_per_launch": round(shots_per_launch, 3)},
        "gradient_roofline_frac": round(value * grad_bytes / HBM_PEAK_GBS, 4),
        "forward_only": {"value": round(fwd_rate, 3), "unit": "Gcell-updates/s",
                         "roofline_frac": round(fwd_rate * 52.0 / HBM_PEAK_GBS, 4)},
        "misfit": misfit,
        "per_shot": ps,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "cpu_baseline_1thread": cpu1,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
