"""The measurement side of bench.py on the CPU: the byte model behind
roofline.achieved, the block shape it assumes, the measured-traffic lookup
and the one-GPU guard."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

CELLS, INTERIOR = 540 * 1040, 497 * 997  # C2: tnz x tnx, imaging region at order 4


@pytest.mark.parametrize("order,shots,spb", [(4, 10, 5), (4, 5, 5), (4, 20, 4), (4, 8, 4),
                                             (4, 1, 4), (8, 10, 4), (4, 6.5, 4)])
def test_shots_per_block(order, shots, spb):
    """Five-shot blocks only at order 4 when the shots fill fives but not fours."""
    assert bench.shots_per_block(order, shots) == spb


def test_byte_model_c2_launches():
    """Per-launch algorithmic bytes of the 10-shot C2 launches, written out
    from the per-cell reads and writes each kernel makes."""
    groups = 2  # ten shots in blocks of five
    assert bench.launch_bytes("fwd_vel", CELLS, INTERIOR, 10) == 32 * CELLS * 10
    assert bench.launch_bytes("fwd_str", CELLS, INTERIOR, 10) == 40 * CELLS * 10
    assert bench.launch_bytes("adj_str", CELLS, INTERIOR, 10) == \
        (36 * CELLS + 8 * INTERIOR) * 10 + 16 * INTERIOR * groups
    assert bench.launch_bytes("adj_vel_grad", CELLS, INTERIOR, 10) == \
        (36 * CELLS + 16 * INTERIOR) * 10 + 8 * INTERIOR * groups
    # the exact adjoint images every active cell
    assert bench.launch_bytes("exact_vel", CELLS, INTERIOR, 20) == 56 * CELLS * 20 + 8 * CELLS * 5
    # order 8 keeps blocks of four: three groups for ten shots
    assert bench.launch_bytes("adj_str", CELLS, INTERIOR, 10, order=8) == \
        (36 * CELLS + 8 * INTERIOR) * 10 + 16 * INTERIOR * 3


def test_one_launch_adjoint_bytes_are_the_adjoint_cell_update():
    """The one-launch adjoint step moves the adjoint cell-update's bytes:
    with one shot a group, 96 B on an interior cell (adjoint state read and
    written, forward state, model, three accumulators read and written), and
    the two launches it replaces move 20 B more per shot on every cell (the
    adjoint stresses read again, the adjoint velocities and b read by both) and
    4 B more on interior cells (b read again for the rho term)."""
    one = bench.launch_bytes("adj_step", INTERIOR, INTERIOR, 1)
    assert one == 96 * INTERIOR
    two = (bench.launch_bytes("adj_str", CELLS, INTERIOR, 1) +
           bench.launch_bytes("adj_vel_grad", CELLS, INTERIOR, 1))
    assert two - bench.launch_bytes("adj_step", CELLS, INTERIOR, 1) == 20 * CELLS + 4 * INTERIOR


def test_tuning_candidates_are_bit_sets():
    from pyfwi_amd import _native, tuning
    c = tuning.default_candidates(20, 10, fused=(0, _native.TUNE_ADJ_FUSED), rows=(12,))
    assert (10, 0, 0) in c and (10, _native.TUNE_ADJ_FUSED, 0) in c
    assert (10, _native.TUNE_ADJ_FUSED, 12) in c
    assert all(g <= 10 for g, _, _ in c)
    # True stands for the fused forward / reverse steps
    assert {f for _, f, _ in tuning.default_candidates(4, fused=(False, True))} == \
        {0, _native.TUNE_FUSED_STEPS}


def test_measured_traffic_lookup():
    assert bench.measured_traffic("no_such_workload", "adj_vel_grad") is None
    with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
        d = json.load(f)
    got = bench.measured_traffic(d["workload"], "adj_vel_grad")
    if d["kernel_source_sha1"] == bench.kernel_source_sha1():
        assert got == d["bytes_per_dispatch"]["adj_vel_grad"]
        n = d["shots_per_dispatch"]
        assert bench.measured_traffic(d["workload"], "adj_vel_grad", n) == got
        # dispatches of another size move other bytes: no figure
        assert bench.measured_traffic(d["workload"], "adj_vel_grad", n - 2) is None
    else:
        assert got is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_pass_bytes_even_shots_match_launch_bytes():
    """Passes of equal shots: the uneven byte model gives nt launches of the
    per-launch bytes (forward) or nt - 1 (adjoint)."""
    nt = 37
    for kind in ("fwd_vel", "fwd_str", "adj_str", "adj_vel_grad"):
        total, n = bench.pass_kernel_bytes(kind, [[nt] * 10, [nt] * 10], CELLS, INTERIOR)
        steps = nt - 1 if kind.startswith("adj") else nt
        assert n == 2 * steps
        assert total == 2 * steps * bench.launch_bytes(kind, CELLS, INTERIOR, 10)


def test_pass_bytes_uneven_shots():
    """Uneven shots: a launch takes only its live shots (a prefix of the pass,
    longest first), in blocks of five when that count is a multiple of five
    but not of four, else four; each block's group moves the accumulators
    once.  Pass of 6 shots with nt 10, 9, 8, 8, 5, 3: live 6, 6, 5, 5, 4, 4,
    4, 2, 1 at t = 1..9, so 2, 2, 1, 1, 1, 1, 1, 1, 1 groups."""
    nts = [10, 9, 8, 8, 5, 3]
    a, b, c = bench.KERNEL_BYTES["adj_vel_grad"]
    want = 0
    groups_seen = []
    for t in range(1, 10):
        live = sum(n > t for n in nts)
        spb = 5 if live % 5 == 0 and live % 4 != 0 else 4
        groups = -(-live // spb)
        groups_seen.append(groups)
        want += (a * CELLS + b * INTERIOR) * live + c * INTERIOR * groups
    assert groups_seen == [2, 2, 1, 1, 1, 1, 1, 1, 1]
    total, n = bench.pass_kernel_bytes("adj_vel_grad", [nts], CELLS, INTERIOR)
    assert n == 9 and total == want
    total, n = bench.pass_kernel_bytes("fwd_vel", [nts], CELLS, INTERIOR)
    assert n == 10 and total == 32 * CELLS * sum(nts)


def test_two_step_pairs_follow_the_live_shots():
    """Two forward steps share a launch only where the same shots are live at
    both; the byte model splits the forward's bytes between the two-step and
    the one-step launches without losing a step."""
    nts = [7, 5, 5]  # live 3,3,3,3,3,1,1 at t = 0..6
    live = bench.launch_live_shots("fwd_step", [nts])
    assert list(live) == [3, 3, 3, 3, 3, 1, 1]
    assert list(bench.two_step_pairs(live)) == [0, 2, 5]  # step 4 alone
    two = bench.launch_live_shots("fwd_two", [nts], two=True)
    one = bench.launch_live_shots("fwd_step", [nts], two=True)
    assert list(two) == [3, 3, 1] and list(one) == [3]
    b2, n2 = bench.pass_kernel_bytes("fwd_two", [nts], CELLS, INTERIOR, two=True)
    b1, n1 = bench.pass_kernel_bytes("fwd_step", [nts], CELLS, INTERIOR, two=True)
    assert (n2, n1) == (3, 1)
    # 72 B for two steps, 52 B for one: 7 shot-steps in all
    assert b2 == 72 * CELLS * 7 and b1 == 52 * CELLS * 3
    assert len(bench.launch_live_shots("fwd_two", [nts])) == 0
