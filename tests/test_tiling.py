"""The tile arithmetic of the multi-stage column-marching kernels, on the CPU.

k_fwd_step and k_adj_fused give each wave 64 columns of which it owns the
middle 64 - 2*H (H = SDO halo lanes a side), k_fwd_two the middle 64 - 6*SDO
(H = 3*SDO); lane l of column block cb sits on column cb*SPAN + l + SDO - H
(tile_fused in pyfwi_amd/csrc/pyfwi_kernels.hip).  Checked here for every
grid width up to a few thousand columns: the owned columns are the active
columns SDO .. nx-SDO, each exactly once.  And for every strip height the
floats the kernels read stay inside the padded fields (HALO zero rows above
and below the grid, a tail after the last row) and the damping profile's
halo: k_fwd_step reads rows
i0 - 2*SDO .. i1 + 2*SDO + FWD_STEP_UNROLL - 1 of a strip i0..i1 (the
unrolled loop may run up to FWD_STEP_UNROLL - 1 rows past the strip), and
k_fwd_two rows i0 - 4*SDO .. i1 + 4*SDO (its loop stops at the strip's end).
The constants are read from the kernels' header.
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = open(os.path.join(ROOT, "pyfwi_amd", "csrc", "pyfwi_kernels.h")).read()


def _const(pattern):
    m = re.search(pattern, HEADER)
    assert m, pattern
    return int(m.group(1))


HALO = _const(r"constexpr int HALO = (\d+);")
FWD_STEP_UNROLL = _const(r"#define FWD_STEP_UNROLL (\d+)")


def _owned_columns(nx, sdo, halo_lanes):
    span = 64 - 2 * halo_lanes
    ncb = -(-(nx - 2 * sdo + 1) // span)
    owned = []
    for cb in range(ncb):
        for lane in range(64):
            j = cb * span + lane + sdo - halo_lanes
            if halo_lanes <= lane < 64 - halo_lanes and sdo <= j <= nx - sdo:
                owned.append(j)
    return owned


@pytest.mark.parametrize("sdo,halo_lanes", [(2, 2), (4, 4), (2, 6)])
def test_owned_columns_partition_the_active_columns(sdo, halo_lanes):
    for nx in list(range(2 * sdo + 1, 700)) + [1040, 1041, 2040, 2041, 6040]:
        owned = _owned_columns(nx, sdo, halo_lanes)
        assert sorted(owned) == list(range(sdo, nx - sdo + 1)), nx
        assert len(owned) == len(set(owned)), nx


def _strips(nz, sdo, rows):
    i0 = sdo
    while i0 <= nz - sdo:
        yield i0, min(i0 + rows - 1, nz - sdo)
        i0 += rows


def _layout(nz, nx, sdo):
    """Padded field of the plan (pyfwi_plan.hip): row pitch, offset of cell
    (0, 0) and floats per field."""
    pitch = -(-(nx + HALO) // 64) * 64
    origin = HALO * pitch + 64 - sdo
    fs = -(-((nz + 2 * HALO) * pitch + 64) // 64) * 64
    return pitch, origin, fs


@pytest.mark.parametrize("sdo,kernel", [(2, "step"), (4, "step"), (2, "two")])
def test_reads_stay_in_the_padded_fields(sdo, kernel):
    """Every float a strip reads lies inside its field: rows -HALO ..
    nz + HALO - 1 (which the damping profile's halo covers too), and the
    columns of the last column block, which run past the row into the next
    row's padding, inside the field's tail on its last row."""
    halo_lanes = sdo if kernel == "step" else 3 * sdo
    span = 64 - 2 * halo_lanes
    for nz, nx in ((2 * sdo + 1, 2 * sdo + 1), (17, 33), (100, 100), (540, 1040),
                   (1040, 2040), (301, 1999)):
        pitch, origin, fs = _layout(nz, nx, sdo)
        ncb = -(-(nx - 2 * sdo + 1) // span)
        # taps reach SDO columns beyond the wave's 64
        col_lo = sdo - halo_lanes - sdo
        col_hi = (ncb - 1) * span + 63 + sdo - halo_lanes + sdo
        for rows in range(1, 65):
            for i0, i1 in _strips(nz, sdo, rows):
                if kernel == "step":
                    lo, hi = i0 - 2 * sdo, i1 + 2 * sdo + FWD_STEP_UNROLL - 1
                else:
                    lo, hi = i0 - 4 * sdo, i1 + 4 * sdo
                assert lo >= -HALO and hi <= nz + HALO - 1, (nz, rows, i0, i1)
                assert origin + lo * pitch + col_lo >= 0, (nz, nx, rows)
                assert origin + hi * pitch + col_hi < fs, (nz, nx, rows)
