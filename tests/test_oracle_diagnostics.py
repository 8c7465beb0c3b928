"""Closed-form pins of the oracle's diagnostics (compute_KE + compute_divB,
driver.f90:492-553, plus the mass and max|div B| the library adds).

The reference's compute_KE / compute_divB are internal procedures of the driver program,
so no reference build can call them; these states fix every quantity exactly instead:

* dyadic uniform states on a non-cubic box (dx = 1 / max(n), so dx^3 = 2^-9 for 8 x 4 x 2):
  every product and partial sum is exact, so ke, v_rms^2, momentum and mass are known to
  the last bit -- a lost 1/2, a dx^2 for dx^3 or a 1/nx for 1/max(n) changes them;
* the clamps rho in [1e-15, 1e10] and v^2 <= 1e5 (driver.f90:535-536), one planted cell at
  a time: a density below the lower clamp, a speed above the v^2 cap, a density above the
  upper clamp (which the mass, unclamped, still counts in full);
* div B: a uniform face field has none; a planted face perturbation delta adds +delta to
  the cell on its low side and -delta to the cell on its high side (driver.f90:504-505),
  so an interior face leaves the signed sum at 0 and max|div B| at delta, while a face on
  the box's upper (lower) boundary has only its low (high) cell inside: sum +delta
  (-delta).  A swapped index (i+1 for i-1) or a dropped component fails one of them.
"""
import numpy as np
import pytest

from refutil import Oracle

NB = 6
DIMS = (8, 4, 2)
DX3 = (1.0 / 8) ** 3
NCELL = 8 * 4 * 2
I = tuple(slice(NB, NB + n) for n in DIMS)


def state(rho=1.0, m=(0.0, 0.0, 0.0), bface=(0.5, 0.25, -1.0)):
    sh = tuple(n + 2 * NB for n in DIMS)
    q = np.zeros(sh + (7,), order="F")
    bf = np.zeros(sh + (3,), order="F")
    q[..., 0] = rho
    for d in range(3):
        q[..., 1 + d] = m[d]
        q[..., 4 + d] = bface[d]
        bf[..., d] = bface[d]
    return q, bf


def diag(q, bf):
    O = Oracle(*DIMS, npass=0)
    O.set_state(q, bf)
    return O.diagnostics()


def cell(i, j, k):
    """padded index of interior cell (i, j, k), reference numbering from 1"""
    return (NB - 1 + i, NB - 1 + j, NB - 1 + k)


def test_uniform_dyadic_state_exact():
    q, bf = state(rho=2.0, m=(1.0, -0.5, 0.25))
    d = diag(q, bf)
    v2 = (1.0 + 0.25 + 0.0625) / 4.0          # 0.328125, exact
    assert d["vrms_sq"] == NCELL * v2 * DX3 == 0.041015625
    assert d["ke"] == 0.5 * NCELL * 2.0 * v2 * DX3 == 0.041015625
    assert d["momx"] == 0.125 and d["momy"] == -0.0625 and d["momz"] == 0.03125
    assert d["mass"] == 0.25
    assert d["divb_sum"] == 0.0 and d["divb_max"] == 0.0


def test_ghost_cells_do_not_count():
    q, bf = state(rho=1.0, m=(0.5, 0.0, 0.0))
    q[:NB, ..., 0] = 1e6          # ghost slabs: never summed
    q[-NB:, ..., 1] = 123.0
    d = diag(q, bf)
    assert d["mass"] == NCELL * DX3
    assert d["momx"] == NCELL * 0.5 * DX3
    assert d["ke"] == 0.5 * NCELL * 0.25 * DX3


def test_lower_density_clamp_and_speed_cap():
    q, bf = state()
    c = cell(5, 3, 2)
    q[c + (0,)] = 1e-20           # rho -> 1e-15; v^2 = 1e-20 / 1e-30 = 1e10 -> 1e5
    q[c + (1,)] = 1e-10
    d = diag(q, bf)
    assert d["vrms_sq"] == 1e5 * DX3
    assert d["ke"] == 0.5 * (1e-15 * 1e5) * DX3
    assert d["momx"] == 1e-10 * DX3
    assert d["mass"] == ((NCELL - 1) + 1e-20) * DX3


def test_speed_cap_alone():
    q, bf = state()
    c = cell(1, 1, 1)
    q[c + (0,)] = 0.5
    q[c + (2,)] = -512.0           # v^2 = 2^18 / 2^-2 = 2^20 > 1e5
    d = diag(q, bf)
    assert d["vrms_sq"] == 1e5 * DX3
    assert d["ke"] == 0.5 * (0.5 * 1e5) * DX3
    assert d["momy"] == -512.0 * DX3


def test_upper_density_clamp_mass_unclamped():
    q, bf = state()
    c = cell(8, 4, 1)
    q[c + (0,)] = 4e10            # rho -> 1e10 for v^2 and ke; v^2 = (2e10)^2 / 1e20 = 4
    q[c + (3,)] = 2e10
    d = diag(q, bf)
    assert d["vrms_sq"] == 4.0 * DX3
    assert d["ke"] == 0.5 * (1e10 * 4.0) * DX3
    assert d["mass"] == ((NCELL - 1) + 4e10) * DX3


DELTA = 2.0 ** -10


@pytest.mark.parametrize("comp,face,expect_sum", [
    (0, (3, 2, 1), 0.0),          # interior x-face between cells 3 and 4
    (0, (8, 2, 1), DELTA),        # upper x boundary: only cell 8 is inside
    (0, (0, 3, 2), -DELTA),       # lower x boundary face (ghost side): only cell 1
    (1, (5, 2, 2), 0.0),
    (1, (5, 4, 1), DELTA),
    (1, (2, 0, 1), -DELTA),
    (2, (7, 3, 1), 0.0),          # the one interior z-face of a 2-plane box
    (2, (7, 3, 2), DELTA),
    (2, (1, 1, 0), -DELTA),
])
def test_planted_face_divergence(comp, face, expect_sum):
    q, bf = state()
    bf[cell(*face) + (comp,)] += DELTA
    d = diag(q, bf)
    assert d["divb_sum"] == expect_sum
    assert d["divb_max"] == DELTA


def test_two_planted_faces_max_not_sum():
    q, bf = state()
    bf[cell(2, 2, 1) + (0,)] += DELTA          # cells (2,2,1) +d, (3,2,1) -d
    bf[cell(3, 2, 1) + (1,)] -= 2 * DELTA      # cells (3,2,1) -2d, (3,3,1) +2d
    d = diag(q, bf)
    assert d["divb_max"] == 3 * DELTA          # cell (3,2,1): -d - 2d
    assert d["divb_sum"] == 0.0
