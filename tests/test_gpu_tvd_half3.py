"""The three CTU half updates of the K98 TVD corrector in one pass (k_tvd_half3: each
direction's half state into the set its sweep reads) against three k_tvd_half launches
into one set (IMHD_TVDHALF3=0).  The per-direction arithmetic is shared
(tvd_half_calc), so FP64 agrees bit for bit over native-loop runs (Roe-TVD and HLLE,
both boundaries, passives, an ensemble of uneven boxes, the density-floor count
included); FP32 within rounding."""
import sys

import numpy as np
import pytest

from refutil import random_state

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(monkeypatch, half3, dims, bc, riemann="roe", precision=64, steps=5, npass=0):
    from isothermal_mhd_amd import TvdSolver
    monkeypatch.setenv("IMHD_TVDHALF3", half3)
    try:
        S = TvdSolver(dims, npass=npass, riemann=riemann, bc=bc, precision=precision)
    finally:
        monkeypatch.delenv("IMHD_TVDHALF3")
    for b, d in enumerate(S.boxes):
        q, bf, qp = random_state(*d, npass=npass, seed=70 + b)
        S.set_state(q, bf, qp if npass else None, box=b)
    S.boundary_conditions()
    S.run(steps, graph=False)
    out = [S.get_state(box=b) for b in range(len(S.boxes))], S.get_time(), S.floor_count()
    S.close()
    return out


CASES = [((16, 16, 16), "periodic", "roe", 2), ((13, 9, 11), "open", "roe", 0), ((20, 12, 14), "periodic", "hlle", 1),
         ([(16, 12, 10), (24, 16, 8), (10, 10, 40)], "periodic", "roe", 1)]


@pytest.mark.parametrize("case", CASES, ids=["single_roe_periodic", "single_roe_open", "single_hlle", "ensemble"])
def test_tvd_half3_bitwise_fp64(monkeypatch, case):
    dims, bc, riemann, npass = case
    say(f"[gpu] tvd half3 {case}")
    ref = run(monkeypatch, "0", dims, bc, riemann=riemann, npass=npass)
    got = run(monkeypatch, "1", dims, bc, riemann=riemann, npass=npass)
    for b in range(len(ref[0])):
        for x, y in zip(ref[0][b], got[0][b]):
            if x is not None:
                assert np.array_equal(x, y), f"box {b}"
    assert np.array_equal(ref[1][0], got[1][0]) and np.array_equal(ref[1][1], got[1][1])
    assert ref[2] == got[2]


def test_tvd_half3_fp32(monkeypatch):
    dims = (24, 20, 16)
    say("[gpu] tvd half3 fp32")
    ref = run(monkeypatch, "0", dims, "periodic", precision=32, npass=1)
    got = run(monkeypatch, "1", dims, "periodic", precision=32, npass=1)
    for x, y in zip(ref[0][0][:2], got[0][0][:2]):
        assert np.max(np.abs(x - y)) <= 1e-5 * max(1.0, np.max(np.abs(x)))
