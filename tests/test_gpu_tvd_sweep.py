"""The fused Roe-TVD sweep (eigen-system and flux in one launch per direction and stage:
k_tvd_sweep_x with g~_k in LDS, k_tvd_sweep_m marching y / z with a register window)
against the two-kernel path it replaces (k_tvd_eig -> k_tvd_flux, IMHD_TVDFUSE=0).
FP64 must agree bit for bit over native-loop runs, including ensembles of uneven boxes,
open boundaries and several x segments per row; FP32 within rounding."""
import sys

import numpy as np
import pytest

from refutil import random_state

pytestmark = pytest.mark.gpu

CASES = [((16, 16, 16), "periodic"), ((13, 9, 11), "open"), ((300, 12, 20), "periodic"),
         ((20, 24, 130), "periodic")]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(monkeypatch, fuse, dims, bc, precision=64, steps=4, npass=0):
    from isothermal_mhd_amd import TvdSolver
    monkeypatch.setenv("IMHD_TVDFUSE", fuse)
    try:
        S = TvdSolver(dims, npass=npass, riemann="roe", bc=bc, precision=precision)
    finally:
        monkeypatch.delenv("IMHD_TVDFUSE")
    boxes = S.boxes
    for b, d in enumerate(boxes):
        q, bf, qp = random_state(*d, npass=npass, seed=60 + b)
        S.set_state(q, bf, qp if npass else None, box=b)
    S.boundary_conditions()
    S.run(steps, graph=False)
    out = [S.get_state(box=b) for b in range(len(boxes))], S.get_time()
    S.close()
    return out


@pytest.mark.parametrize("case", CASES, ids=["x".join(map(str, c[0])) + "_" + c[1] for c in CASES])
def test_fused_tvd_sweep_bitwise_fp64(monkeypatch, case):
    dims, bc = case
    say(f"[gpu] fused tvd sweep {case}")
    ref = run(monkeypatch, "0", dims, bc)
    got = run(monkeypatch, "1", dims, bc)
    for x, y in zip(ref[0][0], got[0][0]):
        assert np.array_equal(x, y)
    assert np.array_equal(ref[1][0], got[1][0]) and np.array_equal(ref[1][1], got[1][1])


def test_fused_tvd_sweep_ensemble_fp64(monkeypatch):
    dims = [(16, 12, 10), (24, 16, 8), (10, 10, 40)]
    say("[gpu] fused tvd sweep ensemble")
    ref = run(monkeypatch, "0", dims, "periodic", npass=1)
    got = run(monkeypatch, "1", dims, "periodic", npass=1)
    for b in range(len(dims)):
        for x, y in zip(ref[0][b], got[0][b]):
            assert np.array_equal(x, y), f"box {b}"


def test_fused_tvd_sweep_fp32(monkeypatch):
    dims = (24, 20, 16)
    say("[gpu] fused tvd sweep fp32")
    ref = run(monkeypatch, "0", dims, "periodic", precision=32)
    got = run(monkeypatch, "1", dims, "periodic", precision=32)
    for x, y in zip(ref[0][0][:2], got[0][0][:2]):
        assert np.max(np.abs(x - y)) <= 1e-5 * max(1.0, np.max(np.abs(x)))
