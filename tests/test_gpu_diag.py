"""Diagnostics and spectrum set-up on the GPU.

* imhd_diagnostics reduces in a fixed order (per-workgroup partial rows, then one
  fixed-order pass per box): two calls on one state return the same bits, every box of an
  ensemble returns the bits of the same box alone, and the values agree with the oracle's
  compute_KE / compute_divB restatement (driver.f90:492-553) to summation-order rounding,
  1e-12 relative.
* A spectrum set-up that fails after allocating releases everything it made, so the next
  call plans again and returns the spectrum of a fresh context.
"""
import sys

import numpy as np
import pytest

from refutil import Oracle, random_state

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


KEYS = ("ke", "vrms_sq", "momx", "momy", "momz", "divb_sum", "divb_max", "mass")


@pytest.mark.parametrize("prec", [64, 32])
def test_diagnostics_repeatable_and_vs_oracle(prec):
    from isothermal_mhd_amd import VanLeerSolver
    dims = (40, 24, 33)
    say(f"[gpu] diagnostics prec={prec} dims={dims}")
    q, bf, qp = random_state(*dims, npass=0, seed=61)
    S = VanLeerSolver(dims, npass=0, riemann="hlld", bc="periodic", precision=prec)
    S.set_state(q, bf)
    S.boundary_conditions()
    S.run(3, graph=False)
    a, b = S.diagnostics(), S.diagnostics()
    for k in KEYS:
        assert a[k] == b[k], k
    if prec == 64:
        O = Oracle(*dims, npass=0, mhd=2, riemann=2, recon=0, bc=2)
        qs, bs, _ = S.get_state()
        O.set_state(qs, bs)
        do = O.diagnostics()
        for k in ("ke", "vrms_sq", "momx", "momy", "momz", "mass"):
            assert abs(a[k] - do[k]) <= 1e-12 * max(1.0, abs(do[k])), (k, a[k], do[k])
        assert abs(a["divb_sum"] - do["divb_sum"]) <= 1e-12
        assert a["divb_max"] == do["divb_max"]
    S.close()


def test_diagnostics_ensemble_equals_single_boxes():
    from isothermal_mhd_amd import VanLeerSolver
    boxes = [(12, 10, 8), (40, 24, 33), (9, 14, 11)]
    say(f"[gpu] diagnostics ensemble {boxes}")
    states = [random_state(*d, npass=0, seed=70 + i) for i, d in enumerate(boxes)]
    E = VanLeerSolver(boxes, npass=0, riemann="hlle", bc="periodic", precision=64)
    for i, (q, bf, _) in enumerate(states):
        E.set_state(q, bf, None, box=i)
    de = E.diagnostics()
    for i, d in enumerate(boxes):
        S = VanLeerSolver(d, npass=0, riemann="hlle", bc="periodic", precision=64)
        S.set_state(states[i][0], states[i][1])
        ds = S.diagnostics()
        for k in KEYS:
            assert ds[k] == de[i][k], (i, k)
        S.close()
    E.close()


def test_spectrum_replans_after_failed_setup(monkeypatch):
    from isothermal_mhd_amd import VanLeerSolver
    n = 16
    say("[gpu] spectrum set-up failure, then re-plan")
    q, bf, qp = random_state(n, n, n, npass=0, seed=62)
    S = VanLeerSolver((n, n, n), npass=0, riemann="hlld", bc="periodic", precision=64)
    S.set_state(q, bf)
    b0 = S.device_bytes()
    monkeypatch.setenv("IMHD_TEST_FFT_SETUP_FAIL", "1")
    with pytest.raises(RuntimeError):
        S.spectrum("velocity")
    assert S.device_bytes() == b0          # everything made by the failed set-up is released
    monkeypatch.delenv("IMHD_TEST_FFT_SETUP_FAIL")
    k1, p1 = S.spectrum("velocity")
    assert S.device_bytes() > b0           # plan work area and buffers are counted
    T = VanLeerSolver((n, n, n), npass=0, riemann="hlld", bc="periodic", precision=64)
    T.set_state(q, bf)
    k2, p2 = T.spectrum("velocity")
    assert np.array_equal(k1, k2)
    assert np.allclose(p1, p2, rtol=1e-12, atol=0)
    S.close()
    T.close()
