"""Device spectra and density PDF (imhd_spectrum, imhd_density_pdf) against the analysis
program's restatement (oracle/spectra.py, pinned in tests/test_spectra.py).  Tolerance:
the device transform is hipFFT in FP32 and the oracle's numpy in FP64 (the reference uses
its own REAL*4 FFT), so each shell agrees to 2e-5 of the largest shell plus 1e-4 of its
own value; the bin rules are the same REAL*4 arithmetic, checked exactly by single-mode
fields.  The PDF's counts are exact."""
import os
import sys

import numpy as np
import pytest

from refutil import random_state

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import spectra as SP  # noqa: E402

pytestmark = pytest.mark.gpu
FIELDS = ["density", "velocity", "bfield", "vpar", "vperp"]


def solver(n, precision=64):
    from isothermal_mhd_amd import VanLeerSolver
    return VanLeerSolver((n, n, n), npass=0, precision=precision)


@pytest.mark.parametrize("precision", [64, 32])
def test_spectra_match_oracle(precision):
    n = 32
    print(f"[gpu] spectra n={n} fp{precision}", flush=True)
    q, bf, _ = random_state(n, n, n, npass=0, seed=17, amp_v=0.6)
    S = solver(n, precision)
    S.set_state(q, bf, None)
    qs = S.get_state()[0]          # the state as the context holds it (FP32 rounded)
    for which, name in enumerate(FIELDS):
        k, pk = S.spectrum(name)
        ko, po = SP.spectrum(qs, (n, n, n), which)
        assert np.array_equal(k, ko)
        tol = 2e-5 * np.max(np.abs(po)) + 1e-4 * np.abs(po)
        assert np.all(np.abs(pk - po) <= tol), name
    S.close()


def test_single_mode_bins_exact():
    n = 16
    x = (np.arange(n) + 0.5) / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    from test_spectra import cube
    S = solver(n)
    for kx0 in (1, 3, 5, 7):
        q = cube(n, 1.0 + 0.25 * np.cos(2 * np.pi * kx0 * X), b=(0 * X, 0 * X, 1 + 0 * X))
        bf = np.zeros(q.shape[:3] + (3,), order="F")
        bf[..., 2] = q[..., 6]
        S.set_state(q, bf, None)
        k, pk = S.spectrum("density")
        ik = int(2 * kx0 / np.sqrt(3.0))
        assert pk[ik] == pytest.approx(0.25 ** 2 * n ** 3 / 2, rel=1e-4)
        assert np.max(np.abs(np.delete(pk, [0, ik]))) <= 1e-5 * pk[ik]
    S.close()


def test_density_pdf_matches_oracle():
    n = 24
    q, bf, _ = random_state(n, n, n, npass=0, seed=4, amp_rho=0.8)
    S = solver(n, 32)
    S.set_state(q, bf, None)
    qs = S.get_state()[0]
    # 5000 bins: more than the per-workgroup LDS histogram holds (global counters)
    for nbins, rmax in ((n, 4.0), (50, 2.0), (5000, 3.0)):
        r, pdf = S.density_pdf(nbins, rmax)
        ro, po = SP.density_pdf(qs, (n, n, n), nbins, rmax)
        assert np.array_equal(r, ro) and np.array_equal(pdf, po)
    S.close()


def test_density_pdf_uniform_density():
    """Every cell in one bin (the uniform density of a fresh IC): the wave-aggregated
    counting must still count each cell once."""
    n = 48
    q, bf, _ = random_state(n, n, n, npass=0, seed=5)
    q[..., 0] = 1.0
    S = solver(n, 32)
    S.set_state(q, bf, None)
    for nbins, rmax in ((n, 4.0), (5000, 4.0)):
        r, pdf = S.density_pdf(nbins, rmax)
        ro, po = SP.density_pdf(S.get_state()[0], (n, n, n), nbins, rmax)
        assert np.array_equal(pdf, po)
        ib = int(1.0 / (rmax / nbins))
        assert np.count_nonzero(pdf) == 1 and pdf[ib] > 0
    S.close()


def test_spectrum_rejects_non_cubic():
    from isothermal_mhd_amd import ImhdError, VanLeerSolver
    S = VanLeerSolver((16, 16, 8), npass=0)
    q, bf, _ = random_state(16, 16, 8, npass=0, seed=1)
    S.set_state(q, bf, None)
    with pytest.raises(ImhdError, match="IMHD_EINVAL"):
        S.spectrum("velocity")
    S.close()


def test_mode_spectra_match_oracle():
    n = 32
    print("[gpu] mode spectra", flush=True)
    q, bf, _ = random_state(n, n, n, npass=0, seed=21, amp_v=0.6, b0=(0.5, 0.2, 0.1))
    S = solver(n)
    S.set_state(q, bf, None)
    k, P = S.mode_spectra()
    ko, Po = SP.mode_spectra(S.get_state()[0], (n, n, n))
    assert np.allclose(k, ko, rtol=0, atol=1e-12)
    for m in range(8):
        tol = 2e-5 * np.max(np.abs(Po[m])) + 1e-4 * np.abs(Po[m])
        assert np.all(np.abs(P[m] - Po[m]) <= tol), m
    S.close()


def test_mode_spectra_alfven_mode():
    """B0 along z, velocity along y varying in x: a pure Alfven mode (tests/test_spectra.py)."""
    from test_spectra import cube
    n, kx0, A = 16, 3, 0.2
    x = (np.arange(n) + 0.5) / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    q = cube(n, 1.0, v=(0 * X, A * np.sin(2 * np.pi * kx0 * X), 0 * X), b=(0 * X, 0 * X, 1.0 + 0 * X))
    bf = np.zeros(q.shape[:3] + (3,), order="F")
    bf[..., 2] = q[..., 6]
    S = solver(n)
    S.set_state(q, bf, None)
    k, P = S.mode_spectra()
    ik = int(2 * kx0 / np.sqrt(3.0))
    want = A ** 2 * n ** 3 / 2
    assert P[0, ik] == pytest.approx(want, rel=1e-4) and P[3, ik] == pytest.approx(want, rel=1e-4)
    for m in (1, 2, 4, 5, 6, 7):
        assert np.max(np.abs(P[m])) <= 1e-5 * want
    S.close()
