"""Pins of the oracle's restatement of the OU forcing (oracle/imhd_oracle_ou.inc), which
the GPU is checked against in tests/test_gpu_ou.py.  The reference has no OU forcing, so
each piece is pinned by what the mathematics fixes:
  * Philox4x32-10 against the published known-answer vectors (Random123 kat_vectors);
  * the OU amplitudes: stationary variance (4 sigma^2 per solenoidal mode: two transverse
    directions x real and imaginary parts) and autocorrelation exp(-lag dt / T);
  * solenoidality k . a = 0 of every mode (and not without the projection);
  * the acceleration field is a real field whose discrete Fourier transform has energy only
    at the chosen +/- wave vectors, each with a 2 Re(a e^{ikx}) amplitude of a;
  * the kick injects no net momentum.
A plausible mistake fails one of them: a wrong Philox constant or round, a dropped sqrt(1-f^2)
(wrong variance), f = exp(-dt/T) vs exp(-2dt/T) (wrong autocorrelation), a missing projection,
a sign or index slip in the twiddles (energy at the wrong wave vector), a missing mean.
"""
import numpy as np
import pytest

from refutil import Oracle, P, oracle_lib, random_state


def philox(ctr, key):
    c = np.array(ctr, dtype=np.uint32)
    k = np.array(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    oracle_lib().orc_philox4x32(P(c), P(k), P(o))
    return [int(x) for x in o]


def test_philox_known_answers():
    assert philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_mode_set_shell_1_2():
    O = Oracle(16, 16, 16)
    n = O.ou_setup(kmin=1.0, kmax=2.0)
    k, _ = O.ou_modes()
    mag2 = (k ** 2).sum(1)
    # |n|^2 in {1, 2, 3, 4}: 6 + 12 + 8 + 6 = 32 vectors, one of each +/- pair
    assert n == 16 and sorted(set(mag2.tolist())) == [1, 2, 3, 4]
    assert len({tuple(v) for v in k} | {tuple(-v) for v in k}) == 32


def test_ou_stationary_variance_and_autocorrelation():
    sigma, T, dt, nsteps, lag = 0.7, 0.5, 0.05, 20000, 10
    O = Oracle(16, 16, 16)
    nm = O.ou_setup(sigma=sigma, t_corr=T, seed=11)
    hist = np.zeros((nsteps, nm, 3), dtype=complex)
    for s in range(nsteps):
        O.ou_modes_step(dt)
        hist[s] = O.ou_modes()[1]
    power = (np.abs(hist) ** 2).sum(2)                      # per mode: |a|^2 over 3 components
    assert abs(power.mean() / (4 * sigma ** 2) - 1.0) < 0.05
    x = hist.real.reshape(nsteps, -1)
    c0 = (x * x).mean()
    cl = (x[:-lag] * x[lag:]).mean()
    assert abs(cl / c0 - np.exp(-lag * dt / T)) < 0.03


@pytest.mark.parametrize("sol", [True, False])
def test_solenoidal_projection(sol):
    O = Oracle(16, 16, 16)
    O.ou_setup(solenoidal=sol, seed=5)
    for _ in range(3):
        O.ou_modes_step(0.05)
    k, a = O.ou_modes()
    dots = np.abs((k * a).sum(1))
    if sol:
        assert dots.max() <= 1e-12 * np.abs(a).max()
    else:
        assert dots.max() > 1e-3 * np.abs(a).max()


def test_field_is_the_discrete_fourier_sum_of_the_modes():
    n = 16
    O = Oracle(n, n, n)
    O.ou_setup(seed=7)
    O.ou_modes_step(0.05)
    k, a = O.ou_modes()
    f = O.ou_field()                                        # (n, n, n, 3), cell centres (i - 1/2)/n
    # e^{i k x_i} with x_i = (i - 1/2) / n: remove the half-cell phase after the FFT
    F = np.fft.fftn(f, axes=(0, 1, 2)) / n ** 3
    idx = np.fft.fftfreq(n, 1.0 / n).astype(int)
    expected = np.zeros_like(F)
    for (nx, ny, nz), amp in zip(k, a):
        ph = np.exp(1j * np.pi * (nx + ny + nz) / n)      # e^{i 2 pi n (i - 1/2)/N}, i from 1
        for sgn, v in ((1, amp * ph), (-1, np.conj(amp * ph))):
            ix = np.where(idx == sgn * nx)[0][0]
            iy = np.where(idx == sgn * ny)[0][0]
            iz = np.where(idx == sgn * nz)[0][0]
            expected[ix, iy, iz] += v
    assert np.abs(F - expected).max() <= 1e-12 * np.abs(expected).max()


def test_kick_injects_no_net_momentum():
    dims = (16, 12, 10)
    O = Oracle(*dims, npass=0)
    q, bf, _ = random_state(*dims, npass=0, seed=3)
    O.set_state(q, bf)
    O.ou_setup(sigma=2.0, seed=9)
    q0 = O.get_state()[0]
    O.ou_step(0.01)
    q1 = O.get_state()[0]
    I = (slice(6, -6),) * 3
    dm = (q1 - q0)[I][..., 1:4]
    assert np.abs(dm).max() > 1e-4
    assert np.abs(dm.sum((0, 1, 2))).max() <= 1e-12 * np.abs(dm).sum()
    assert np.array_equal(q1[..., 0], q0[..., 0]) and np.array_equal(q1[..., 4:], q0[..., 4:])
