"""The oracle's restatement of the reference's random force (generate_rand_force,
turb_driver_mod.f90:197-391, with its unassigned modes zeroed).

Pinned against the reference where it is deterministic -- the taus88 stream and the 1-D
direct inverse DFT, bit for bit, from golden vectors (tests/golden/force_vectors.json) and
the live reference build when present -- and, for the whole field, against what the
mathematics fixes: the field equals an independent evaluation of its Fourier sum, is real,
its DFT holds exactly the octant modes and their mirrors, axis-plane modes carry their own
draws' Box-Muller amplitudes and phases, and off-axis modes are normal to k.
"""
import json
import os

import numpy as np
import pytest

from refutil import Reference, oracle_lib, oracle_rand_force, oracle_taus88, ref_available, P

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "force_vectors.json")))
NB = 6


def test_taus88_matches_reference_stream():
    for g in GOLD["taus"]:
        got, _ = oracle_taus88(g["seeds"], len(g["values"]))
        assert np.array_equal(got, np.array(g["values"]))
    if ref_available("vl_hlld_p16_drive"):
        R = Reference("vl_hlld_p16_drive")
        for s in [(1, 2, 3), (2147483647, -2147483648, 77), (31337, 4, 100000)]:
            assert np.array_equal(oracle_taus88(s, 5000)[0], R.rand_taus(s, 5000))


def test_idft_1d_matches_reference():
    import ctypes
    L = oracle_lib()
    for g in GOLD["idft"]:
        m = g["ihi"] - g["ilow"] + 1
        a, b = np.array(g["in_re"]), np.array(g["in_im"])
        ore, oim = np.zeros(m), np.zeros(m)
        L.orc_idft_1d(P(a), P(b), P(ore), P(oim), g["ilow"], g["ihi"], g["nnx"], g["k_max"])
        assert np.array_equal(ore, np.array(g["out_re"])) and np.array_equal(oim, np.array(g["out_im"]))
    if ref_available("vl_hlld_p16_drive"):
        R = Reference("vl_hlld_p16_drive")
        rng = np.random.default_rng(3)
        for _ in range(5):
            a, b = rng.standard_normal(4) * 1e3, rng.standard_normal(4)
            ore, oim = np.zeros(28), np.zeros(28)
            L.orc_idft_1d(P(a), P(b), P(ore), P(oim), -5, 22, 16, 2)
            rr, ri = R.idft_1d(a, b, -5, 22, 16)
            assert np.array_equal(ore, rr) and np.array_equal(oim, ri)


def fourier_sum(modes, n, K, nb=NB):
    """Independent evaluation: f(x) = (1/N) sum_k c_k exp(+2 pi i k.x / n), x = global cell
    index 1-nb .. n+nb per axis, over the summed wave numbers |k_d| <= K-1."""
    ax = [np.arange(1 - nb, n[d] + nb + 1) for d in range(3)]
    ks = np.arange(-(K - 1), K)
    E = [np.exp(2j * np.pi * np.outer(ax[d], ks) / n[d]) for d in range(3)]   # (P_d, 2K-1)
    sub = modes[:, 1:2 * K, 1:2 * K, 1:2 * K]                                  # [d, kz, ky, kx]
    f = np.einsum("dzyx,ix,jy,kz->ijkd", sub, E[0], E[1], E[2])
    return f / (n[0] * n[1] * n[2])


@pytest.mark.parametrize("n,K", [((16, 16, 16), 2), ((12, 10, 8), 3), ((20, 12, 16), 4)])
def test_field_is_the_real_fourier_sum_of_its_modes(n, K):
    delv, modes, _ = oracle_rand_force(*n, K=K, seeds=(11, 22, 33))
    f = fourier_sum(modes, n, K)
    scale = np.max(np.abs(delv))
    assert scale > 0
    assert np.max(np.abs(f.imag)) <= 1e-13 * scale           # Hermitian mirror -> real field
    assert np.max(np.abs(f.real - delv)) <= 1e-13 * scale      # passes, signs, offsets, 1/N


@pytest.mark.parametrize("n,K", [((16, 16, 16), 2), ((12, 10, 8), 3)])
def test_modes_and_their_draws(n, K):
    seeds = (1234, 5678, 9123)
    delv, modes, s_out = oracle_rand_force(*n, K=K, seeds=seeds)
    # the state advances by six draws per octant mode
    draws, s_end = oracle_taus88(seeds, 6 * K ** 3)
    assert s_out == s_end
    # DFT of the periodic interior: coefficient k = c_k e^{+2 pi i k.(1,1,1)/n}
    interior = delv[NB:NB + n[0], NB:NB + n[1], NB:NB + n[2], :]
    F = np.stack([np.fft.fftn(interior[..., d]) for d in range(3)])          # [d, kx, ky, kz]
    kk = [np.fft.fftfreq(n[d], 1.0 / n[d]).astype(int) for d in range(3)]
    c = F * np.exp(-2j * np.pi * (kk[0][:, None, None] / n[0] + kk[1][None, :, None] / n[1] +
                                  kk[2][None, None, :] / n[2]))[None]
    scale = np.max(np.abs(c))
    twopi = 8.0 * np.arctan(1.0)
    L = [n[d] / max(n) for d in range(3)]
    k0 = [twopi / L[d] for d in range(3)]
    k_pk = twopi / max(L)
    seen = np.zeros(c.shape[1:], bool)
    m = 0
    for kz in range(K):
        for ky in range(K):
            for kx in range(K):
                rn = draws[6 * m:6 * m + 6]
                m += 1
                for sgn in (1, -1):
                    idx = (sgn * kx % n[0], sgn * ky % n[1], sgn * kz % n[2])
                    seen[idx] = True
                    ck = c[(slice(None),) + idx]
                    if sgn == -1:
                        ck = np.conj(ck)
                    kvec = np.array([kx * k0[0], ky * k0[1], kz * k0[2]])
                    kmag = np.sqrt(np.sum(kvec ** 2))
                    Pk = kmag ** 6 * np.exp(-8.0 * kmag / k_pk)
                    if kx == 0 and ky == 0 and kz == 0:
                        assert np.max(np.abs(ck)) <= 1e-12 * scale
                    elif min(kx, ky, kz) == 0:
                        # axis plane: no projection, the draws themselves
                        amp = np.sqrt(-2.0 * Pk * np.log(rn[:3]))
                        assert np.allclose(np.abs(ck), amp, rtol=1e-10, atol=1e-12 * scale)
                        big = amp > 1e-6 * scale
                        ph = np.angle(ck[big]) % twopi
                        assert np.allclose(ph, (twopi * rn[3:][big]) % twopi, atol=1e-9)
                    else:
                        # solenoidal: normal to k, and no longer than the draw
                        assert abs(np.dot(kvec, ck)) <= 1e-10 * kmag * max(np.max(np.abs(ck)), 1e-300)
                        amp = np.sqrt(-2.0 * Pk * np.log(rn[:3]))
                        assert np.linalg.norm(ck) <= np.linalg.norm(amp) * (1 + 1e-12)
    # every other wave vector is empty (the zeroed modes)
    assert np.max(np.abs(c[:, ~seen])) <= 1e-12 * scale


def test_planted_mirror_sign_error_is_caught():
    """The reality check fails if the mirror copies the imaginary part unconjugated (the
    mistake is planted in a copy of the oracle's modes, on the CPU)."""
    n, K = (16, 16, 16), 2
    _, modes, _ = oracle_rand_force(*n, K=K, seeds=(5, 6, 7))
    bad = modes.copy()
    W = 2 * K + 1
    for kz in range(K):
        for ky in range(K):
            for kx in range(K):
                if kx + ky + kz:
                    bad[:, K - kz, K - ky, K - kx] = modes[:, K + kz, K + ky, K + kx]
    f = fourier_sum(bad, n, K)
    assert np.max(np.abs(f.imag)) > 1e-3 * np.max(np.abs(f.real))
    assert W == 5
