"""The device-generated random-field start (imhd_init_turbulence) against an independent
numpy construction of the same field.

* Philox4x32-10 draws: a vectorised numpy Philox, itself checked here against the oracle's
  (which the Random123 known-answer vectors pin, tests/test_oracle_ou.py);
* velocity: the same modes placed in a full complex spectrum and transformed with
  numpy's inverse FFT -- a different evaluation from the device's separable sums;
* exact properties: the interior rms of v is the requested Mach number, the field is
  solenoidal when asked (k . v_k = 0 for the physical wave vectors n_d / n_cells,d), B is
  uniform in every cell and face, ghost cells are the periodic images;
* an FP32 context holds the FP64 context's values rounded once, bit for bit.
"""
import ctypes
import sys

import numpy as np
import pytest

from refutil import oracle_lib

NB = 6
M32 = np.uint64(0xFFFFFFFF)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def philox(c0, c1, c2, c3, k0, k1):
    """Philox4x32-10 on uint64 arrays holding 32-bit words"""
    c = [np.asarray(x, dtype=np.uint64) & M32 for x in (c0, c1, c2, c3)]
    c = np.broadcast_arrays(*c)
    c = [x.copy() for x in c]
    k0, k1 = np.uint64(k0), np.uint64(k1)
    for _ in range(10):
        p0 = np.uint64(0xD2511F53) * c[0]
        p1 = np.uint64(0xCD9E8D57) * c[2]
        y0 = ((p1 >> np.uint64(32)) ^ c[1] ^ k0) & M32
        y1 = p1 & M32
        y2 = ((p0 >> np.uint64(32)) ^ c[3] ^ k1) & M32
        y3 = p0 & M32
        c = [y0, y1, y2, y3]
        k0 = (k0 + np.uint64(0x9E3779B9)) & M32
        k1 = (k1 + np.uint64(0xBB67AE85)) & M32
    return c


def box_muller(a, b):
    s = 2.3283064365386963e-10
    u1 = (a.astype(np.float64) + 0.5) * s
    u2 = (b.astype(np.float64) + 0.5) * s
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)


def test_expected_state_rms_is_exact():
    """the construction itself: Parseval makes the grid rms the requested Mach number"""
    dims, p = CASES[0]
    rho, vel, ps, modes, vh = expected_state(dims, npass=0, **p)
    assert abs(np.sqrt(np.mean((vel ** 2).sum(axis=0))) - p["mach"]) <= 1e-12 * p["mach"]
    assert len(modes) > 100 and abs(rho.mean() - 1.0) < 0.01


def test_numpy_philox_equals_oracle():
    L = oracle_lib()
    rng = np.random.default_rng(1)
    for _ in range(20):
        ctr = rng.integers(0, 2 ** 32, 4, dtype=np.uint64)
        key = rng.integers(0, 2 ** 32, 2, dtype=np.uint64)
        c = (ctypes.c_uint * 4)(*[int(x) for x in ctr])
        k = (ctypes.c_uint * 2)(*[int(x) for x in key])
        o = (ctypes.c_uint * 4)()
        L.orc_philox4x32(c, k, o)
        mine = philox(*ctr, *key)
        assert [int(x) for x in mine] == list(o)


def expected_state(dims, rho0, sigma, b0, kmin, kmax, slope, mach, seed, solenoidal, npass):
    nx, ny, nz = dims
    K = int(np.floor(kmax))
    slo, shi = seed & 0xFFFFFFFF, seed >> 32
    modes = []
    for kz in range(K + 1):
        for ky in range(-K, K + 1):
            for kx in range(-K, K + 1):
                if not (kz > 0 or (kz == 0 and ky > 0) or (kz == 0 and ky == 0 and kx > 0)):
                    continue
                kn = np.sqrt(kx * kx + ky * ky + kz * kz)
                if kmin <= kn <= kmax:
                    modes.append((kx, ky, kz, kn))
    m = np.arange(len(modes), dtype=np.uint64)
    n = np.array([md[:3] for md in modes], dtype=np.float64)
    amp = np.array([md[3] for md in modes]) ** (-0.5 * (slope + 2.0))
    vh = np.zeros((len(modes), 3), dtype=np.complex128)
    for d in range(3):
        o = philox(m, d, 1, 0, slo, shi)
        vh[:, d] = amp * box_muller(o[0], o[1]) + 1j * amp * box_muller(o[2], o[3])
    if solenoidal:
        k = n / np.array(dims, dtype=np.float64)
        vh = vh - k * ((k * vh).sum(axis=1) / (k * k).sum(axis=1))[:, None]
    S = 2.0 * (np.abs(vh) ** 2).sum()
    vh *= mach / np.sqrt(S)
    F = np.zeros((3,) + dims, dtype=np.complex128)
    for (kx, ky, kz, _), v in zip(modes, vh):
        F[:, kx % nx, ky % ny, kz % nz] += v
        F[:, -kx % nx, -ky % ny, -kz % nz] += np.conj(v)
    vel = np.real(np.fft.ifftn(F, axes=(1, 2, 3))) * (nx * ny * nz)
    cell = (np.arange(nx)[:, None, None] + nx * (np.arange(ny)[None, :, None] +
                                                 ny * np.arange(nz)[None, None, :])).astype(np.uint64)
    if sigma:
        o = philox(cell & M32, cell >> np.uint64(32), 2, 0, slo, shi)
        rho = rho0 * np.exp(sigma * box_muller(o[0], o[1]) - 0.5 * sigma * sigma)
    else:
        rho = np.full(dims, rho0)
    ps = []
    for p in range(npass):
        o = philox(cell & M32, cell >> np.uint64(32), 3 + p, 0, slo, shi)
        ps.append(((o[0] << np.uint64(32) | o[1]) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53)
    return rho, vel, ps, modes, vh


CASES = [((24, 20, 16), dict(rho0=1.0, sigma=0.1, b0=(0.05, 0.05, 0.0), kmin=1.0, kmax=5.0, slope=5 / 3,
                             mach=3.0, seed=5005, solenoidal=False)),
         ((16, 16, 16), dict(rho0=2.0, sigma=0.0, b0=(0.0, 0.0, 1.0), kmin=1.0, kmax=4.0, slope=2.0, mach=2.0,
                             seed=(7 << 32) + 2002, solenoidal=True))]


@pytest.mark.gpu
@pytest.mark.parametrize("dims,p", CASES)
def test_device_start_matches_numpy(dims, p):
    from isothermal_mhd_amd import VanLeerSolver
    say(f"[gpu] init_turbulence {dims} {p}")
    npass = 2
    rho, vel, ps, modes, vh = expected_state(dims, npass=npass, **p)
    S = VanLeerSolver(dims, npass=npass, riemann="hlld", bc="periodic", precision=64)
    S.init_turbulence(rho0=p["rho0"], sigma_ln=p["sigma"], b0=p["b0"], kmin=p["kmin"], kmax=p["kmax"],
                      slope=p["slope"], mach=p["mach"], seed=p["seed"], solenoidal=p["solenoidal"], passive=True)
    q, bf, qp = S.get_state()
    I = tuple(slice(NB, NB + n) for n in dims)
    g_rho = q[I + (0,)]
    assert np.allclose(g_rho, rho, rtol=1e-14, atol=0), np.abs(g_rho / rho - 1).max()
    g_v = np.stack([q[I + (1 + d,)] / g_rho for d in range(3)])
    err = np.abs(g_v - vel).max() / np.abs(vel).max()
    say(f"  {len(modes)} modes, max |v - v_fft| / max|v| = {err:.2e}")
    assert err <= 1e-12
    ms = np.mean((g_v ** 2).sum(axis=0))
    assert abs(np.sqrt(ms) - p["mach"]) <= 1e-12 * p["mach"]
    if p["solenoidal"]:
        vk = np.fft.fftn(g_v, axes=(1, 2, 3))
        k = [np.fft.fftfreq(n) for n in dims]       # n_d / n_cells,d
        KX, KY, KZ = np.meshgrid(*k, indexing="ij")
        div = np.abs(KX * vk[0] + KY * vk[1] + KZ * vk[2]).max() / np.abs(vk).max()
        say(f"  max |k . v_k| / max |v_k| = {div:.2e}")
        assert div <= 1e-12
    for d in range(3):
        assert np.all(q[..., 4 + d] == p["b0"][d]) and np.all(bf[..., d] == p["b0"][d])
    Ip = tuple(slice(NB - 2, NB - 2 + n) for n in dims)
    for pp in range(npass):
        assert np.array_equal(qp[Ip + (pp,)], ps[pp])
    # ghosts: periodic images of the interior
    nx = dims[0]
    assert np.array_equal(q[:NB, NB:-NB, NB:-NB, :], q[nx:nx + NB, NB:-NB, NB:-NB, :])
    assert np.array_equal(q[-NB:, NB:-NB, NB:-NB, :], q[NB:2 * NB, NB:-NB, NB:-NB, :])
    # both state sets hold the start: one step from it equals one step from set_state
    T = VanLeerSolver(dims, npass=npass, riemann="hlld", bc="periodic", precision=64)
    T.set_state(q, bf, qp)
    for X in (S, T):
        X.run(2, graph=False)
    a, b = S.get_state(), T.get_state()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    # FP32 context: the FP64 values rounded once
    S32 = VanLeerSolver(dims, npass=npass, riemann="hlld", bc="periodic", precision=32)
    S32.init_turbulence(rho0=p["rho0"], sigma_ln=p["sigma"], b0=p["b0"], kmin=p["kmin"], kmax=p["kmax"],
                        slope=p["slope"], mach=p["mach"], seed=p["seed"], solenoidal=p["solenoidal"],
                        passive=True)
    q32, bf32, qp32 = S32.get_state()
    assert np.array_equal(q32, q.astype(np.float32).astype(np.float64))
    assert np.array_equal(bf32, bf)
    assert np.array_equal(qp32[Ip + (slice(0, npass),)],
                          qp[Ip + (slice(0, npass),)].astype(np.float32).astype(np.float64))
    for f in (0, 3, 8, 10):
        assert np.array_equal(S32.get_field(f), (q32[..., f] if f < 7 else bf32[..., f - 7] if f < 10
                                                  else qp32[..., f - 10]))
    for X in (S, T, S32):
        X.close()


@pytest.mark.gpu
def test_too_many_modes_rejected():
    from isothermal_mhd_amd import VanLeerSolver
    from isothermal_mhd_amd._lib import ImhdError
    S = VanLeerSolver((16, 16, 12), npass=0, riemann="hlld", bc="periodic", precision=64)
    with pytest.raises(ImhdError):
        S.init_turbulence(kmax=6.0)      # 2 * 6 >= 12
    S.init_turbulence(kmax=5.0)
    S.close()
