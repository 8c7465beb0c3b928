"""Closed-form pins of the oracle where the reference build cannot reach: the primitive
PLM branch (vanleer_mod.f90:1160-1215; unreachable because vanleer_mod.f90:83 fixes
piecewise_parabolic = .TRUE.), and whole-step invariants of the PLM integrator.

Each check is chosen so that a plausible mistake fails it:
  * linear data reproduce the exact interface values (a wrong index or a dropped 1/2 fails);
  * the MC-beta cap with beta = 1.01 (minmod or MC beta = 2 give different numbers);
  * the slope is NOT zeroed at an extremum (a "textbook" limiter gives 1.0, not 1.125);
  * mirror symmetry of left/right states;
  * uniform states stay exactly uniform through a whole step; mass and momentum are
    conserved to roundoff in a periodic box; raw div B stays at roundoff (CT).
"""
import ctypes

import numpy as np

from refutil import Oracle, P, oracle_lib, random_state


def recon(w, wb, kind=0, beta=1.01):
    """w: (ni, nj, 6) primitive cells along axis 0; returns qL, qR (ni, nj, 7)."""
    ni, nj = w.shape[:2]
    w = np.asfortranarray(w, dtype=np.float64)
    wb = np.asfortranarray(wb, dtype=np.float64)
    qL = np.zeros((ni, nj, 7), order="F")
    qR = np.zeros((ni, nj, 7), order="F")
    oracle_lib().orc_recon_plane(kind, ni, nj, 6, 1.0, beta, P(w), P(wb), P(qL), P(qR))
    return qL, qR


def column(vals):
    """one pencil with the same profile in every primitive slot; rho column kept positive"""
    v = np.asarray(vals, dtype=np.float64)
    w = np.zeros((len(v), 1, 6))
    for m in range(6):
        w[:, 0, m] = v
    w[:, 0, 0] = 2.0 + v
    return w


def test_plm_linear_exact():
    v = 0.25 * np.arange(10)                  # dyadic: exact in binary
    w = column(v)
    qL, qR = recon(w, np.zeros((10, 1)))
    for i in range(1, 7):                      # interfaces i + 1/2 with a full stencil
        face = v[i] + 0.125                    # midpoint of the linear profile
        assert qL[i, 0, 5] == face and qR[i, 0, 5] == face
        assert qL[i, 0, 0] == 2.0 + face
        assert qL[i, 0, 1] == face * (2.0 + face)   # momentum = rho * v at the face


def test_plm_mc_beta_cap():
    w = column([0.0, 0.0, 1.0, 3.0, 3.0, 3.0])
    qL, _ = recon(w, np.zeros((6, 1)))
    # cell 2: dL = 1, dR = 2, dC = 1.5 -> slope min(1.01*1, 1.01*2, 1.5) = 1.01
    assert qL[2, 0, 5] == 1.0 + 0.5 * 1.01


def test_plm_no_zeroing_at_extremum():
    w = column([0.0, 0.0, 1.0, 0.5, 0.5, 0.5])
    qL, _ = recon(w, np.zeros((6, 1)))
    # cell 2 is a maximum: dL = 1, dR = -0.5, dC = 0.25 -> slope 0.25 (not 0)
    assert qL[2, 0, 5] == 1.0 + 0.125


def test_plm_mirror_symmetry():
    rng = np.random.default_rng(0)
    v = rng.standard_normal(12)
    w = column(v)
    wm = column(v[::-1])
    qL, qR = recon(w, np.zeros((12, 1)))
    qLm, qRm = recon(wm, np.zeros((12, 1)))
    for i in range(1, 9):
        j = 12 - 2 - i                         # mirrored interface
        assert np.isclose(qL[i, 0, 5], qRm[j, 0, 5], rtol=0, atol=1e-15)
        assert np.isclose(qR[i, 0, 5], qLm[j, 0, 5], rtol=0, atol=1e-15)


def test_plm_step_uniform_state_exact():
    n = 8
    q, bf, qp = random_state(n, n, n, npass=1, seed=0, amp_v=0.0, amp_rho=0.0, amp_b=0.0,
                             b0=(0.3, -0.2, 0.7))
    I = (slice(6, -6),) * 3
    q[I + (1,)] = 0.4
    q[I + (2,)] = -0.1
    q[I + (3,)] = 0.25
    O = Oracle(n, n, n, npass=1, mhd=2, riemann=2, recon=0, bc=2)
    O.set_state(q, bf, qp)
    O.bc()
    q0 = O.get_state()[0].copy()
    for dt in (1e-3, 2e-3):
        O.copy_velocity(0)
        O.step(dt)
        O.copy_velocity(1)
        O.passive_step(dt)
        O.bc()
    q1 = O.get_state()[0]
    assert np.array_equal(q0[I], q1[I])


def test_plm_step_conservation_and_divb():
    n = 12
    q, bf, qp = random_state(n, n, n, npass=1, seed=4)
    for rs in (1, 2, 3):
        O = Oracle(n, n, n, npass=1, mhd=2, riemann=rs, recon=0, bc=2)
        O.set_state(q, bf, qp)
        O.bc()
        d0 = O.diagnostics()
        for _ in range(5):
            O.copy_velocity(0)
            O.step(2e-3)
            O.copy_velocity(1)
            O.passive_step(2e-3)
            O.bc()
        d1 = O.diagnostics()
        assert abs(d1["mass"] - d0["mass"]) <= 1e-14 * d0["mass"]
        for k in ("momx", "momy", "momz"):
            assert abs(d1[k] - d0[k]) <= 1e-14
        assert d1["divb_max"] <= 1e-14
        assert O.floor_count() == 0


def test_riemann_consistency_and_supersonic_limits():
    """F(u, u) equals the physical flux; a supersonic left state returns F(uL)."""
    L = oracle_lib()
    rng = np.random.default_rng(1)
    n = 64
    u = np.zeros((n, 7))
    u[:, 0] = 0.5 + rng.random(n)
    u[:, 1:4] = u[:, [0]] * rng.normal(0, 1, (n, 3))
    u[:, 4:7] = rng.normal(0, 1, (n, 3))
    rho, vx, vy, vz = u[:, 0], u[:, 1] / u[:, 0], u[:, 2] / u[:, 0], u[:, 3] / u[:, 0]
    Bx, By, Bz = u[:, 4], u[:, 5], u[:, 6]
    Fex = np.stack([rho * vx, rho * vx * vx + rho + 0.5 * (By ** 2 + Bz ** 2 - Bx ** 2),
                    rho * vx * vy - Bx * By, rho * vx * vz - Bx * Bz, 0 * rho,
                    vx * By - vy * Bx, vx * Bz - vz * Bx], axis=1)
    A = np.asfortranarray(u.reshape(n, 1, 7))
    for kind in (1, 2, 3):
        F = np.zeros((n, 1, 7), order="F")
        sm = ctypes.c_double()
        L.orc_riemann_plane(kind, n, 1, 1.0, P(A), P(A), P(F), ctypes.byref(sm))
        assert np.allclose(F[:, 0, :], Fex, rtol=1e-12, atol=1e-12), kind
    # supersonic to the right: HLLD and HLLE return the left flux exactly
    us = u.copy()
    us[:, 1] = us[:, 0] * 20.0
    B = np.asfortranarray(us.reshape(n, 1, 7))
    C = np.asfortranarray((us * [1, 1.01, 1, 1, 1, 1, 1]).reshape(n, 1, 7))
    for kind in (2, 3):
        F = np.zeros((n, 1, 7), order="F")
        sm = ctypes.c_double()
        L.orc_riemann_plane(kind, n, 1, 1.0, P(B), P(C), P(F), ctypes.byref(sm))
        rho, vx = us[:, 0], us[:, 1] / us[:, 0]
        assert np.allclose(F[:, 0, 0], rho * vx, rtol=1e-14)
