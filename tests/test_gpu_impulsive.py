"""The reference driver's impulsive turbulence forcing inside the native loop
(imhd_set_impulsive_forcing + imhd_run; driver.f90:57-70, 102-111, 336-344) against the
oracle driven on the same schedule (refutil.oracle_driven).  The kick count and the
taus88 state match exactly (the fields are drawn bit for bit, tests/test_gpu_force_gen.py);
state, clock and next drive time to the forcing application's tolerance, 1e-12 relative
(tests/test_gpu_forcing.py: the normalisation sums are reduced in a different order)."""
import sys

import numpy as np
import pytest

from refutil import oracle_driven, random_state

pytestmark = pytest.mark.gpu

DIMS = (16, 12, 10)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_native(recon, precision=64, graph=False, nsteps=14, dt_turb=0.02, seed=63, riemann="hlld"):
    from isothermal_mhd_amd import VanLeerSolver
    q, bf, qp = random_state(*DIMS, npass=1, seed=seed)
    S = VanLeerSolver(DIMS, npass=1, riemann=riemann, recon=recon, bc="periodic", precision=precision, cour=0.35)
    S.set_state(q, bf, qp)
    S.boundary_conditions()
    S.set_time(t=0.0, dt=1.0e-15)
    if dt_turb is not None:
        S.set_impulsive_forcing(k_max=2, edot=1e-3, dt_turb=dt_turb, seeds=(1234, 5678, 9123))
    S.run(nsteps, graph=graph)
    state = S.get_state()
    t, dt = S.get_time()
    imp = S.impulsive_forcing_state() if dt_turb is not None else None
    S.close()
    return (q, bf, qp), state, float(t[0]), float(dt[0]), imp


@pytest.mark.parametrize("recon,orecon", [("plm", 0), ("ppm_char", 2)])
def test_native_impulsive_forcing_matches_oracle(recon, orecon):
    say(f"[gpu] native impulsive forcing {recon}")
    (q, bf, qp), (qg, bg, pg), tg, dtg, (sg, ttg) = run_native(recon)
    (qo, bo, po), to, dto, tto, so, kicks = oracle_driven(DIMS, q, bf, qp, 14, riemann=2, recon=orecon)
    assert kicks >= 3, kicks
    assert sg == so
    for g, o in ((tg, to), (dtg, dto), (ttg, tto)):
        assert abs(g - o) <= 1e-12 * abs(o), (g, o)
    for f in range(7):
        assert np.max(np.abs(qg[..., f] - qo[..., f])) <= 1e-12 * np.max(np.abs(qo[..., f])), f
    for d in range(3):
        assert np.max(np.abs(bg[..., d] - bo[..., d])) <= 1e-12 * np.max(np.abs(bo[..., d])), d
    assert np.max(np.abs(pg[..., 0] - po[..., 0])) <= 1e-12 * np.max(np.abs(po[..., 0]))


@pytest.mark.parametrize("precision", [64, 32])
def test_native_impulsive_forcing_graph_equals_eager(precision):
    """Non-kick steps replay the captured step graph, kick steps run eagerly: the result is
    the eager loop's, bit for bit."""
    say(f"[gpu] native impulsive forcing graph vs eager fp{precision}")
    _, ea, te, dte, ie = run_native("plm", precision=precision, graph=False)
    _, ga, tg, dtg, ig = run_native("plm", precision=precision, graph=True)
    for x, y in zip(ea, ga):
        assert np.array_equal(x, y)
    assert te == tg and dte == dtg and ie == ig


def test_native_impulsive_forcing_without_kick_is_the_plain_loop():
    """A drive time beyond the run leaves every step the plain native step."""
    say("[gpu] native impulsive forcing, no kick")
    _, a, ta, dta, _ = run_native("plm", dt_turb=1e9)
    _, b, tb, dtb, _ = run_native("plm", dt_turb=None)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert ta == tb and dta == dtb


def test_native_impulsive_forcing_rejects():
    from isothermal_mhd_amd import ImhdError, TvdSolver, VanLeerSolver
    q, bf, qp = random_state(*DIMS, npass=1, seed=63)
    S = VanLeerSolver(DIMS, npass=1, riemann="hlld", recon="plm", bc="periodic")
    S.set_state(q, bf, qp)
    with pytest.raises(ImhdError, match="IMHD_EINVAL"):
        S.set_impulsive_forcing(k_max=5)                      # k_max < min(n)/2 = 5
    with pytest.raises(ImhdError, match="IMHD_EINVAL"):
        S.set_impulsive_forcing(dt_turb=0.0)
    S.set_ou_forcing(sigma=1.0)
    with pytest.raises(ImhdError, match="IMHD_EINVAL"):
        S.set_impulsive_forcing()
    S.set_ou_forcing(sigma=0.0)
    S.set_impulsive_forcing()
    with pytest.raises(ImhdError, match="IMHD_EINVAL"):
        S.set_ou_forcing(sigma=1.0)
    S.set_impulsive_forcing(k_max=0)                         # off
    S.close()
    T = TvdSolver(DIMS, npass=1, bc="periodic")
    T.set_state(q, bf, qp)
    with pytest.raises(ImhdError, match="IMHD_EINVAL"):
        T.set_impulsive_forcing()
    T.close()
