"""Device generation of the reference's random force (imhd_generate_force ==
generate_rand_force, turb_driver_mod.f90:197-391, with the unassigned modes zeroed)
against the oracle's restatement (oracle/imhd_oracle_force.inc, pinned in
tests/test_oracle_force.py).  The library forms the modes, twiddles and x pass on the
host with the same arithmetic and runs the y and z passes on the device without FMA
contraction: the field and the advanced taus88 state must equal the oracle's bit for bit.
A drive step with the generated field then matches the oracle driven with the same field
(the tolerance of the forcing application, tests/test_gpu_forcing.py)."""
import sys

import numpy as np
import pytest

from refutil import Oracle, oracle_rand_force, random_state

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [((16, 16, 16), 2, (1234, 5678, 9123)), ((12, 10, 8), 3, (7, 8, 9)),
         ((20, 12, 16), 4, (-1, 2147483647, 99)), ((32, 24, 40), 2, (5, 5, 5)),
         ((24, 12, 14), 5, (3, 1, 4))]


@pytest.mark.parametrize("case", CASES, ids=["x".join(map(str, c[0])) + f"_k{c[1]}" for c in CASES])
def test_generated_field_equals_oracle(case):
    from isothermal_mhd_amd import VanLeerSolver
    dims, K, seeds = case
    say(f"[gpu] generate_rand_force {dims} k_max={K}")
    delv_o, _, s_o = oracle_rand_force(*dims, K=K, seeds=seeds)
    for prec in (64, 32):
        S = VanLeerSolver(dims, npass=0, precision=prec)
        s_g = S.generate_rand_force(k_max=K, seeds=seeds)
        delv_g = S.get_force()
        S.close()
        assert s_g == s_o
        assert np.array_equal(delv_g, delv_o), f"precision {prec}: max diff {np.max(np.abs(delv_g - delv_o))}"


def test_invalid_k_max_rejected():
    from isothermal_mhd_amd import ImhdError, VanLeerSolver
    S = VanLeerSolver((12, 10, 8), npass=0)
    for K in (0, 4):
        with pytest.raises(ImhdError, match="IMHD_EINVAL"):
            S.generate_rand_force(k_max=K)
    S.close()


def test_drive_cycle_with_generated_force():
    """Two drive cycles: generate, step with drive_next, plain step, regenerate from the
    advanced seeds, drive again -- against the oracle fed the oracle's fields."""
    from isothermal_mhd_amd import VanLeerSolver
    dims, K, edot, dt_turb = (16, 16, 16), 2, 1e-3, 0.05
    say("[gpu] drive cycle with generated force")
    q, bf, qp = random_state(*dims, npass=0, seed=91)
    S = VanLeerSolver(dims, npass=0, riemann="hlld", recon="plm", bc="periodic", precision=64)
    S.set_state(q, bf, None)
    S.boundary_conditions()
    O = Oracle(*dims, npass=0, mhd=2, riemann=2, recon=0, bc=2)
    O.set_state(q, bf, None)
    O.bc()
    seeds = (1234, 5678, 9123)
    for cycle in range(2):
        delv_o, _, s_o = oracle_rand_force(*dims, K=K, seeds=seeds)
        seeds = S.generate_rand_force(k_max=K, edot=edot, dt_turb=dt_turb, seeds=seeds)
        assert seeds == s_o
        O.set_force(delv_o, edot, dt_turb)
        for dt, drive in ((1.5e-3, True), (2e-3, False)):
            S.vanleer_solve_3d(dt, drive_next=drive)
            O.step_drive(dt) if drive else O.step(dt)
            S.boundary_conditions()
            O.bc()
    qg, bg, _ = S.get_state()
    qo, bo, _ = O.get_state()[:3]
    S.close()
    for f in range(7):
        assert np.max(np.abs(qg[..., f] - qo[..., f])) <= 1e-12 * np.max(np.abs(qo[..., f])), f
    for d in range(3):
        assert np.max(np.abs(bg[..., d] - bo[..., d])) <= 1e-12 * np.max(np.abs(bo[..., d])), d
