"""The oracle against the reference's own outputs committed in tests/golden/
(written by tests/golden/make_golden.py from the compiled reference).  Runs wherever the
oracle builds; this is how the pins of tests/test_oracle_vs_reference.py travel to hosts
without /root/reference."""
import ctypes
import os

import numpy as np
import pytest

from golden_inputs import (DRIVE_CASE, DRIVE_DT_TURB, DRIVE_EDOT, STEP_CASES, STEP_DTS, TVD_FLUX_CASES,
                           TVD_FLUX_DX, drive_force, recon_inputs, riemann_inputs, step_case_params,
                           tvd_flux_inputs)
from refutil import Oracle, P, oracle_lib, random_state

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


@pytest.mark.parametrize("kind", [1, 2, 3])
def test_riemann_golden(gold, kind):
    A, B = riemann_inputs(kind)
    ni, nj = A.shape[:2]
    F = np.zeros((ni, nj, 7), order="F")
    s = ctypes.c_double()
    oracle_lib().orc_riemann_plane(kind, ni, nj, 1.0, P(A), P(B), P(F), ctypes.byref(s))
    assert np.array_equal(F, gold[f"riemann_{kind}"])
    assert s.value == float(gold[f"riemann_{kind}_smax"])


@pytest.mark.parametrize("kind", [1, 2])
def test_recon_golden(gold, kind):
    w, wb = recon_inputs(kind)
    ni, nj = w.shape[:2]
    qL = np.zeros((ni, nj, 7), order="F")
    qR = np.zeros((ni, nj, 7), order="F")
    oracle_lib().orc_recon_plane(kind, ni, nj, 6, 1.0, 1.01, P(w), P(wb), P(qL), P(qR))
    assert np.array_equal(qL[2:ni - 3], gold[f"recon_{kind}_L"])
    assert np.array_equal(qR[2:ni - 3], gold[f"recon_{kind}_R"])


@pytest.mark.parametrize("n", range(len(TVD_FLUX_CASES)))
def test_tvd_flux_golden(gold, n):
    kind, pred, dt = TVD_FLUX_CASES[n]
    A = tvd_flux_inputs()
    ni, nj = A.shape[:2]
    F = np.zeros((ni, nj, 7), order="F")
    oracle_lib().orc_tvd_flux_plane(kind, ni, nj, 1.0, dt, TVD_FLUX_DX, pred, P(A), P(F))
    assert np.array_equal(F, gold[f"tvdflux_{kind}_{pred}_{n}"])


@pytest.mark.parametrize("name", list(STEP_CASES))
def test_step_golden(gold, name):
    build, dims, npass, seed = STEP_CASES[name]
    mhd, rs, bc = step_case_params(name)
    O = Oracle(*dims, npass=npass, mhd=mhd, riemann=rs, recon=2, bc=bc)
    q, bf, qp = random_state(*dims, npass=npass, seed=seed)
    O.set_state(q, bf, qp)
    O.bc()
    for dt in STEP_DTS:
        if npass:
            O.copy_velocity(0)
        O.step(dt)
        if npass:
            O.copy_velocity(1)
            O.passive_step(dt)
        O.bc()
    qo, bo, po, _ = O.get_state()
    I = (slice(6, -6),) * 3
    Ip = (slice(4, -4),) * 3
    assert np.array_equal(qo[I], gold[f"step_{name}_q"])
    assert np.array_equal(bo[I], gold[f"step_{name}_b"])
    assert np.array_equal(po[Ip][..., :max(npass, 1)], gold[f"step_{name}_p"])
    assert O.smax() == float(gold[f"step_{name}_smax"])
    assert O.max_speed() == float(gold[f"step_{name}_cmax"])


def test_drive_golden(gold):
    """forcing application (get_rho_turb + add_random_force) against the reference's output"""
    build, dims, npass, seed, fseed, dts, drives = DRIVE_CASE
    O = Oracle(*dims, npass=npass, mhd=2, riemann=2, recon=2, bc=2)
    q, bf, qp = random_state(*dims, npass=npass, seed=seed)
    O.set_state(q, bf, qp)
    O.bc()
    O.set_force(drive_force(O.sh, fseed), DRIVE_EDOT, DRIVE_DT_TURB)
    for dt, drive in zip(dts, drives):
        O.copy_velocity(0)
        O.step_drive(dt) if drive else O.step(dt)
        O.copy_velocity(1)
        O.passive_step(dt)
        O.bc()
    qo, bo, po, _ = O.get_state()
    assert np.array_equal(qo, gold["drive_q"])
    assert np.array_equal(bo, gold["drive_b"])
    assert np.array_equal(po, gold["drive_p"])
    assert O.max_speed() == float(gold["drive_cmax"])
