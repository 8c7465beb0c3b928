"""The oracle's flux protection (update_hydro_state_3d with flux_protection,
vanleer_mod.f90:2080-2207; Roe with HLLE protection fluxes and corner EMFs).

The reference fixes flux_protection = .FALSE. as a module PARAMETER (vanleer_mod.f90:81),
so no reference build can run this branch; it is pinned by what its definition fixes:
* no cell below the threshold: the step equals the unprotected Roe step bit for bit;
* every cell flagged (threshold above every density): every face flux and corner EMF the
  update reads is replaced, so the step equals a pure HLLE step bit for bit -- this checks
  the replaced face and corner index sets and the recompute from the backup;
* a strong rarefaction that drives the unprotected Roe update below the density floor
  (54 floor events in 3 steps) runs with no floor event and positive density.
The partial-flag case (some cells flagged) has no closed form: "parity unpinned" against
the reference; the GPU matches this restatement bit for bit (tests/test_gpu_protection.py).
"""
import numpy as np

from refutil import Oracle, random_state

NB = 6


def rarefaction(n, vx0, seed=0):
    nx, ny, nz = n
    q, bf, qp = random_state(nx, ny, nz, npass=0, seed=seed, amp_v=0.1)
    I = (slice(NB, NB + nx), slice(NB, NB + ny), slice(NB, NB + nz))
    x = (np.arange(nx) + 0.5) / nx
    v = np.where(x < 0.5, -vx0, vx0)[:, None, None] * np.ones((nx, ny, nz))
    q[I + (1,)] = q[I + (0,)] * v
    return q, bf


def run(dims, q, bf, steps, dt, **kw):
    O = Oracle(*dims, npass=0, mhd=2, recon=0, bc=2, **kw)
    O.set_state(q, bf, None)
    O.bc()
    for _ in range(steps):
        O.step(dt)
        O.bc()
    return O.get_state()[:2], O.floor_count()


def test_untriggered_protection_is_the_roe_step():
    dims = (12, 10, 8)
    q, bf, _ = random_state(*dims, npass=0, seed=5)
    a, _ = run(dims, q, bf, 3, 2e-3, riemann=1)
    b, _ = run(dims, q, bf, 3, 2e-3, riemann=1, protection=True)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_all_flagged_is_the_hlle_step():
    dims = (12, 10, 8)
    q, bf, _ = random_state(*dims, npass=0, seed=3)
    a, _ = run(dims, q, bf, 2, 2e-3, riemann=1, protection=True, prot_dens=1e300)
    b, _ = run(dims, q, bf, 2, 2e-3, riemann=3)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_rarefaction_keeps_density_positive():
    dims = (16, 8, 8)
    q, bf = rarefaction(dims, 10.0)
    dt = 0.3 / 12.0 / 16
    (qa, _), nfa = run(dims, q, bf, 3, dt, riemann=1)
    (qb, _), nfb = run(dims, q, bf, 3, dt, riemann=1, protection=True)
    I = (slice(NB, -NB),) * 3
    assert nfa > 0 and qa[I + (0,)].min() == 1e-20
    assert nfb == 0 and qb[I + (0,)].min() > 1e-3
