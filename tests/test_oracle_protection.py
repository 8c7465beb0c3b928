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
* partial flags (some cells flagged): a cell-by-cell brute force written from the
  reference text (below) gives the same state, face field and corner EMFs bit for bit on
  arbitrary fluxes with planted flagged cells, for the predictor's and the corrector's
  update ranges.
The GPU matches this restatement bit for bit (tests/test_gpu_protection.py).
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


# ---------------------------------------------------------------------------------------
# Partial flags, brute force.  update_hydro_state_3d and update_CT_3d (vanleer_mod.f90:
# 2080-2259) are restated below cell by cell from the reference text, independently of the
# oracle's loops, and both are run on the same arbitrary Fs, Fs_prot, emf_corner and
# emf_corner_prot with a planted set of cells that the first pass takes below the
# threshold: an isolated cell, a face-adjacent pair, an L-shaped triple, a cell on each
# edge of the update range and two cells whose one common neighbour is not flagged.  The
# face set (six faces of each flagged cell, all seven components), the corner set
# ((i,j,k), (i-1,j,k), (i,j-1,k), (i,j,k-1) of all three EMF components), the redo set
# (flagged cells and their six face neighbours, from the backup) and the CT update that
# reads the swapped corners must all agree bit for bit.

def _brute_force_update(q, bf, Fs, Fsp, ec, ecp, dtdx, off, thr):
    """vanleer_mod.f90:2080-2259 on padded 0-based arrays (Fs[..., comp, dir])"""
    q, bf, Fs, ec = q.copy(), bf.copy(), Fs.copy(), ec.copy()
    N = q.shape[:3]
    lo = 1 + off                                   # reference klow-nb+1 with klow = 1+off
    hi = [n - 2 - off for n in N]                  # reference khi+nb-1
    qprot = q.copy()

    def upd(base, f, i, j, k):
        return base - dtdx * ((Fs[i, j, k, f, 0] - Fs[i - 1, j, k, f, 0]) +
                              (Fs[i, j, k, f, 1] - Fs[i, j - 1, k, f, 1]) +
                              (Fs[i, j, k, f, 2] - Fs[i, j, k - 1, f, 2]))

    cells = [(i, j, k) for k in range(lo, hi[2] + 1) for j in range(lo, hi[1] + 1)
             for i in range(lo, hi[0] + 1)]
    for (i, j, k) in cells:
        for f in range(4):
            q[i, j, k, f] = upd(q[i, j, k, f], f, i, j, k)
    flag = np.zeros(N, dtype=bool)
    for (i, j, k) in cells:
        if q[i, j, k, 0] < thr:
            flag[i, j, k] = True
            for (a, b, c, d) in ((i, j, k, 0), (i - 1, j, k, 0), (i, j, k, 1), (i, j - 1, k, 1),
                                 (i, j, k, 2), (i, j, k - 1, 2)):
                Fs[a, b, c, :, d] = Fsp[a, b, c, :, d]
            for m in range(3):
                for (a, b, c) in ((i, j, k), (i - 1, j, k), (i, j - 1, k), (i, j, k - 1)):
                    ec[a, b, c, m] = ecp[a, b, c, m]
    for (i, j, k) in cells:
        if (flag[i, j, k] or flag[i - 1, j, k] or flag[i + 1, j, k] or flag[i, j - 1, k] or
                flag[i, j + 1, k] or flag[i, j, k - 1] or flag[i, j, k + 1]):
            for f in range(4):
                q[i, j, k, f] = upd(qprot[i, j, k, f], f, i, j, k)
    for (i, j, k) in cells:
        if q[i, j, k, 0] < 1e-20:
            q[i, j, k, 0] = 1e-20
    for (i, j, k) in cells:
        bf[i, j, k, 0] = bf[i, j, k, 0] - dtdx * ((ec[i, j, k, 2] - ec[i, j - 1, k, 2]) -
                                                 (ec[i, j, k, 1] - ec[i, j, k - 1, 1]))
        bf[i, j, k, 1] = bf[i, j, k, 1] - dtdx * ((ec[i, j, k, 0] - ec[i, j, k - 1, 0]) -
                                                 (ec[i, j, k, 2] - ec[i - 1, j, k, 2]))
        bf[i, j, k, 2] = bf[i, j, k, 2] - dtdx * ((ec[i, j, k, 1] - ec[i - 1, j, k, 1]) -
                                                 (ec[i, j, k, 0] - ec[i, j - 1, k, 0]))
    for (i, j, k) in cells:
        if min(i, j, k) < lo + 1:
            continue
        q[i, j, k, 4] = 0.5 * (bf[i, j, k, 0] + bf[i - 1, j, k, 0])
        q[i, j, k, 5] = 0.5 * (bf[i, j, k, 1] + bf[i, j - 1, k, 1])
        q[i, j, k, 6] = 0.5 * (bf[i, j, k, 2] + bf[i, j, k - 1, 2])
    return q, bf, ec, flag


import pytest  # noqa: E402


@pytest.mark.parametrize("off,dtdx", [(0, 0.25), (3, 0.5)])
def test_partial_flags_brute_force(off, dtdx):
    from refutil import P
    dims = (10, 8, 6)
    N = tuple(n + 2 * NB for n in dims)
    rng = np.random.default_rng(11 + off)
    q = np.asfortranarray(rng.uniform(0.9, 1.1, N + (7,)))
    bf = np.asfortranarray(rng.uniform(-1, 1, N + (3,)))
    Fs = np.asfortranarray(rng.uniform(-0.05, 0.05, N + (7, 3)))
    Fsp = np.asfortranarray(rng.uniform(-0.05, 0.05, N + (7, 3)))
    ec = np.asfortranarray(rng.uniform(-1, 1, N + (3,)))
    ecp = np.asfortranarray(rng.uniform(-1, 1, N + (3,)))
    lo, hx, hy, hz = 1 + off, N[0] - 2 - off, N[1] - 2 - off, N[2] - 2 - off
    planted = [(lo, lo, lo), (hx, hy, hz), (lo + 4, lo + 3, lo + 2), (lo + 5, lo + 3, lo + 2),
               (lo + 2, lo + 1, hz), (lo + 2, lo + 2, hz), (lo + 3, lo + 2, hz),
               (hx - 2, lo + 2, lo + 1), (hx - 2, lo + 4, lo + 1)]
    for c in planted:
        q[c + (0,)] = 0.1
    thr = 0.5
    eq, ebf, eec, flag = _brute_force_update(q, bf, Fs, Fsp, ec, ecp, dtdx, off, thr)
    assert sorted(zip(*np.nonzero(flag))) == sorted(planted)
    O = Oracle(*dims, npass=0, mhd=2, riemann=1, recon=0, bc=2, protection=True, prot_dens=thr)
    O.set_state(q, bf, None)
    assert O.L.orc_update_test(O.c, dtdx, off, P(Fs), P(Fsp), P(ec), P(ecp)) == 0
    gq, gbf, _, gec = O.get_state()
    assert np.array_equal(gq, eq)
    assert np.array_equal(gbf, ebf)
    assert np.array_equal(gec, eec)
    # the swap reached exactly the reference's corner slots: each flagged cell's four
    assert int((eec != ec).sum()) == 3 * len({(a, b, c) for (i, j, k) in planted
                                             for (a, b, c) in ((i, j, k), (i - 1, j, k), (i, j - 1, k),
                                                               (i, j, k - 1))})
