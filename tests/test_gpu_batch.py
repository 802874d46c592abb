"""Batched CG over systems of different lattices, sizes, Lt and masses."""
import numpy as np
import pytest

import qed3_amd as q
from conftest import rel
from pairs import links, oracle_setup, source

SYSTEMS = [(0, 2, 4, 0.1), (1, 2, 3, 0.05), (0, 4, 8, 0.02), (1, 4, 2, 0.2), (0, 1, 1, 0.15),
           (1, 1, 5, 0.08), (0, 3, 6, 0.12)]


def _build(systems, seed0=40):
    ops, refs, bs = [], [], []
    for i, (lat, n, Lt, M5) in enumerate(systems):
        L, Do, at = oracle_setup(lat, n, Lt, M5=M5)
        sp, tp = links(lat, n, Lt, seed=seed0 + i, kappa=1.4)
        Do.update(sp, tp)
        D = q.WilsonDirac(lat, n, Lt, r=1.0, M5=M5, at=at)
        D.update(sp, tp)
        ops.append(D)
        refs.append(Do)
        bs.append(source(lat, n, Lt, seed=100 + i))
    return ops, refs, bs


@pytest.mark.gpu
def test_batch_matches_oracle_per_system(gpu):
    ops, refs, bs = _build(SYSTEMS)
    d_b = [q.DeviceArray.from_numpy(b) for b in bs]
    d_x = [q.DeviceArray(b.nbytes) for b in bs]
    reps = q.batch_solve(ops, d_x, d_b, tol=1e-12)
    iters = []
    for D, Do, b, x, rep in zip(ops, refs, bs, d_x, reps):
        assert rep.status == 0 and rep.true_resid <= 1e-12
        x_ref, _, _ = Do.solve(b, tol=1e-13)
        assert rel(x.to_numpy(np.complex128), x_ref) < 1e-10
        # same answer as the single-system solver
        xs, rs = D.solve(b, tol=1e-12, solver=q.CG_MIXED)
        assert rel(x.to_numpy(np.complex128), xs) < 1e-10
        iters.append(rep.iterations)
    assert len(set(iters)) > 1  # systems converged independently


@pytest.mark.gpu
def test_batch_masks_a_failing_system(gpu):
    ops, refs, bs = _build(SYSTEMS[:3], seed0=60)
    bad = bs[1].copy()
    bad[5] = np.nan
    d_b = [q.DeviceArray.from_numpy(b) for b in (bs[0], bad, bs[2])]
    d_x = [q.DeviceArray(b.nbytes) for b in bs]
    reps = q.batch_solve(ops, d_x, d_b, tol=1e-11, raise_on_fail=False)
    assert reps[1].status == 4
    for k in (0, 2):
        assert reps[k].status == 0 and reps[k].true_resid <= 1e-11
