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
        # same answer and iteration count as the single-system half-storage solver
        xs, rs = D.solve(b, tol=1e-12, solver=q.CG_HALF)
        assert rel(x.to_numpy(np.complex128), xs) < 1e-10
        assert abs(rs.iterations - rep.iterations) <= 2, (rs.iterations, rep.iterations)
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


@pytest.mark.gpu
def test_batch_with_autotuned_shapes_and_scaled_sources(gpu):
    """Systems whose contexts carry autotuned launch shapes (different time tiles per
    system) and sources far outside the binary16 range (per-system scales)."""
    systems = [(0, 4, 16, 0.05), (1, 4, 8, 0.1), (1, 2, 12, 0.02)]
    ops, refs, bs = _build(systems, seed0=80)
    for D in ops:
        D.autotune()
    scales = (1e-25, 1.0, 1e25)
    d_b = [q.DeviceArray.from_numpy(s * b) for s, b in zip(scales, bs)]
    d_x = [q.DeviceArray(b.nbytes) for b in bs]
    reps = q.batch_solve(ops, d_x, d_b, tol=1e-11)
    for Do, b, s, x, rep in zip(refs, bs, scales, d_x, reps):
        assert rep.status == 0 and rep.true_resid <= 1e-11
        x_ref, _, _ = Do.solve(b, tol=1e-13)
        assert rel(x.to_numpy(np.complex128) / s, x_ref) < 1e-9


@pytest.mark.gpu
def test_batch_mixes_large_and_small_systems(gpu):
    """A system whose own grid fills the GPU (dual n=32, Lt=64: 1,280 blocks at the default
    time tile) is solved by the single-system kernels inside the batch call, the small ones
    by the batched launches; every answer equals the single-system half-storage solver's."""
    systems = [(1, 32, 64, 0.1), (0, 2, 4, 0.1), (1, 4, 8, 0.05)]
    ops, bs = [], []
    for i, (lat, n, Lt, M5) in enumerate(systems):
        _, _, at = oracle_setup(lat, n, Lt, M5=M5)
        D = q.WilsonDirac(lat, n, Lt, r=1.0, M5=M5, at=at)
        sp, tp = links(lat, n, Lt, seed=90 + i, kappa=1.4)
        D.update(sp, tp)
        ops.append(D)
        bs.append(source(lat, n, Lt, seed=190 + i))
    d_b = [q.DeviceArray.from_numpy(b) for b in bs]
    d_x = [q.DeviceArray(b.nbytes) for b in bs]
    reps = q.batch_solve(ops, d_x, d_b, tol=1e-11)
    for D, b, x, rep in zip(ops, bs, d_x, reps):
        assert rep.status == 0 and rep.true_resid <= 1e-11
        xs, rs = D.solve(b, tol=1e-11, solver=q.CG_HALF)
        assert rel(x.to_numpy(np.complex128), xs) < 1e-9


@pytest.mark.gpu
def test_batch_rounds_of_large_and_small_systems(gpu):
    """Systems large enough to cover the GPU (>= 256 blocks of 256 at their time tile) run
    in rounds of one chunk per system on their own streams, the small ones in the shared
    launches; every solution equals the oracle's, a NaN source in a large system is reported
    while the others converge."""
    systems = [(1, 8, 64, 0.05), (0, 8, 96, 0.1), (1, 2, 4, 0.1), (1, 4, 64, 0.02)]
    ops, refs, bs = _build(systems, seed0=70)
    for D in ops:  # one slice per block: ceil(V / 256) * Lt blocks
        D.set_tile(256, 1, 1, 0, ops=(q.OP_D, q.OP_DDAG), precs=(q.F32, q.F64))
    large = [-(-D.V // 256) * D.Lt >= 256 for D in ops]
    assert large == [True, True, False, False], large
    d_b = [q.DeviceArray.from_numpy(b) for b in bs]
    d_x = [q.DeviceArray(b.nbytes) for b in bs]
    reps = q.batch_solve(ops, d_x, d_b, tol=1e-11)
    for Do, b, x, rep in zip(refs, bs, d_x, reps):
        assert rep.status == 0 and rep.true_resid <= 1e-11
        x_ref, _, _ = Do.solve(b, tol=1e-13)
        assert rel(x.to_numpy(np.complex128), x_ref) < 1e-9
    bn = bs[0].copy()
    bn[7] = np.nan
    d_b[0] = q.DeviceArray.from_numpy(bn)
    reps = q.batch_solve(ops, d_x, d_b, tol=1e-11, raise_on_fail=False)
    assert reps[0].status == 4
    assert all(r.status == 0 and r.true_resid <= 1e-11 for r in reps[1:])
