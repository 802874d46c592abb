"""GPU time-exact Horner preconditioner and the preconditioned flexible CG (the metric's
learned method) against the oracle: the network and its adjoint, the training loss and
gradient, and the solutions of D^dag D x = b."""
import numpy as np
import pytest

import oracle as O
import qed3_amd as q
from conftest import rel
from pairs import links, oracle_setup, source


def _pair(lattice, n, Lt, M5, bc=0, seed=3, kappa=1.4):
    L, Do, at = oracle_setup(lattice, n, Lt, M5=M5, bc=bc)
    sp, tp = links(lattice, n, Lt, seed=seed, kappa=kappa)
    Do.update(sp, tp)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=M5, at=at, bc=bc)
    Dq.update(sp, tp)
    return Do, Dq


def _learned(layers, seed):
    rng = np.random.default_rng(seed)
    return np.concatenate([1 + 0.3 * (rng.normal(size=layers) + 1j * rng.normal(size=layers)),
                           0.9 + 0.2 * (rng.normal(size=4) + 1j * rng.normal(size=4))])


@pytest.mark.gpu
@pytest.mark.parametrize("lattice,n,Lt,bc,layers", [(0, 2, 4, 0, 3), (1, 2, 4, 0, 4), (1, 4, 8, 1, 2),
                                                    (0, 4, 6, 0, 1), (1, 3, 16, 0, 4)])
def test_precond_apply_matches_oracle(gpu, lattice, n, Lt, bc, layers):
    """K and K^dagger (FP32 on the GPU) = the oracle's dense-LU evaluation (FP64)."""
    print(f"case lattice={lattice} n={n} Lt={Lt} bc={bc} layers={layers}", flush=True)
    Do, Dq = _pair(lattice, n, Lt, 0.05, bc=bc)
    K = q.Precond(Dq, layers=layers)
    v = source(lattice, n, Lt, seed=7)
    for p in (np.ones(layers + 4, complex), _learned(layers, lattice + Lt)):
        K.params = p
        assert np.allclose(K.params, p)
        for adj in (False, True):
            ref = O.horner_apply(Do, layers, p, v, adjoint=adj)
            assert rel(K.apply(v, adjoint=adj), ref) < 2e-6, (adj, rel(K.apply(v, adjoint=adj), ref))


@pytest.mark.gpu
@pytest.mark.parametrize("lattice,n,Lt,bc,layers", [(1, 2, 4, 0, 3), (0, 2, 16, 0, 4), (1, 2, 20, 1, 4),
                                                    (0, 2, 48, 0, 2), (1, 2, 64, 0, 4), (1, 3, 1, 0, 3),
                                                    (0, 4, 33, 0, 3), (0, 3, 64, 1, 4)])
def test_solver_kernels_match_oracle(gpu, lattice, n, Lt, bc, layers):
    """The solver's parallel-in-time layers (binary16 fields in the temporal gauge, time
    chunks of 16 slices per lane with carries between lanes) apply K and K^dagger as the
    oracle does, to binary16 accuracy: every chunking case (one lane, a partial last lane,
    an empty padding lane, four lanes of one wave each on both lattices) and the 2D
    operator."""
    print(f"case lattice={lattice} n={n} Lt={Lt} bc={bc} layers={layers}", flush=True)
    Do, Dq = _pair(lattice, n, Lt, 0.05, bc=bc)
    K = q.Precond(Dq, layers=layers)
    v = source(lattice, n, Lt, seed=9)
    for p in (np.ones(layers + 4, complex), _learned(layers, Lt)):
        K.params = p
        for adj in (False, True):
            ref = O.horner_apply(Do, layers, p, v, adjoint=adj)
            got = K.apply_solver(v, adjoint=adj)
            assert rel(got, ref) < 5e-3, (adj, rel(got, ref))
        # the solver's M = K K^dagger, K's first layer fused into K^dagger's last
        ref = O.horner_apply(Do, layers, p, O.horner_apply(Do, layers, p, v, adjoint=True))
        got = K.apply_solver(v, both=True)
        assert rel(got, ref) < 8e-3, rel(got, ref)


@pytest.mark.gpu
def test_precond_rejects_r_not_one(gpu):
    L, Do, at = oracle_setup(0, 2, 4, M5=0.1, r=0.5)
    Dq = q.WilsonDirac(0, 2, 4, r=0.5, M5=0.1, at=at)
    with pytest.raises(q.Qed3Error):
        q.Precond(Dq, layers=2)


@pytest.mark.gpu
def test_precond_loss_grad_matches_oracle(gpu):
    lattice, n, Lt, layers = 1, 2, 6, 3
    Do, Dq = _pair(lattice, n, Lt, 0.05)
    K = q.Precond(Dq, layers=layers)
    p = _learned(layers, 3)
    K.params = p
    v = np.stack([source(lattice, n, Lt, seed=s) for s in (1, 2, 3)])
    e = np.stack([source(lattice, n, Lt, seed=s) for s in (4, 5, 6)])
    loss, g = K.loss_grad(v, e)
    loss_o, g_o = O.horner_loss_grad(Do, layers, p, v, e)
    assert abs(loss - loss_o) < 1e-5 * loss_o
    assert rel(g, g_o) < 1e-4, rel(g, g_o)


@pytest.mark.gpu
@pytest.mark.parametrize("lattice,n,Lt,M5,bc,layers", [
    (0, 4, 8, 0.05, 0, 4), (1, 4, 8, 0.02, 0, 4), (1, 4, 8, 0.05, 1, 3), (0, 8, 4, 0.1, 0, 2),
    (1, 2, 1, 0.1, 0, 3), (1, 8, 16, 0.02, 0, 4)])
def test_pcg_matches_oracle(gpu, lattice, n, Lt, M5, bc, layers):
    """The preconditioned solve reaches the FP64 true residual and equals the oracle's
    MatPoly::solve restatement (tol 1e-13) to 1e-10, in fewer iterations than the
    unpreconditioned half-storage CG."""
    print(f"case lattice={lattice} n={n} Lt={Lt} M5={M5} bc={bc} layers={layers}", flush=True)
    Do, Dq = _pair(lattice, n, Lt, M5, bc=bc)
    K = q.Precond(Dq, layers=layers)
    b = source(lattice, n, Lt, seed=17)
    x_ref, _, _ = Do.solve(b, tol=1e-13)
    x, rep = K.solve(b, tol=1e-11)
    assert rep.status == 0, rep.status
    assert rep.true_resid <= 1e-11
    r = b - Do.apply(2, x)
    assert abs(np.linalg.norm(r) / np.linalg.norm(b) - rep.true_resid) < 1e-13
    assert rel(x, x_ref) < 1e-10, rel(x, x_ref)
    _, rep_h = Dq.solve(b, tol=1e-11, solver=q.CG_HALF)
    if Lt > 1:
        assert rep.iterations < rep_h.iterations, (rep.iterations, rep_h.iterations)


@pytest.mark.gpu
def test_pcg_edge_cases(gpu):
    """Zero source, scaled sources, iteration cap and a non-finite source."""
    lattice, n, Lt = 1, 4, 8
    Do, Dq = _pair(lattice, n, Lt, 0.05)
    K = q.Precond(Dq, layers=3)
    b = source(lattice, n, Lt, seed=5)
    x0, rep0 = K.solve(b, tol=1e-10)
    for s in (1e-30, 1e30):
        xs, reps = K.solve(s * b, tol=1e-10)
        assert reps.status == 0 and reps.true_resid <= 1e-10
        assert rel(xs / s, x0) < 1e-9
    xz, repz = K.solve(np.zeros_like(b), tol=1e-10)
    assert repz.status == 0 and np.all(xz == 0)
    d_b = q.DeviceArray.from_numpy(b)
    d_x = q.DeviceArray(b.nbytes)
    rep = K.solve_device(d_x, d_b, tol=1e-12, maxiter=2, raise_on_fail=False)
    assert rep.status == 3 and rep.iterations <= 2 + 8
    bn = b.copy()
    bn[3] = np.nan
    rep = K.solve_device(d_x, q.DeviceArray.from_numpy(bn), tol=1e-10, raise_on_fail=False)
    assert rep.status == 4


@pytest.mark.gpu
def test_pcg_x_update_of_a_full_ring(gpu):
    """With rare reliable updates (delta = 1e-6) more than eight iterations run between
    updates, so directions are still pending when a group of eight starts to reuse the ring
    and are first folded into x_inner; the solution still equals the oracle's and reaches
    the FP64 true residual."""
    lattice, n, Lt = 1, 8, 16
    Do, Dq = _pair(lattice, n, Lt, 0.02)
    K = q.Precond(Dq, layers=2)
    b = source(lattice, n, Lt, seed=23)
    x, rep = K.solve(b, tol=1e-10, delta=1e-6)
    assert rep.status == 0 and rep.true_resid <= 1e-10
    assert rep.iterations > 8 * rep.reliable_updates, (rep.iterations, rep.reliable_updates)
    x_ref, _, _ = Do.solve(b, tol=1e-13)
    assert rel(x, x_ref) < 1e-9


@pytest.mark.gpu
def test_pcg_kernel_times(gpu):
    lattice, n, Lt = 1, 8, 16
    Do, Dq = _pair(lattice, n, Lt, 0.02)
    K = q.Precond(Dq, layers=4)
    b = source(lattice, n, Lt, seed=5)
    d_b = q.DeviceArray.from_numpy(b)
    d_x = q.DeviceArray(b.nbytes)
    rep, ms = K.solve_kernel_times(d_x, d_b, tol=1e-10)
    assert rep.status == 0 and rep.true_resid <= 1e-10
    assert all(t > 0 for t in ms[:2] + ms[3:]) and ms[2] >= 0, ms
    x_ref, _, _ = Do.solve(b, tol=1e-13)
    assert rel(d_x.to_numpy(np.complex128), x_ref) < 1e-9


@pytest.mark.gpu
def test_training_gain(gpu):
    """Training lowers the loss on the solver's own (residual, error) pairs; the parameters
    kept after a validation solve (the trained ones unless they raise its iteration count,
    else the Neumann start: the true residual of the solver judges) give a solve that beats
    the method without the network, the half-storage CG, by a wide margin in iterations and
    equals the oracle's solution."""
    lattice, n, Lt, M5, layers = 1, 4, 16, 0.02, 3
    Do, Dq = _pair(lattice, n, Lt, M5)
    K = q.Precond(Dq, layers=layers)
    # pairs (D e, e): errors of truncated solves of the plain CG against the converged one
    es, vs = [], []
    for seed in (31, 32, 33):
        b = source(lattice, n, Lt, seed=seed)
        xs, _ = Dq.solve(b, tol=1e-12, solver=q.CG_F64)
        for k in (2, 5, 10, 20):
            xk, _ = Dq.solve(b, tol=1e-12, maxiter=k, solver=q.CG_F64, raise_on_fail=False)
            e = xs - xk
            es.append(e)
            vs.append(Do.apply(0, e))
    v = np.stack(vs)
    e = np.stack(es)
    p0 = K.params
    loss0, _ = K.loss_grad(v, e)
    d_v, d_e = q.DeviceArray.from_numpy(v), q.DeviceArray.from_numpy(e)
    hist = K.train_device(d_v, d_e, len(es), batch=4, epochs=15, lr=0.02)
    loss1, _ = K.loss_grad(v, e)
    assert loss1 < loss0, (loss0, loss1, hist)
    p1 = K.params
    b_val = source(lattice, n, Lt, seed=98)
    its1 = K.solve(b_val, tol=1e-10)[1].iterations
    K.params = p0
    its0 = K.solve(b_val, tol=1e-10)[1].iterations
    print(f"loss {loss0:.4f} -> {loss1:.4f}; validation iterations Neumann {its0}, trained {its1}", flush=True)
    K.params = p1 if its1 <= its0 else p0
    b = source(lattice, n, Lt, seed=99)
    x, rep = K.solve(b, tol=1e-10)
    _, rep_h = Dq.solve(b, tol=1e-10, solver=q.CG_HALF)
    assert rep.status == 0 and rep.true_resid <= 1e-10
    assert rep.iterations * 3 < rep_h.iterations, (rep.iterations, rep_h.iterations)
    x_ref, _, _ = Do.solve(b, tol=1e-13)
    assert rel(x, x_ref) < 1e-9
