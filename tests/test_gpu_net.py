"""Learned preconditioner: forward/backward kernels against the FP64 oracle of the net,
gauge covariance, training on the GPU, and the gain of flexible CG over plain CG."""
import numpy as np
import pytest

import oracle as O
import qed3_amd as q
from conftest import rel
from pairs import geometry, links, oracle_setup, source

C = 16


def _pair(lattice, n, Lt, M5=0.1, seed=3, kappa=1.4):
    L, Do, at = oracle_setup(lattice, n, Lt, M5=M5)
    sp, tp = links(lattice, n, Lt, seed=seed, kappa=kappa)
    Do.update(sp, tp)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=M5, at=at)
    Dq.update(sp, tp)
    return Do, Dq, sp, tp


def _params(net, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = (rng.normal(size=net.nparams) + 1j * rng.normal(size=net.nparams)) * 0.25 * scale
    return p


@pytest.mark.gpu
@pytest.mark.parametrize("lattice,n,Lt,layers", [(0, 2, 4, 2), (0, 2, 4, 4), (1, 2, 3, 4),
                                                 (1, 1, 2, 3)])
def test_net_forward_matches_oracle(gpu, lattice, n, Lt, layers):
    print(f"case lattice={lattice} n={n} Lt={Lt} layers={layers}", flush=True)
    Do, Dq, _, _ = _pair(lattice, n, Lt)
    net = q.Net(Dq, layers=layers, seed=1)
    p = _params(net, 2)
    net.params = p
    r = np.stack([source(lattice, n, Lt, seed=s) for s in (1, 2)])
    ref = O.mpnn_forward(Do, C, layers, p, r)
    got = net.apply(r)
    # BF16 activations and MFMA operands: ~3 significant digits per layer
    assert rel(got, ref) < 3e-2, rel(got, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("lattice,layers", [(0, 4), (1, 3)])
def test_net_gradient_matches_oracle(gpu, lattice, layers):
    n, Lt = 2, 3
    Do, Dq, _, _ = _pair(lattice, n, Lt)
    net = q.Net(Dq, layers=layers, seed=1)
    p = _params(net, 5)
    net.params = p
    r = np.stack([source(lattice, n, Lt, seed=s) for s in (3, 4)])
    e = np.stack([source(lattice, n, Lt, seed=s) for s in (5, 6)])
    loss_ref, g_ref = O.mpnn_loss_grad(Do, C, layers, p, r, e)
    loss, g = net.loss_grad(r, e)
    assert abs(loss - loss_ref) < 3e-2 * loss_ref
    cos = np.real(np.vdot(g, g_ref)) / (np.linalg.norm(g) * np.linalg.norm(g_ref))
    assert cos > 0.995, cos
    assert rel(g, g_ref) < 6e-2, rel(g, g_ref)
    # every parameter block is exercised: heads and mixing layers
    nv = 5 * C
    for blk in (slice(0, nv), slice(nv, net.nparams - nv), slice(net.nparams - nv, None)):
        assert rel(g[blk], g_ref[blk]) < 1e-1


@pytest.mark.gpu
def test_net_is_gauge_covariant(gpu):
    """N[U^g](G r) = G N[U](r) for a random U(1) gauge transformation G."""
    lattice, n, Lt = 0, 2, 4
    Do, Dq, sp, tp = _pair(lattice, n, Lt)
    net = q.Net(Dq, layers=4, seed=3)
    net.params = _params(net, 7)
    r = source(lattice, n, Lt, seed=9)
    out0 = net.apply(r)
    g = geometry(n)
    phi = np.random.default_rng(2).uniform(-np.pi, np.pi, (Lt, Dq.V))
    G = np.repeat(np.exp(1j * phi).ravel(), 2)
    Dq.update(sp + phi[:, g.links[:, 0]] - phi[:, g.links[:, 1]], tp + phi - np.roll(phi, -1, 0))
    out1 = net.apply(G * r)
    assert rel(out1, G * out0) < 3e-2


def make_pairs(Dq, lattice, n, Lt, cfg_seeds, rhs_seeds, ks, betas):
    """(residual, error) pairs from the solver's own solves, per gauge configuration."""
    data = []
    for cs, beta in zip(cfg_seeds, betas):
        sp, tp = links(lattice, n, Lt, seed=cs, kappa=beta)
        Dq.update(sp, tp)
        rs, es = [], []
        for rsd in rhs_seeds:
            b = source(lattice, n, Lt, seed=rsd)
            xs, _ = Dq.solve(b, tol=1e-11, solver=q.CG_F64)
            for k in ks:
                if k == 0:
                    xk = np.zeros_like(b)  # x0 = 0: r = b, e = x*
                else:
                    d_b = q.DeviceArray.from_numpy(b)
                    d_x = q.DeviceArray(b.nbytes)
                    Dq.solve_device(d_x, d_b, tol=1e-14, maxiter=k, solver=q.CG_F64,
                                    raise_on_fail=False)
                    xk = d_x.to_numpy(np.complex128)
                ek = xs - xk
                rk = b - Dq.DHD(xk)
                s = 1.0 / np.linalg.norm(rk)
                rs.append(rk * s)
                es.append(ek * s)
        data.append((sp, tp, np.array(rs), np.array(es)))
    return data


@pytest.mark.gpu
def test_training_and_fcg_gain(gpu):
    """Train on pairs from the solver's own solves (three beta), then flexible CG with the
    net needs fewer iterations than plain CG on an unseen configuration."""
    lattice, n, Lt, M5 = 1, 4, 8, 0.02
    L, Do, at = oracle_setup(lattice, n, Lt, M5=M5)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=M5, at=at)
    data = make_pairs(Dq, lattice, n, Lt, cfg_seeds=[11, 12, 13], rhs_seeds=[1, 2, 3, 4],
                      ks=[0, 1, 2, 4, 8, 16, 32, 64], betas=[1.0, 1.4, 2.0])
    net = q.Net(Dq, layers=4, seed=5)
    first, last = None, None
    for ep in range(12):
        for sp, tp, rs, es in data:
            Dq.update(sp, tp)
            d_r = q.DeviceArray.from_numpy(rs.astype(np.complex64))
            d_e = q.DeviceArray.from_numpy(es.astype(np.complex64))
            h = net.train_device(d_r, d_e, npairs=len(rs), batch=4, epochs=1, lr=3e-3)
            first = h[0] if first is None else first
            last = h[-1]
    print(f"training loss {first:.4f} -> {last:.4f}", flush=True)
    assert last < 0.7 * first
    # unseen configuration and RHS
    sp, tp = links(lattice, n, Lt, seed=99, kappa=1.4)
    Dq.update(sp, tp)
    b = source(lattice, n, Lt, seed=77)
    d_b = q.DeviceArray.from_numpy(b)
    d_x = q.DeviceArray(b.nbytes)
    rep_cg = Dq.solve_device(d_x, d_b, tol=1e-10, solver=q.CG_F64)
    rep, applies, fb = net.fcg(d_x, d_b, tol=1e-10)
    print(f"plain CG iterations {rep_cg.iterations}; FCG iterations {rep.iterations} "
          f"(fallback={fb}, resid={rep.true_resid:.2e})", flush=True)
    assert rep.status == 0 and rep.true_resid <= 1e-10
    assert rep.iterations < rep_cg.iterations
    # the untrained net of the same shape: the gain comes from training
    net0 = q.Net(Dq, layers=4, seed=5)
    rep0, _, fb0 = net0.fcg(d_x, d_b, tol=1e-10, raise_on_fail=False)
    print(f"untrained FCG iterations {rep0.iterations} (fallback={fb0})", flush=True)
    assert rep.iterations <= rep0.iterations


@pytest.mark.gpu
def test_fcg_falls_back_on_a_broken_preconditioner(gpu):
    """A net that returns zero (all parameters 0) cannot precondition: plain CG finishes."""
    Do, Dq, _, _ = _pair(0, 2, 4, M5=0.1)
    net = q.Net(Dq, layers=3, seed=1)
    net.params = np.zeros(net.nparams)
    b = source(0, 2, 4, seed=3)
    d_b = q.DeviceArray.from_numpy(b)
    d_x = q.DeviceArray(b.nbytes)
    rep, applies, fb = net.fcg(d_x, d_b, tol=1e-11)
    assert fb == 1 and rep.status == 0 and rep.true_resid <= 1e-11
    x = d_x.to_numpy(np.complex128)
    x_ref, _, _ = Do.solve(b, tol=1e-13)
    assert rel(x, x_ref) < 1e-9
