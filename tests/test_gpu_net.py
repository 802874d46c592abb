"""Learned preconditioner: forward/backward kernels against the FP64 oracle of the net,
gauge covariance, training on the GPU, and the gain of flexible CG over plain CG."""
import numpy as np
import pytest

import oracle as O
import qed3_amd as q
from qed3_amd import train
from conftest import rel
from pairs import geometry, links, oracle_setup, source, wilson_hop_params

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
    if layers == 4:
        net.skip = 0.4  # a skip comparable to the layers' output
    r = np.stack([source(lattice, n, Lt, seed=s) for s in (1, 2)])
    ref = O.mpnn_forward(Do, C, layers, p, r, skip=net.skip)
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
    net.skip = 0.3
    loss_ref, g_ref = O.mpnn_loss_grad(Do, C, layers, p, r, e, skip=0.3)
    loss, g = net.loss_grad(r, e)
    assert abs(loss - loss_ref) < 3e-2 * loss_ref
    cos = np.real(np.vdot(g, g_ref)) / (np.linalg.norm(g) * np.linalg.norm(g_ref))
    assert cos > 0.995, cos
    assert rel(g, g_ref) < 6e-2, rel(g, g_ref)
    # every parameter block is exercised: heads and mixing layers
    nv = 10 * C
    for blk in (slice(0, nv), slice(nv, net.nparams - nv), slice(net.nparams - nv, None)):
        assert rel(g[blk], g_ref[blk]) < 1e-1


@pytest.mark.gpu
@pytest.mark.parametrize("lattice", [0, 1])
def test_net_reproduces_the_wilson_hops(gpu, lattice):
    """With the hop weights (and no skip) the net applies D - c_x: the spin-dependent time
    projectors and kappa_t are in the device feature maps (BF16 activations: 1e-2)."""
    n, Lt, C_ = 2, 4, 16
    Do, Dq, _, _ = _pair(lattice, n, Lt)
    net = q.Net(Dq, layers=2, seed=1)
    net.params = wilson_hop_params(C_, 1.0)
    net.skip = 0.0
    v = source(lattice, n, Lt, seed=2)
    ref = O.mpnn_forward(Do, C_, 2, net.params, v[None, :])[0]
    assert rel(net.apply(v[None, :])[0], ref) < 1e-2
    # and the oracle's value is (D - c) v, pinned on the CPU against the CSR operator
    d = (Do.apply(0, v) - ref) / v
    assert np.allclose(d.imag, 0, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("lattice", [0, 1])
def test_skip_is_below_inverse_largest_singular_value(gpu, lattice):
    """The default skip c0 satisfies c0 sigma_max(D) <= 1 and is not loose."""
    n, Lt = 2, 4
    Do, Dq, _, _ = _pair(lattice, n, Lt)
    net = q.Net(Dq, layers=2, seed=1)
    N = 2 * Lt * Dq.V
    A = np.stack([Do.apply(2, np.eye(N, dtype=complex)[i]) for i in range(N)], axis=1)
    lmax = np.linalg.eigvalsh(0.5 * (A + A.conj().T)).max()
    assert 0.3 < net.skip * np.sqrt(lmax) <= 1.0, net.skip * np.sqrt(lmax)


@pytest.mark.gpu
@pytest.mark.parametrize("lattice,layers", [(0, 2), (0, 4), (1, 3)])
def test_adjoint_network(gpu, lattice, layers):
    """<u, N v> = <N^dag u, v> for the device adjoint (reverse layers, adjoint features)."""
    n, Lt = 2, 4
    Do, Dq, _, _ = _pair(lattice, n, Lt)
    net = q.Net(Dq, layers=layers, seed=1)
    net.params = _params(net, 4)
    net.skip = 0.2
    u = np.stack([source(lattice, n, Lt, seed=s) for s in (1, 2)])
    v = np.stack([source(lattice, n, Lt, seed=s) for s in (3, 4)])
    Nv, Nu = net.apply(v), net.apply_adj(u)
    for b in range(2):
        lhs, rhs = np.vdot(u[b], Nv[b]), np.vdot(Nu[b], v[b])
        assert abs(lhs - rhs) < 2e-2 * np.linalg.norm(u[b]) * np.linalg.norm(Nv[b]), (lhs, rhs)
    # N^dag N is positive: the preconditioner M = N N^dag of the flexible CG
    assert np.real(np.vdot(Nu[0], Nu[0])) > 0


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


@pytest.mark.gpu
def test_training_and_fcg_gain(gpu):
    """Train the net as an approximate inverse of D on (D e, e) pairs (three beta), then
    flexible CG on D^dag D with M = N N^dag needs fewer iterations than plain CG on an
    unseen configuration and right-hand side."""
    lattice, n, Lt, M5 = 1, 4, 8, 0.02
    L, Do, at = oracle_setup(lattice, n, Lt, M5=M5)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=M5, at=at)
    data = train.make_pairs(Dq, lattice, n, Lt, cfg_seeds=[11, 12, 13], betas=[1.0, 1.4, 2.0])
    net = q.Net(Dq, layers=4, seed=5)
    hist = train.train(net, Dq, data, epochs=40, batch=4, lr=3e-3)
    first, last = hist[0], hist[-1]
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
    # the solution is the oracle's (MatPoly::solve restated, on the same links)
    Do.update(sp, tp)
    x_ref, _, _ = Do.solve(b, tol=1e-13)
    assert rel(d_x.to_numpy(np.complex128), x_ref) < 1e-8
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
    net.skip = 0.0
    b = source(0, 2, 4, seed=3)
    d_b = q.DeviceArray.from_numpy(b)
    d_x = q.DeviceArray(b.nbytes)
    rep, applies, fb = net.fcg(d_x, d_b, tol=1e-11)
    assert fb == 1 and rep.status == 0 and rep.true_resid <= 1e-11
    x = d_x.to_numpy(np.complex128)
    x_ref, _, _ = Do.solve(b, tol=1e-13)
    assert rel(x, x_ref) < 1e-9
