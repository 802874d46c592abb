"""Pins the oracle itself: known answers, closed forms and invariants a plausible mistake
(dropped term, wrong sign or index, transposed operand) would break."""
import numpy as np
import pytest

import oracle as O
from conftest import rel
from pairs import geometry, links, oracle_setup, source, wilson_hop_params


def test_mt19937_64_known_answer():
    # C++ standard [rand.predef]: 10000th output of default-seeded mt19937_64
    assert O.mt64(5489, 10000) == 9981545732273789042


@pytest.mark.parametrize("lattice,n,Lt", [(0, 1, 1), (0, 1, 3), (1, 1, 2), (0, 2, 2)])
def test_dense_operator_properties(lattice, n, Lt):
    L, D, at = oracle_setup(lattice, n, Lt, M5=0.1)
    sp, tp = links(lattice, n, Lt, seed=3)
    D.update(sp, tp)
    A = D.dense()
    N = A.shape[0]
    # CSR^H is the explicit adjoint (sparse_dirac.h:132-143)
    for k in range(3):
        e = np.zeros(N, complex)
        e[k * 7 % N] = 1
        assert rel(D.apply(1, e), A.conj().T @ e) < 1e-15
    # nnz = Lt (4*2E + 12V) (dirac_ext.h:49-116)
    assert D.nnz() == Lt * (8 * L.E + 12 * L.V)
    # gamma5-like structure: the Wilson term makes D non-normal but D^dag D is HPD
    H = A.conj().T @ A
    assert np.allclose(H, H.conj().T)
    assert np.linalg.eigvalsh(H).min() > 0


def test_free_field_zero_mode_structure():
    """U=1, r=1, M5=0, 2D: constant spinor is annihilated by the Wilson term (c_x = sum kappa/2)
    only up to the spin connection; the spectrum of D is symmetric under complex conjugation
    pairs lambda, conj(lambda) (D is real-similar for zero gauge field)."""
    L, D, at = oracle_setup(0, 2, 1, M5=0.0)
    D.update(np.zeros((1, L.E)), np.zeros((1, L.V)))
    ev = np.linalg.eigvals(D.dense())
    dist = np.abs(np.conj(ev)[:, None] - ev[None, :]).min(axis=1)
    assert dist.max() < 1e-8


def test_gauge_covariance():
    """D[U^g] = G D[U] G^dag with G = exp(i phi) (gauge_ext.h:185-201, valence.h:84-87)."""
    lattice, n, Lt = 0, 2, 3
    L, D, at = oracle_setup(lattice, n, Lt, M5=0.1)
    sp, tp = links(lattice, n, Lt, seed=5)
    g = geometry(n)
    D.update(sp, tp)
    v = source(lattice, n, Lt, seed=1)
    phi = np.random.default_rng(1).uniform(-np.pi, np.pi, (Lt, L.V))
    G = np.repeat(np.exp(1j * phi).ravel(), 2)
    Dv = D.apply(0, v)
    D.update(sp + phi[:, g.links[:, 0]] - phi[:, g.links[:, 1]], tp + phi - np.roll(phi, -1, 0))
    assert rel(D.apply(0, G * v), G * Dv) < 1e-14


def test_lt1_is_the_2d_operator():
    """Lt=1: kappa_t = 0, so D reduces to c_x = sum_y r kappa/2 + M5 plus spatial hops
    (dirac_simp.h:171-205): D - diag is independent of M5 and r enters linearly."""
    L, D1, at = oracle_setup(0, 2, 1, M5=0.3)
    _, D2, _ = oracle_setup(0, 2, 1, M5=0.0)
    sp, tp = links(0, 2, 1, seed=9)
    D1.update(sp, tp)
    D2.update(sp, tp)
    assert np.allclose(D1.dense() - D2.dense(), 0.3 * np.eye(2 * L.V), atol=1e-15)


def test_cg_matches_dense_solve():
    L, D, at = oracle_setup(1, 1, 2, M5=0.1)
    sp, tp = links(1, 1, 2, seed=4)
    D.update(sp, tp)
    A = D.dense()
    b = source(1, 1, 2, seed=2)
    x, it, mu = D.solve(b, tol=1e-13)
    assert rel(x, np.linalg.solve(A.conj().T @ A, b)) < 1e-11


def test_force_full_equals_local_and_fd():
    lattice, n, Lt = 0, 1, 3
    L, D, at = oracle_setup(lattice, n, Lt, M5=0.2)
    sp, tp = links(lattice, n, Lt, seed=8)
    D.update(sp, tp)
    phi = D.apply(1, source(lattice, n, Lt, seed=3))
    eta, _, _ = D.solve(phi, tol=1e-14)
    fs1, ft1 = D.force(eta, full=True)
    fs2, ft2 = D.force(eta, full=False)
    assert rel(fs1, fs2) < 1e-13 and rel(ft1, ft2) < 1e-13

    def S(a, b):
        D.update(a, b)
        e, _, _ = D.solve(phi, tol=1e-14)
        return np.vdot(phi, e).real

    eps = 1e-5
    for s, l in [(0, 2), (2, 11)]:
        a, b = sp.copy(), sp.copy()
        a[s, l] += eps
        b[s, l] -= eps
        assert abs((S(a, tp) - S(b, tp)) / (2 * eps) - fs1[s, l]) < 1e-6
    for s, x in [(0, 1), (2, 4)]:
        a, b = tp.copy(), tp.copy()
        a[s, x] += eps
        b[s, x] -= eps
        assert abs((S(sp, a) - S(sp, b)) / (2 * eps) - ft1[s, x]) < 1e-6


@pytest.mark.parametrize("compact", [False, True])
@pytest.mark.parametrize("lattice", [0, 1])
def test_gauge_action_force_fd(compact, lattice):
    n, Lt = 2, 3
    L, D, at = oracle_setup(lattice, n, Lt)
    Sg = O.GaugeAction(L, Lt, gsq=1 / 1.4, at=at, compact=compact)
    sp, tp = links(lattice, n, Lt, seed=6, kappa=0.5)
    fs, ft = Sg.force(sp, tp)
    eps = 1e-6
    for s, l in [(0, 0), (2, 7)]:
        a, b = sp.copy(), sp.copy()
        a[s, l] += eps
        b[s, l] -= eps
        fd = (Sg.S(a, tp) - Sg.S(b, tp)) / (2 * eps)
        assert abs(fd - fs[s, l]) < 1e-6 * max(1, abs(fd))
    for s, x in [(1, 2), (2, 3)]:
        a, b = tp.copy(), tp.copy()
        a[s, x] += eps
        b[s, x] -= eps
        fd = (Sg.S(sp, a) - Sg.S(sp, b)) / (2 * eps)
        assert abs(fd - ft[s, x]) < 1e-6 * max(1, abs(fd))
    if not compact:  # quadratic: S(lambda U) = lambda^2 S(U)
        assert abs(Sg.S(2 * sp, 2 * tp) - 4 * Sg.S(sp, tp)) < 1e-9 * Sg.S(sp, tp)


@pytest.mark.parametrize("lattice,bc", [(0, 0), (1, 0), (0, 1)])
def test_net_features_reproduce_the_wilson_hops(lattice, bc):
    """The net's five feature maps with spin-dependent head weights give D r - c_x r, checked
    against the CSR operator and the closed-form diagonal c_x = sum 1/2 r kappa + r kappa_t + M5."""
    n, Lt, C, r, M5 = 2, 4, 16, 1.0, 0.1
    L, D, at = oracle_setup(lattice, n, Lt, M5=M5, bc=bc)
    sp, tp = links(lattice, n, Lt, seed=8)
    D.update(sp, tp)
    v = source(lattice, n, Lt, seed=2)
    out = O.mpnn_forward(D, C, 2, wilson_hop_params(C, r), v[None, :])[0]
    d = (D.apply(0, v) - out).reshape(Lt, L.V, 2) / v.reshape(Lt, L.V, 2)
    g = geometry(n)
    lk = g.links if lattice == 0 else g.dual_links
    der = L.derived()
    kap = np.zeros(L.V)
    np.add.at(kap, lk[:, 0], 0.5 * r * der["kappa"])
    np.add.at(kap, lk[:, 1], 0.5 * r * der["kappa"])
    kt = der["dual_area"] / (L.checks()["mean_ell"] * at)
    c = kap + r * kt + M5
    assert np.abs(d - c[None, :, None]).max() < 1e-12 * c.max()


def test_net_gradient_matches_finite_differences():
    lattice, n, Lt, C, layers, B = 0, 1, 2, 16, 3, 2
    L, D, at = oracle_setup(lattice, n, Lt, M5=0.1)
    sp, tp = links(lattice, n, Lt, seed=4)
    D.update(sp, tp)
    npar = 20 * C + (layers - 2) * 10 * C * C
    rng = np.random.default_rng(3)
    p = (rng.normal(size=npar) + 1j * rng.normal(size=npar)) * 0.25
    r = np.stack([source(lattice, n, Lt, seed=s) for s in (1, 2)])
    e = np.stack([source(lattice, n, Lt, seed=s) for s in (3, 4)])
    loss, g = O.mpnn_loss_grad(D, C, layers, p, r, e, skip=0.3)
    h = 1e-6
    for k in range(4):
        dlt = rng.normal(size=npar) + 1j * rng.normal(size=npar)
        lp, _ = O.mpnn_loss_grad(D, C, layers, p + h * dlt, r, e, skip=0.3)
        lm, _ = O.mpnn_loss_grad(D, C, layers, p - h * dlt, r, e, skip=0.3)
        fd = (lp - lm) / (2 * h)
        an = np.real(np.vdot(g, dlt))
        assert abs(fd - an) < 1e-6 * abs(an), (k, fd, an)
    assert B == r.shape[0]
