"""Pins the oracle's time-exact Horner preconditioner (the learned method's network):
an independent scipy evaluation of the same polynomial, the adjoint identity, U(1) gauge
covariance, the Neumann-series identity and finite differences of the training gradient."""
import numpy as np
import pytest
import scipy.sparse as sps
import scipy.sparse.linalg as spla

import oracle as O
from conftest import rel
from pairs import geometry, links, oracle_setup, source


def split_T_S(D, L, Lt, with_wrap=False):
    """T_open (diagonal + temporal entries between neighbouring slices, no wrap) and
    S = -(spatial entries), classified from the CSR by slice and site of row and column
    (needs Lt >= 3, where the wrap is the only temporal hop with |s - s'| = Lt - 1);
    with_wrap also returns W = -(wrap entries)."""
    A = D.csr().tocoo()
    Nx = 2 * L.V
    rs, cs = A.row // Nx, A.col // Nx
    rx, cx = (A.row % Nx) // 2, (A.col % Nx) // 2
    same_site = rx == cx
    T_keep = same_site & ((rs == cs) | (np.abs(rs - cs) == 1))
    S_keep = (rs == cs) & ~same_site
    W_keep = same_site & (np.abs(rs - cs) == Lt - 1)
    assert np.all(T_keep | S_keep | W_keep)
    T = sps.csr_matrix((A.data * T_keep, (A.row, A.col)), shape=A.shape)
    S = sps.csr_matrix((-A.data * S_keep, (A.row, A.col)), shape=A.shape)
    W = sps.csr_matrix((-A.data * W_keep, (A.row, A.col)), shape=A.shape)
    return (T, S, W) if with_wrap else (T, S + W)


def scaled_S(S, A_, B_):
    """S with entry (row spin a, col spin b) scaled by A_[a] if a == b else B_[a]."""
    C = S.tocoo()
    a, b = C.row % 2, C.col % 2
    f = np.where(a == b, np.asarray(A_)[a], np.asarray(B_)[a])
    return sps.csr_matrix((C.data * f, (C.row, C.col)), shape=C.shape)


def poly_K(T, S, c, v, adjoint=False):
    """sum_j c_j (T^-1 S)^j T^-1 v summed term by term (not in Horner order)."""
    if adjoint:
        T, S, c = T.conj().T.tocsc(), S.conj().T.tocsr(), np.conj(c)
    lu = spla.splu(T.tocsc())
    term = lu.solve(v)
    out = c[0] * term
    for j in range(1, len(c)):
        term = lu.solve(S @ term)
        out = out + c[j] * term
    return out


def params(L, rng, learned):
    if not learned:
        return np.ones(L + 4, complex)
    return np.concatenate([1 + 0.3 * (rng.normal(size=L) + 1j * rng.normal(size=L)),
                           0.9 + 0.2 * (rng.normal(size=4) + 1j * rng.normal(size=4))])


@pytest.mark.parametrize("lattice,n,Lt,bc,layers,learned", [
    (0, 2, 4, 0, 3, False), (1, 2, 3, 0, 4, True), (0, 1, 6, 1, 2, True), (1, 1, 5, 0, 1, True)])
def test_oracle_horner_matches_polynomial(lattice, n, Lt, bc, layers, learned):
    rng = np.random.default_rng(lattice * 7 + Lt)
    L, D, at = oracle_setup(lattice, n, Lt, M5=0.05, bc=bc)
    D.update(*links(lattice, n, Lt, seed=5))
    T, S, W = split_T_S(D, L, Lt, with_wrap=True)
    p = params(layers, rng, learned)
    S_th = scaled_S(S, p[layers:layers + 2], p[layers + 2:layers + 4]) + W  # the wrap is not scaled
    v = source(lattice, n, Lt, seed=3)
    # Horner evaluates sum_j c_j (T^-1 S)^j T^-1 with c_j = p[j]
    for adj in (False, True):
        ref = poly_K(T, S_th, p[:layers], v, adjoint=adj)
        assert rel(O.horner_apply(D, layers, p, v, adjoint=adj), ref) < 1e-12


def test_neumann_identity_and_adjoint():
    """With c = A = B = 1, D K = (T_open - S) K = 1 - (S T_open^-1)^L (S holding the spatial
    hops and the time wrap), and <u, K v> = <K^dag u, v>."""
    lattice, n, Lt, layers = 1, 2, 4, 3
    L, D, at = oracle_setup(lattice, n, Lt, M5=0.02)
    D.update(*links(lattice, n, Lt, seed=8))
    T, S = split_T_S(D, L, Lt)
    p = np.ones(layers + 4, complex)
    v = source(lattice, n, Lt, seed=4)
    u = source(lattice, n, Lt, seed=9)
    Kv = O.horner_apply(D, layers, p, v)
    lu = spla.splu(T.tocsc())
    X = v.copy()
    for _ in range(layers):
        X = S @ lu.solve(X)
    assert rel((T - S) @ Kv, v - X) < 1e-12
    assert rel(D.apply(0, Kv), v - X) < 1e-12  # T_open - S is D itself
    lhs = np.vdot(u, Kv)
    rhs = np.vdot(O.horner_apply(D, layers, p, u, adjoint=True), v)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # a planted mistake (K in place of K^dagger) breaks the identity
    wrong = poly_K(T, S, p[:layers], u, adjoint=False)
    assert abs(np.vdot(wrong, v) - lhs) > 1e-6 * abs(lhs)


def test_horner_gauge_covariance():
    """K[U^g] = G K[U] G^dag with G = exp(i phi) (gauge_ext.h:185-201)."""
    lattice, n, Lt, layers = 0, 2, 4, 3
    L, D, at = oracle_setup(lattice, n, Lt, M5=0.1)
    sp, tp = links(lattice, n, Lt, seed=5)
    g = geometry(n)
    rng = np.random.default_rng(2)
    p = params(layers, rng, True)
    D.update(sp, tp)
    v = source(lattice, n, Lt, seed=1)
    phi = rng.uniform(-np.pi, np.pi, (Lt, L.V))
    G = np.repeat(np.exp(1j * phi).ravel(), 2)
    Kv = O.horner_apply(D, layers, p, v)
    D.update(sp + phi[:, g.links[:, 0]] - phi[:, g.links[:, 1]], tp + phi - np.roll(phi, -1, 0))
    assert rel(O.horner_apply(D, layers, p, G * v), G * Kv) < 1e-13


def test_horner_gradient_matches_finite_differences():
    lattice, n, Lt, layers = 1, 1, 4, 3
    L, D, at = oracle_setup(lattice, n, Lt, M5=0.05)
    D.update(*links(lattice, n, Lt, seed=4))
    rng = np.random.default_rng(6)
    p = params(layers, rng, True)
    v = np.stack([source(lattice, n, Lt, seed=s) for s in (1, 2)])
    e = np.stack([source(lattice, n, Lt, seed=s) for s in (3, 4)])
    loss, g = O.horner_loss_grad(D, layers, p, v, e)
    h = 1e-6
    for k in range(4):
        dlt = rng.normal(size=p.size) + 1j * rng.normal(size=p.size)
        lp, _ = O.horner_loss_grad(D, layers, p + h * dlt, v, e)
        lm, _ = O.horner_loss_grad(D, layers, p - h * dlt, v, e)
        fd = (lp - lm) / (2 * h)
        an = np.real(np.vdot(g, dlt))
        assert abs(fd - an) < 1e-6 * abs(an), (k, fd, an)
    # the loss itself is the definition: mean |K v - e|^2 / |e|^2
    ref = np.mean([np.linalg.norm(O.horner_apply(D, layers, p, v[b]) - e[b]) ** 2
                   / np.linalg.norm(e[b]) ** 2 for b in range(2)])
    assert abs(loss - ref) < 1e-12 * ref
