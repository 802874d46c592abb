"""Local spin-frame covariance (src/both_3d/random_gauge.cu:198-236): rotating the tangent
frame at every site x by delta_x shifts alpha_x. by delta_x and omega_xy by
delta_x - delta_y; the operator must then transform as D' = R D R^dagger with
R = exp(-i delta_x sigma3 / 2) at every site and slice, and the solution of
D'^dag D' x' = R b as x' = R x.  This fixes where gamma(x,y) and Omega(x,y) enter D: a
gamma or Omega placed on the wrong side of the hop, or built from alpha_yx instead of
alpha_xy, breaks it (checked below with the shift applied the wrong way, which is what
such a misplacement amounts to)."""
import copy

import numpy as np
import pytest

import oracle as O
import qed3_amd as q
from conftest import rel
from pairs import geometry, links, source


def rotated_geometry(g, lattice, delta, wrong=None):
    """Tables of g with the local frame of every site of `lattice` rotated by delta[site].
    wrong = "alpha": the shift goes to alpha_yx instead of alpha_xy; "omega": omega_xy
    shifts by delta_y - delta_x (the opposite sign)."""
    h = copy.deepcopy(g)
    if lattice == 0:
        lk, al, om = h.links, h.alpha, h.omega
    else:
        lk, al, om = h.dual_links, h.dual_alpha, h.dual_omega
    i, j = lk[:, 0], lk[:, 1]
    if wrong == "alpha":
        al[:, 0] += delta[j]
        al[:, 1] += delta[i]
    else:
        al[:, 0] += delta[i]
        al[:, 1] += delta[j]
    om += (delta[j] - delta[i]) if wrong == "omega" else (delta[i] - delta[j])
    return h


def spin_rotation(delta, Lt):
    """R as the diagonal of the reference layout i = Nx s + 2x + a: exp(-+ i delta_x / 2)."""
    r = np.stack([np.exp(-0.5j * delta), np.exp(0.5j * delta)], axis=1).ravel()
    return np.tile(r, Lt)


def oracle_dirac(g, lattice, Lt, M5, at):
    L = O.Lattice(g, lattice)
    return L, O.Dirac(L, Lt, r=1.0, M5=M5, at=at)


@pytest.mark.parametrize("lattice,n,Lt", [(0, 2, 3), (1, 2, 4), (0, 3, 2), (1, 1, 1)])
def test_oracle_frame_covariance(lattice, n, Lt):
    g = geometry(n)
    L0 = O.Lattice(g, lattice)
    at = 0.5 * L0.checks()["mean_ell"]
    delta = np.random.default_rng(n + Lt).normal(0.0, 2.0, L0.V)
    sp, tp = links(lattice, n, Lt, seed=7)
    _, D = oracle_dirac(g, lattice, Lt, 0.1, at)
    D.update(sp, tp)
    gr = rotated_geometry(g, lattice, delta)
    Lr, Dr = oracle_dirac(gr, lattice, Lt, 0.1, at)
    assert Lr.checks()["spin_dev"] < 1e-12  # the reference's consistency assert still holds
    Dr.update(sp, tp)
    R = spin_rotation(delta, Lt)
    v = source(lattice, n, Lt, seed=3)
    for op in (0, 1, 2):
        assert rel(Dr.apply(op, R * v), R * D.apply(op, v)) < 1e-13
    # the solution rotates with the source
    b = source(lattice, n, Lt, seed=5)
    x, _, _ = D.solve(b, tol=1e-13)
    xr, _, _ = Dr.solve(R * b, tol=1e-13)
    assert rel(xr, R * x) < 1e-10
    # planted misplacements: the check fails
    for wrong in ("alpha", "omega"):
        gw = rotated_geometry(g, lattice, delta, wrong=wrong)
        _, Dw = oracle_dirac(gw, lattice, Lt, 0.1, at)
        Dw.update(sp, tp)
        assert rel(Dw.apply(0, R * v), R * D.apply(0, v)) > 1e-2, wrong


@pytest.mark.gpu
@pytest.mark.parametrize("lattice,n,Lt", [(0, 4, 4), (1, 4, 4), (1, 8, 8)])
def test_gpu_frame_covariance(gpu, lattice, n, Lt):
    """The GPU operator and CG on rotated tables: D'(R v) = R (D v) and x' = R x."""
    print(f"case lattice={lattice} n={n} Lt={Lt}", flush=True)
    g = geometry(n)
    L0 = O.Lattice(g, lattice)
    at = 0.5 * L0.checks()["mean_ell"]
    delta = np.random.default_rng(11).normal(0.0, 2.0, L0.V)
    sp, tp = links(lattice, n, Lt, seed=7)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=0.1, at=at)
    Dq.update(sp, tp)
    Dr = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=0.1, at=at, geometry=rotated_geometry(g, lattice, delta))
    Dr.update(sp, tp)
    R = spin_rotation(delta, Lt)
    v = source(lattice, n, Lt, seed=3)
    for op in (q.OP_D, q.OP_DDAG, q.OP_DDAGD, q.OP_DDDAG):
        assert rel(Dr.apply(op, R * v), R * Dq.apply(op, v)) < 1e-13, op
    b = source(lattice, n, Lt, seed=5)
    for solver in (q.CG_F64, q.CG_HALF):
        x, rep = Dq.solve(b, tol=1e-11, solver=solver)
        xr, repr_ = Dr.solve(R * b, tol=1e-11, solver=solver)
        assert rep.status == 0 and repr_.status == 0
        assert rel(xr, R * x) < 1e-9, solver
