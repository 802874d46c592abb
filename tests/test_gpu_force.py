"""GPU pseudofermion force against the oracle (per-link route of dirac_pf.h:107-121)."""
import numpy as np
import pytest

import qed3_amd as q
from conftest import rel
from pairs import links, oracle_setup, source


@pytest.mark.gpu
@pytest.mark.parametrize("lattice,n,Lt,full", [(0, 1, 2, True), (1, 1, 3, True), (0, 2, 4, False),
                                               (1, 2, 4, False), (0, 4, 1, False)])
def test_force_matches_oracle(gpu, lattice, n, Lt, full):
    print(f"case lattice={lattice} n={n} Lt={Lt} full={full}", flush=True)
    L, Do, at = oracle_setup(lattice, n, Lt, M5=0.1)
    sp, tp = links(lattice, n, Lt, seed=13)
    Do.update(sp, tp)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=0.1, at=at)
    Dq.update(sp, tp)
    eta = source(lattice, n, Lt, seed=19)
    fs_r, ft_r = Do.force(eta, full=full)
    fs, ft = Dq.force(eta)
    assert rel(fs, fs_r) < 1e-12
    if Lt > 1:
        assert rel(ft, ft_r) < 1e-12
    else:
        assert np.all(ft == 0)


@pytest.mark.gpu
def test_force_is_gradient_of_action(gpu):
    """F_l = dS_f/dtheta_l by central differences, eps=1e-5 (hmc.cu:228-306)."""
    lattice, n, Lt = 0, 2, 4
    L, Do, at = oracle_setup(lattice, n, Lt, M5=0.2)
    sp, tp = links(lattice, n, Lt, seed=2)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=0.2, at=at)
    Dq.update(sp, tp)
    phi = Dq.adj(source(lattice, n, Lt, seed=6))  # phi = D^dag xi (pseudofermion.h:36-49)
    eta, _ = Dq.solve(phi, tol=1e-13)
    fs, ft = Dq.force(eta)

    def S(sp_, tp_):
        Dq.update(sp_, tp_)
        e, _ = Dq.solve(phi, tol=1e-13)
        return np.vdot(phi, e).real

    eps = 1e-5
    for (s, l) in [(0, 3), (Lt - 1, 17)]:
        a, b = sp.copy(), sp.copy()
        a[s, l] += eps
        b[s, l] -= eps
        fd = (S(a, tp) - S(b, tp)) / (2 * eps)
        assert abs(fd - fs[s, l]) < 1e-6 * max(1, abs(fd))
    for (s, x) in [(1, 5), (Lt - 1, 0)]:
        a, b = tp.copy(), tp.copy()
        a[s, x] += eps
        b[s, x] -= eps
        fd = (S(sp, a) - S(sp, b)) / (2 * eps)
        assert abs(fd - ft[s, x]) < 1e-6 * max(1, abs(fd))


@pytest.mark.gpu
@pytest.mark.parametrize("lattice,n,Lt", [(0, 2, 4), (1, 4, 8)])
def test_pf_action(gpu, lattice, n, Lt):
    """qed3_pf_action (PseudoFermion::S, pseudofermion.h:53-58) on phi = D^dag xi and
    eta = (D^dag D)^{-1} phi: S_f = Re <phi, eta> = |xi|^2 exactly, and = the oracle's."""
    L, Do, at = oracle_setup(lattice, n, Lt, M5=0.1)
    sp, tp = links(lattice, n, Lt, seed=3)
    Do.update(sp, tp)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=0.1, at=at)
    Dq.update(sp, tp)
    xi = source(lattice, n, Lt, seed=44)
    d_xi = q.DeviceArray.from_numpy(xi)
    d_phi = q.DeviceArray(xi.nbytes)
    Dq.apply_device(q.OP_DDAG, q.F64, d_phi, d_xi)
    d_eta = q.DeviceArray(xi.nbytes)
    rep = Dq.solve_device(d_eta, d_phi, tol=1e-12)
    assert rep.status == 0
    S = Dq.pf_action(d_phi, d_eta)
    assert abs(S - np.vdot(xi, xi).real) < 1e-10 * S
    phi = Do.apply(1, xi)
    eta, _, _ = Do.solve(phi, tol=1e-13)
    assert abs(S - np.vdot(phi, eta).real) < 1e-10 * S
