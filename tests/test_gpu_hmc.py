"""HMC trajectory (HMC2 + ExplicitLeapfrogML) against the oracle on the same seeds."""
import numpy as np
import pytest

import oracle as O
import qed3_amd as q
from conftest import rel
from pairs import links, oracle_setup


@pytest.mark.gpu
@pytest.mark.parametrize("lattice,compact", [(0, False), (0, True), (1, True)])
def test_gauge_action_matches_oracle(gpu, lattice, compact):
    n, Lt = 2, 4
    L, Do, at = oracle_setup(lattice, n, Lt)
    sp, tp = links(lattice, n, Lt, seed=4, kappa=0.7)
    Sg = O.GaugeAction(L, Lt, gsq=1 / 1.4, at=at, compact=compact)
    Dq = q.WilsonDirac(lattice, n, Lt, M5=0.1, at=at)
    S, fs, ft = Dq.gauge_action(sp, tp, 1 / 1.4, compact=compact, force=True)
    fs_r, ft_r = Sg.force(sp, tp)
    assert abs(S - Sg.S(sp, tp)) < 1e-12 * abs(S)
    assert rel(fs, fs_r) < 1e-13 and rel(ft, ft_r) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("lattice,compact,npf", [(0, True, 1), (1, False, 2), (1, True, 0)])
def test_hmc_trajectory_matches_oracle(gpu, lattice, compact, npf):
    n, Lt, M5, gsq = 1, 4, 0.1, 1 / 1.4
    print(f"case lattice={lattice} compact={compact} npf={npf}", flush=True)
    L, Do, at = oracle_setup(lattice, n, Lt, M5=M5)
    sp, tp = links(lattice, n, Lt, seed=3, kappa=1.4)
    Sg = O.GaugeAction(L, Lt, gsq=gsq, at=at, compact=compact)
    ref = O.hmc_trajectory(Do, Sg, seed=12, npf=npf, tau=0.5, nsteps=4, nsteps_inner=3,
                           tol=1e-13, sp=sp, tp=tp)
    Dq = q.WilsonDirac(lattice, n, Lt, M5=M5, at=at)
    H = q.HMC(Dq, gsq, compact=compact, npf=npf, tau=0.5, nsteps=4, nsteps_inner=3, tol=1e-12,
              solver=q.CG_F64, seed=12)
    u, t = sp.copy(), tp.copy()
    rep = H.trajectory(u, t, no_reject=True)
    assert abs(rep["h0"] - ref["h0"]) < 1e-9 * abs(ref["h0"])
    assert abs(rep["dH"] - ref["dH"]) < 1e-9 * max(1.0, abs(ref["dH"]))
    assert rel(u, ref["u"]) < 1e-10 and rel(t, ref["t"]) < 1e-10
    assert rep["cg_solves"] == npf * (1 + 4)
    # same accept decision from the same master stream (no_reject only forces the field)
    H2 = q.HMC(Dq, gsq, compact=compact, npf=npf, tau=0.5, nsteps=4, nsteps_inner=3, tol=1e-12,
               solver=q.CG_F64, seed=12)
    rep2 = H2.trajectory(sp.copy(), tp.copy())
    assert bool(rep2["accept"]) == ref["accept"]


@pytest.mark.gpu
def test_hmc_dH_scales_as_tau_squared(gpu):
    """Leapfrog: dH = O(eps^2) (grad.cu:344-369; figs/scaling.plt guide 100 tau^2)."""
    lattice, n, Lt, M5 = 0, 2, 4, 0.2
    L, Do, at = oracle_setup(lattice, n, Lt, M5=M5)
    sp, tp = links(lattice, n, Lt, seed=5, kappa=2.0)
    Dq = q.WilsonDirac(lattice, n, Lt, M5=M5, at=at)
    dh = []
    for nsteps in (4, 8):
        H = q.HMC(Dq, 1 / 2.0, compact=True, npf=1, tau=0.2, nsteps=nsteps, nsteps_inner=8,
                  tol=1e-12, solver=q.CG_F64, seed=7)
        rep = H.trajectory(sp.copy(), tp.copy(), no_reject=True)
        dh.append(abs(rep["dH"]))
    ratio = dh[0] / dh[1]
    assert 2.5 < ratio < 6.0, (dh, ratio)


@pytest.mark.gpu
def test_hmc_parallel_host_path_matches_oracle(gpu):
    """At n=8, Lt=16 (10,272 points, 30,720 links) the momentum and noise draws and the
    leapfrog updates run on host threads; the trajectory still equals the oracle's."""
    lattice, n, Lt, M5, gsq = 0, 8, 16, 0.1, 1 / 1.4
    L, Do, at = oracle_setup(lattice, n, Lt, M5=M5)
    sp, tp = links(lattice, n, Lt, seed=8, kappa=1.4)
    Sg = O.GaugeAction(L, Lt, gsq=gsq, at=at, compact=True)
    ref = O.hmc_trajectory(Do, Sg, seed=21, npf=1, tau=0.2, nsteps=3, nsteps_inner=2,
                           tol=1e-13, sp=sp, tp=tp)
    Dq = q.WilsonDirac(lattice, n, Lt, M5=M5, at=at)
    H = q.HMC(Dq, gsq, compact=True, npf=1, tau=0.2, nsteps=3, nsteps_inner=2, tol=1e-12,
              solver=q.CG_F64, seed=21)
    u, t = sp.copy(), tp.copy()
    rep = H.trajectory(u, t, no_reject=True)
    assert abs(rep["h0"] - ref["h0"]) < 1e-9 * abs(ref["h0"])
    assert abs(rep["dH"] - ref["dH"]) < 1e-9 * max(1.0, abs(ref["h0"]))
    assert rel(u, ref["u"]) < 1e-10 and rel(t, ref["t"]) < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("solver", [q.CG_MIXED, q.CG_HALF], ids=["mixed", "half"])
def test_hmc_with_reduced_precision_solvers_matches_oracle(gpu, solver):
    """The FP32 and binary16-storage inner solvers stop on the FP64 true residual, so the
    trajectory matches the oracle's FP64 one to the solver tolerance times the condition
    number (here ~1e-12 x 1e2)."""
    lattice, n, Lt, M5, gsq = 0, 8, 16, 0.1, 1 / 1.4
    L, Do, at = oracle_setup(lattice, n, Lt, M5=M5)
    sp, tp = links(lattice, n, Lt, seed=8, kappa=1.4)
    Sg = O.GaugeAction(L, Lt, gsq=gsq, at=at, compact=True)
    ref = O.hmc_trajectory(Do, Sg, seed=21, npf=1, tau=0.2, nsteps=3, nsteps_inner=2,
                           tol=1e-13, sp=sp, tp=tp)
    Dq = q.WilsonDirac(lattice, n, Lt, M5=M5, at=at)
    H = q.HMC(Dq, gsq, compact=True, npf=1, tau=0.2, nsteps=3, nsteps_inner=2, tol=1e-12,
              solver=solver, seed=21)
    u, t = sp.copy(), tp.copy()
    rep = H.trajectory(u, t, no_reject=True)
    assert abs(rep["h0"] - ref["h0"]) < 1e-8 * abs(ref["h0"])
    assert abs(rep["dH"] - ref["dH"]) < 1e-7 * max(1.0, abs(ref["h0"]))
    assert rel(u, ref["u"]) < 1e-8 and rel(t, ref["t"]) < 1e-8


@pytest.mark.gpu
def test_hmc_with_learned_solver_matches_oracle(gpu):
    """The pseudofermion solves by the learned method (flexible CG with the time-exact
    preconditioner): the trajectory still equals the oracle's (h0, dH, field)."""
    lattice, n, Lt, M5, gsq = 0, 2, 8, 0.1, 1 / 1.4
    L, Do, at = oracle_setup(lattice, n, Lt, M5=M5)
    sp, tp = links(lattice, n, Lt, seed=6, kappa=1.4)
    Sg = O.GaugeAction(L, Lt, gsq=gsq, at=at, compact=True)
    ref = O.hmc_trajectory(Do, Sg, seed=31, npf=1, tau=0.5, nsteps=4, nsteps_inner=3,
                           tol=1e-13, sp=sp, tp=tp)
    Dq = q.WilsonDirac(lattice, n, Lt, M5=M5, at=at)
    H = q.HMC(Dq, gsq, compact=True, npf=1, tau=0.5, nsteps=4, nsteps_inner=3, tol=1e-12,
              solver=q.PCG, seed=31)
    u, t = sp.copy(), tp.copy()
    rep = H.trajectory(u, t, no_reject=True)
    assert rep["max_resid"] <= 1e-12
    assert abs(rep["h0"] - ref["h0"]) < 1e-9 * abs(ref["h0"])
    assert abs(rep["dH"] - ref["dH"]) < 1e-8 * max(1.0, abs(ref["dH"]))
    assert rel(u, ref["u"]) < 1e-10 and rel(t, ref["t"]) < 1e-10


@pytest.mark.gpu
def test_hmc_set_steps(gpu):
    """set_steps before a trajectory gives the trajectory of a handle created with those
    steps (same seed), and is rejected for a non-positive length."""
    lattice, n, Lt, M5, gsq = 1, 1, 4, 0.1, 1 / 1.4
    L, Do, at = oracle_setup(lattice, n, Lt, M5=M5)
    sp, tp = links(lattice, n, Lt, seed=2, kappa=1.4)
    Dq = q.WilsonDirac(lattice, n, Lt, M5=M5, at=at)
    H1 = q.HMC(Dq, gsq, compact=True, npf=1, tau=1.0, nsteps=2, nsteps_inner=2, tol=1e-12,
               solver=q.CG_F64, seed=5)
    H1.set_steps(0.3, 3, 4)
    u1, t1 = sp.copy(), tp.copy()
    r1 = H1.trajectory(u1, t1, no_reject=True)
    H2 = q.HMC(Dq, gsq, compact=True, npf=1, tau=0.3, nsteps=3, nsteps_inner=4, tol=1e-12,
               solver=q.CG_F64, seed=5)
    u2, t2 = sp.copy(), tp.copy()
    r2 = H2.trajectory(u2, t2, no_reject=True)
    assert r1["dH"] == r2["dH"] and np.array_equal(u1, u2) and np.array_equal(t1, t2)
    with pytest.raises(q.Qed3Error):
        H1.set_steps(0.0, 3, 4)
