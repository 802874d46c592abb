"""Zolotarev sign approximation, multi-shift CG and the overlap operator (overlap.h)."""
import numpy as np
import pytest

import qed3_amd as q
from conftest import rel
from pairs import links, oracle_setup, source


@pytest.mark.parametrize("n,k", [(21, 0.01), (13, 0.05), (8, 0.2)])
def test_zolotarev_is_the_optimal_sign_approximation(n, k):
    """max_{k<=x<=1} |eps(x) - 1| = Delta, reached with alternating sign (equioscillation
    of the optimal rational approximation); eps is odd; partial fractions = product form."""
    z = q.zolotarev(n, k)
    x = np.concatenate([np.geomspace(k, 1, 200001), [k, 1.0]])
    A, cp, c = z["A"][1:], z["cp"][1:], z["c"][1:]

    def eps(t):
        return z["C"] * t * (1 + np.sum(A[None, :] / (t[:, None] ** 2 - k * k / cp[None, :]), axis=1))

    def eps_prod(t):
        r = 2 / (1 + z["lambda_inv"]) * t / (k * z["M"])
        return r * np.prod((k * k - c[None, :] * t[:, None] ** 2)
                           / (k * k - cp[None, :] * t[:, None] ** 2), axis=1)

    e = eps(x) - 1
    D = z["Delta"]
    assert 0 < D < 0.1
    assert np.abs(e).max() <= D * (1 + 1e-6)
    assert np.abs(e).max() >= D * (1 - 1e-3)
    # alternation: the error reaches +Delta and -Delta
    assert e.max() >= D * (1 - 1e-3) and e.min() <= -D * (1 - 1e-3)
    assert np.allclose(eps(-x), -eps(x), rtol=1e-14)
    assert np.allclose(eps_prod(x), eps(x), rtol=1e-10)
    assert np.all(z["shifts"] > 0)


def test_zolotarev_error_falls_with_degree():
    d = [q.zolotarev(n, 0.05)["Delta"] for n in (6, 10, 14, 18)]
    assert all(a > 10 * b for a, b in zip(d, d[1:]))


def test_zolotarev_rejects_bad_modulus():
    with pytest.raises(q.Qed3Error):
        q.zolotarev(21, 1.5)


def _dense(Do, N, op):
    return np.stack([Do.apply(op, np.eye(N, dtype=complex)[i]) for i in range(N)], axis=1)


@pytest.mark.gpu
@pytest.mark.parametrize("lattice", [0, 1])
def test_multishift_matches_dense_solves(gpu, lattice):
    n, Lt = 2, 4
    L, Do, at = oracle_setup(lattice, n, Lt, M5=0.05)
    sp, tp = links(lattice, n, Lt, seed=6)
    Do.update(sp, tp)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=0.05, at=at)
    Dq.update(sp, tp)
    N = 2 * Lt * Dq.V
    A = _dense(Do, N, 2)
    b = source(lattice, n, Lt, seed=2)
    shifts = [0.0, 0.03, 0.4, 5.0]
    xs, reps = q.multishift_solve(Dq, b, shifts, tol=1e-11)
    for s, x, r in zip(shifts, xs, reps):
        ref = np.linalg.solve(A + s * np.eye(N), b)
        assert rel(x, ref) < 1e-9, (s, rel(x, ref))
        assert r.status == 0 and r.true_resid <= 1e-11
    # larger shifts converge in fewer iterations
    its = [r.iterations for r in reps]
    assert its == sorted(its, reverse=True)


@pytest.mark.gpu
def test_spectral_bounds_match_dense_eigenvalues(gpu):
    lattice, n, Lt = 1, 2, 4
    L, Do, at = oracle_setup(lattice, n, Lt, M5=0.1)
    sp, tp = links(lattice, n, Lt, seed=9)
    Do.update(sp, tp)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=0.1, at=at)
    Dq.update(sp, tp)
    ev = np.linalg.eigvalsh(_dense(Do, 2 * Lt * Dq.V, 2))
    lo, hi = q.spectral_bounds(Dq, tol=1e-10, maxiter=20000)
    assert abs(lo - ev[0]) < 1e-6 * ev[0] and abs(hi - ev[-1]) < 1e-6 * ev[-1]


@pytest.mark.gpu
@pytest.mark.parametrize("lattice", [0, 1])
def test_overlap_matches_dense(gpu, lattice):
    """D_ov = the Zolotarev rational function of the dense operator (1e-9), close to the
    exact 1 + X (X^dag X)^{-1/2} (within Delta), its adjoint, and |(D_ov - 1) v| = |v|."""
    n, Lt, nz = 2, 4, 21
    L, Do, at = oracle_setup(lattice, n, Lt, M5=-0.5)
    sp, tp = links(lattice, n, Lt, seed=4)
    Do.update(sp, tp)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=-0.5, at=at)
    Dq.update(sp, tp)
    N = 2 * Lt * Dq.V
    Dd = _dense(Do, N, 0)
    ev = np.linalg.eigvalsh(Dd.conj().T @ Dd)
    lmin, lmax = np.sqrt(ev[0]) * 0.999, np.sqrt(ev[-1]) * 1.001
    ov = q.Overlap(Dq, n=nz, lambda_min=lmin, lambda_max=lmax, tol=1e-12)
    z = q.zolotarev(nz, lmin / lmax)
    H = Dd.conj().T @ Dd / lmax ** 2
    R = np.eye(N, dtype=complex)
    for m in range(1, z["A"].size):
        R += z["A"][m] * np.linalg.inv(H + z["shifts"][m - 1] * np.eye(N))
    Dzol = np.eye(N) + z["C"] / lmax * Dd @ R
    w, U = np.linalg.eigh(H)
    Dex = np.eye(N) + (Dd / lmax) @ (U @ np.diag(w ** -0.5) @ U.conj().T)
    for seed in (1, 2):
        v = source(lattice, n, Lt, seed=seed)
        got = ov.apply(v)
        assert rel(got, Dzol @ v) < 1e-9
        assert rel(got, Dex @ v) < 4 * z["Delta"] + 1e-9
        assert rel(ov.apply(v, dag=True), Dzol.conj().T @ v) < 1e-9
        assert abs(np.linalg.norm(got - v) / np.linalg.norm(v) - 1) < 4 * z["Delta"] + 1e-9


def _dense_zolotarev(Dd, lmax, z):
    """The dense D_ov of the Zolotarev partial fractions at fixed lambda_max and k."""
    N = Dd.shape[0]
    H = Dd.conj().T @ Dd / lmax ** 2
    R = np.eye(N, dtype=complex)
    for m in range(1, z["A"].size):
        R += z["A"][m] * np.linalg.inv(H + z["shifts"][m - 1] * np.eye(N))
    return np.eye(N) + z["C"] / lmax * Dd @ R


def _overlap_pair(lattice, n, Lt, nz, seed, tol):
    L, Do, at = oracle_setup(lattice, n, Lt, M5=-0.5)
    sp, tp = links(lattice, n, Lt, seed=seed)
    Do.update(sp, tp)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=-0.5, at=at)
    Dq.update(sp, tp)
    N = 2 * Lt * Dq.V
    Dd = _dense(Do, N, 0)
    ev = np.linalg.eigvalsh(Dd.conj().T @ Dd)
    lmin, lmax = np.sqrt(ev[0]) * 0.999, np.sqrt(ev[-1]) * 1.001
    ov = q.Overlap(Dq, n=nz, lambda_min=lmin, lambda_max=lmax, tol=tol)
    return Do, Dq, ov, q.zolotarev(nz, lmin / lmax), sp, tp, N, Dd


@pytest.mark.gpu
@pytest.mark.parametrize("lattice", [0, 1])
def test_overlap_force_is_the_derivative_of_the_dense_overlap(gpu, lattice):
    """f_l = -2 d/dtheta_l Re <eta, D_ov eta> for every spatial and temporal link, against
    central differences (eps = 1e-5) of the dense Zolotarev operator with lambda_max and the
    coefficients held fixed, as the reference's grad does (overlap.h:393-465)."""
    n, Lt, nz = 1, 4, 13
    Do, Dq, ov, z, sp, tp, N, _ = _overlap_pair(lattice, n, Lt, nz, seed=5, tol=1e-13)
    lmax = ov.lambda_max
    eta = source(lattice, n, Lt, seed=3)
    fs, ft = ov.force(eta)

    def g(sp2, tp2):
        Do.update(sp2, tp2)
        return np.real(np.vdot(eta, _dense_zolotarev(_dense(Do, N, 0), lmax, z) @ eta))

    eps = 1e-5
    ref_s, ref_t = np.zeros_like(sp), np.zeros_like(tp)
    for ref, base, spatial in ((ref_s, sp, True), (ref_t, tp, False)):
        for idx in np.ndindex(base.shape):
            up, dn = base.copy(), base.copy()
            up[idx] += eps
            dn[idx] -= eps
            gp = g(up, tp) if spatial else g(sp, up)
            gm = g(dn, tp) if spatial else g(sp, dn)
            ref[idx] = -2.0 * (gp - gm) / (2 * eps)
    assert rel(fs, ref_s) < 1e-7, rel(fs, ref_s)
    assert rel(ft, ref_t) < 1e-7, rel(ft, ref_t)
    # the pole terms matter: without them the force is far off
    Do.update(sp, tp)
    assert rel(fs, Dq.force(eta)[0]) > 1e-2


@pytest.mark.gpu
@pytest.mark.parametrize("lattice", [0, 1])
def test_overlap_pseudofermion_solve_matches_dense(gpu, lattice):
    """eta = (D_ov + D_ov^dag)^{-1} phi (PseudoFermion<Overlap>::update_eta) against the
    dense solve, and the true residual the report states."""
    n, Lt, nz = 2, 4, 13
    Do, Dq, ov, z, sp, tp, N, Dd = _overlap_pair(lattice, n, Lt, nz, seed=8, tol=1e-13)
    Dz = _dense_zolotarev(Dd, ov.lambda_max, z)
    A = Dz + Dz.conj().T
    b = source(lattice, n, Lt, seed=12)
    x, rep = ov.solve(b, tol=1e-11)
    ref = np.linalg.solve(A, b)
    assert rel(x, ref) < 1e-9, rel(x, ref)
    assert rep.status == 0 and rep.true_resid <= 1e-11 and rep.iterations > 0
    assert abs(rep.true_resid - np.linalg.norm(b - A @ x) / np.linalg.norm(b)) < 1e-12


def _overlap_hmc_setup(kappa=2.0):
    lattice, n, Lt, M5 = 0, 2, 4, -0.5
    L, Do, at = oracle_setup(lattice, n, Lt, M5=M5)
    sp, tp = links(lattice, n, Lt, seed=5, kappa=kappa)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=M5, at=at)
    Dq.update(sp, tp)
    mu_min, mu_max = q.spectral_bounds(Dq, tol=1e-10, maxiter=20000)
    ov = q.Overlap(Dq, n=13, lambda_min=np.sqrt((1 - 1e-6) * mu_min),
                   lambda_max=np.sqrt((1 + 1e-6) * mu_max), tol=1e-12)
    return Dq, ov, sp, tp


@pytest.mark.gpu
def test_overlap_hmc_dH_scales_as_tau_squared(gpu):
    """HMC with overlap fermions (hmc.cu with IS_OVERLAP) at fixed Zolotarev coefficients:
    H is then a smooth function of the links and the overlap force its gradient, so the
    leapfrog's dH falls as tau^2 (ratio 4 when the step count doubles)."""
    Dq, ov, sp, tp = _overlap_hmc_setup()
    dh = []
    for nsteps in (4, 8):
        H = q.HMC(Dq, 1 / 2.0, compact=True, npf=1, tau=0.2, nsteps=nsteps, nsteps_inner=8,
                  tol=1e-11, solver=q.CG_F64, seed=7)
        H.set_overlap(ov, refresh=False)
        rep = H.trajectory(sp.copy(), tp.copy(), no_reject=True)
        assert rep["cg_solves"] == nsteps + 1 and rep["max_resid"] <= 1e-11
        dh.append(abs(rep["dH"]))
    ratio = dh[0] / dh[1]
    assert 2.5 < ratio < 6.0, (dh, ratio)


@pytest.mark.gpu
def test_overlap_hmc_with_refreshed_bounds_and_back_to_wilson(gpu):
    """With refresh the bounds and coefficients follow every gauge update (Overlap::update);
    the trajectory runs and gives a finite dH.  set_overlap(None) returns the handle to
    Wilson fermions: the same trajectory as a fresh Wilson HMC."""
    Dq, ov, sp, tp = _overlap_hmc_setup()
    H = q.HMC(Dq, 1 / 2.0, compact=True, npf=1, tau=0.2, nsteps=3, nsteps_inner=4,
              tol=1e-11, solver=q.CG_F64, seed=9)
    H.set_overlap(ov, refresh=True)
    rep = H.trajectory(sp.copy(), tp.copy(), no_reject=True)
    assert np.isfinite(rep["dH"]) and rep["max_resid"] <= 1e-11
    H2 = q.HMC(Dq, 1 / 2.0, compact=True, npf=1, tau=0.2, nsteps=3, nsteps_inner=4,
               tol=1e-12, solver=q.CG_F64, seed=9)
    H2.set_overlap(ov)
    H2.set_overlap(None)
    H3 = q.HMC(Dq, 1 / 2.0, compact=True, npf=1, tau=0.2, nsteps=3, nsteps_inner=4,
               tol=1e-12, solver=q.CG_F64, seed=9)
    u2, t2, u3, t3 = sp.copy(), tp.copy(), sp.copy(), tp.copy()
    r2 = H2.trajectory(u2, t2, no_reject=True)
    r3 = H3.trajectory(u3, t3, no_reject=True)
    assert abs(r2["dH"] - r3["dH"]) <= 1e-9 * max(1.0, abs(r3["dH"]))
    assert np.allclose(u2, u3, rtol=0, atol=1e-12) and np.allclose(t2, t3, rtol=0, atol=1e-12)
