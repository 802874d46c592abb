"""GPU CG on D^dag D against the oracle's MatPoly::solve restatement and a dense solve."""
import numpy as np
import pytest

import qed3_amd as q
from conftest import rel
from pairs import links, oracle_setup, source


def _pair(lattice, n, Lt, M5, seed=3, kappa=1.4):
    L, Do, at = oracle_setup(lattice, n, Lt, M5=M5)
    sp, tp = links(lattice, n, Lt, seed=seed, kappa=kappa)
    Do.update(sp, tp)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=M5, at=at)
    Dq.update(sp, tp)
    return Do, Dq


@pytest.mark.gpu
def test_cg_dense_known_answer(gpu):
    """n=1, Lt=2: the solution equals a dense LU solve of D^dag D (prop_dirac.cu:222-268)."""
    Do, Dq = _pair(0, 1, 2, M5=0.1)
    A = Do.dense()
    b = source(0, 1, 2, seed=5)
    x_ref = np.linalg.solve(A.conj().T @ A, b)
    for solver in (q.CG_F64, q.CG_MIXED, q.CG_HALF):
        x, rep = Dq.solve(b, tol=1e-12, solver=solver)
        assert rep.status == 0 and rep.true_resid <= 1e-12
        assert rel(x, x_ref) < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("lattice,n,Lt,M5", [(0, 4, 8, 0.05), (1, 4, 8, 0.05), (0, 8, 4, 0.1),
                                             (1, 2, 1, 0.1)])
def test_cg_matches_oracle(gpu, lattice, n, Lt, M5):
    print(f"case lattice={lattice} n={n} Lt={Lt} M5={M5}", flush=True)
    Do, Dq = _pair(lattice, n, Lt, M5)
    b = source(lattice, n, Lt, seed=17)
    x_ref, it_ref, _ = Do.solve(b, tol=1e-13)
    for solver in (q.CG_F64, q.CG_MIXED, q.CG_HALF):
        x, rep = Dq.solve(b, tol=1e-12, solver=solver)
        assert rep.status == 0, rep.status
        assert rep.true_resid <= 1e-12
        assert rel(x, x_ref) < 1e-10, (solver, rel(x, x_ref))
        # the reported residual is the true one
        r = b - Do.apply(2, x)
        assert abs(np.linalg.norm(r) / np.linalg.norm(b) - rep.true_resid) < 1e-13


@pytest.mark.gpu
def test_config1_2d_simplicial_uniform_links(gpu):
    """the 2D operator (Lt = 1, so kappa_t = 0, dirac_ext.h:307-320) on
    the simplicial n=8 refinement (642 sites, the reference's largest N_REFINE,
    src/both/prop.cu:22), r=1, m=0.1, link angles uniform in [-pi, pi), Gaussian RHS, FP64 CG
    to 1e-10: the solution equals the oracle's MatPoly::solve and a dense solve."""
    lattice, n, Lt, M5 = 0, 8, 1, 0.1
    L, Do, at = oracle_setup(lattice, n, Lt, M5=M5)
    assert L.V == 642
    sp, tp = links(lattice, n, Lt, seed=1, kappa=0.0)  # kappa <= 0: uniform angles
    assert sp.min() >= -np.pi and sp.max() < np.pi and np.std(sp) > 1.5
    Do.update(sp, tp)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=M5, at=at)
    Dq.update(sp, tp)
    b = source(lattice, n, Lt, seed=3)
    x_ref, it_ref, _ = Do.solve(b, tol=1e-13)
    A = Do.dense()
    x_dense = np.linalg.solve(A.conj().T @ A, b)
    assert rel(x_ref, x_dense) < 1e-10
    for solver in (q.CG_F64, q.CG_MIXED, q.CG_HALF):
        x, rep = Dq.solve(b, tol=1e-10, solver=solver)
        assert rep.status == 0 and rep.true_resid <= 1e-10
        assert rel(x, x_ref) < 1e-8, (solver, rel(x, x_ref))
        r = b - Do.apply(2, x)
        assert abs(np.linalg.norm(r) / np.linalg.norm(b) - rep.true_resid) < 1e-13
    x, rep = Dq.solve(b, tol=1e-12, solver=q.CG_F64)
    assert rel(x, x_ref) < 1e-10


@pytest.mark.gpu
def test_cg_iterations_close_to_oracle(gpu):
    """FP64 CG runs the same Krylov recursion: iteration counts agree to a few."""
    Do, Dq = _pair(0, 4, 4, 0.1)
    b = source(0, 4, 4, seed=23)
    _, it_ref, _ = Do.solve(b, tol=1e-10)
    _, rep = Dq.solve(b, tol=1e-10, solver=q.CG_F64)
    assert abs(rep.iterations - (it_ref + 1)) <= 3, (rep.iterations, it_ref)


@pytest.mark.gpu
def test_cg_zero_rhs_and_errors(gpu):
    Do, Dq = _pair(0, 2, 2, 0.1)
    x, rep = Dq.solve(np.zeros(Dq.N, np.complex128), tol=1e-10)
    assert rep.status == 0 and np.all(x == 0)
    b = source(0, 2, 2, seed=1)
    rep = Dq.solve_device(q.DeviceArray(b.nbytes), q.DeviceArray.from_numpy(b), tol=1e-12,
                          maxiter=2, raise_on_fail=False)
    assert rep.status == 3  # QED3_ERR_NOCONV
    bad = b.copy()
    bad[3] = np.nan
    rep = Dq.solve_device(q.DeviceArray(b.nbytes), q.DeviceArray.from_numpy(bad), tol=1e-12,
                          raise_on_fail=False)
    assert rep.status == 4  # QED3_ERR_NAN


@pytest.mark.gpu
def test_half_storage_cg_edge_cases(gpu):
    """Zero source, iteration cap and a non-finite source on the binary16-storage solver."""
    Do, Dq = _pair(1, 2, 2, 0.1)
    x, rep = Dq.solve(np.zeros(Dq.N, np.complex128), tol=1e-10, solver=q.CG_HALF)
    assert rep.status == 0 and np.all(x == 0)
    b = source(1, 2, 2, seed=1)
    rep = Dq.solve_device(q.DeviceArray(b.nbytes), q.DeviceArray.from_numpy(b), tol=1e-12,
                          maxiter=2, solver=q.CG_HALF, raise_on_fail=False)
    assert rep.status == 3 and rep.iterations == 2
    bad = b.copy()
    bad[3] = np.inf
    rep = Dq.solve_device(q.DeviceArray(b.nbytes), q.DeviceArray.from_numpy(bad), tol=1e-12,
                          solver=q.CG_HALF, raise_on_fail=False)
    assert rep.status == 4
    # a source far outside the binary16 range: the stored fields are scaled to unit rms
    for s in (1e-30, 1e30):
        x, rep = Dq.solve(s * b, tol=1e-11, solver=q.CG_HALF)
        x0, _ = Dq.solve(b, tol=1e-11, solver=q.CG_F64)
        assert rep.status == 0 and rep.true_resid <= 1e-11
        assert rel(x / s, x0) < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("lattice,n,Lt,M5", [(1, 8, 16, 0.02), (0, 8, 16, 0.02)])
def test_half_storage_cg_light_mass(gpu, lattice, n, Lt, M5):
    """Near the chiral limit (cond(D^dag D) ~ 1e4) the binary16-storage solver still reaches
    the FP64 solution and needs at most a third more iterations than the FP32 one."""
    print(f"case lattice={lattice} n={n} Lt={Lt} M5={M5}", flush=True)
    _, Dq = _pair(lattice, n, Lt, M5, kappa=2.0)
    b = source(lattice, n, Lt, seed=41)
    x64, r64 = Dq.solve(b, tol=1e-12, solver=q.CG_F64)
    _, r32 = Dq.solve(b, tol=1e-10, solver=q.CG_MIXED)
    x16, r16 = Dq.solve(b, tol=1e-10, solver=q.CG_HALF)
    assert r16.status == 0 and r16.true_resid <= 1e-10
    assert r16.iterations <= 1.34 * r32.iterations + 10, (r16.iterations, r32.iterations)
    # |x - x64| / |x64| <= cond(D^dag D) * residual, cond ~ 1e4 here
    assert rel(x16, x64) < 1e-6, rel(x16, x64)


@pytest.mark.gpu
def test_gauge_covariance_of_solution(gpu):
    """Solve with U^g and G b; the sink transforms back (prop.cu:246-274)."""
    lattice, n, Lt = 0, 4, 4
    L, Do, at = oracle_setup(lattice, n, Lt, M5=0.1)
    g = Do  # noqa: F841
    sp, tp = links(lattice, n, Lt, seed=31)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=0.1, at=at)
    Dq.update(sp, tp)
    b = source(lattice, n, Lt, seed=8)
    x0, _ = Dq.solve(b, tol=1e-12)
    rng = np.random.default_rng(4)
    phi = rng.uniform(-np.pi, np.pi, (Lt, Dq.V))
    from pairs import geometry
    geo = geometry(n)
    links_arr = geo.links
    sp2 = sp + phi[:, links_arr[:, 0]] - phi[:, links_arr[:, 1]]  # gauge_ext.h:185-201
    tp2 = tp + phi - np.roll(phi, -1, axis=0)
    G = np.repeat(np.exp(1j * phi).ravel(), 2)
    Dq.update(sp2, tp2)
    x1, _ = Dq.solve(G * b, tol=1e-12)
    assert rel(np.conj(G) * x1, x0) < 1e-10


@pytest.mark.gpu
def test_half_storage_cg_truncated_iterates(gpu):
    """Stopped after k iterations (no reliable update yet), the binary16-storage iterate
    equals the FP32 one to binary16 accuracy for odd and even k: the x term that an even
    iteration leaves pending is flushed at the end."""
    _, Dq = _pair(1, 4, 8, 0.05)
    b = source(1, 4, 8, seed=61)
    d_b = q.DeviceArray.from_numpy(b)
    for k in (1, 2, 3, 4, 7):
        xs = []
        for solver in (q.CG_MIXED, q.CG_HALF):
            d_x = q.DeviceArray(b.nbytes)
            rep = Dq.solve_device(d_x, d_b, tol=1e-14, maxiter=k, solver=solver, delta=1e-6,
                                  raise_on_fail=False)
            assert rep.status == 3 and rep.iterations == k and rep.reliable_updates == 0
            xs.append(d_x.to_numpy(np.complex128))
        assert rel(xs[1], xs[0]) < 1e-2, (k, rel(xs[1], xs[0]))


@pytest.mark.gpu
def test_autotuned_shapes_keep_parity(gpu):
    """The autotuner (bare-D sweep, then the CG-iteration refinement) only changes launch
    shapes: the three solvers still reach the oracle's solution."""
    lattice, n, Lt, M5 = 1, 4, 16, 0.05
    Do, Dq = _pair(lattice, n, Lt, M5)
    Dq.autotune()
    t = Dq.tile(q.OP_D, q.F32)
    assert t["block"] in (64, 128, 256) and 1 <= t["ts"] <= Lt, t
    assert Dq.tile(q.OP_DDAG, q.F32) == t  # D and D^dag share the refined shape
    b = source(lattice, n, Lt, seed=71)
    x_ref, _, _ = Do.solve(b, tol=1e-13)
    for solver in (q.CG_F64, q.CG_MIXED, q.CG_HALF):
        x, rep = Dq.solve(b, tol=1e-12, solver=solver)
        assert rep.status == 0 and rel(x, x_ref) < 1e-10, (solver, rel(x, x_ref))


@pytest.mark.gpu
@pytest.mark.parametrize("solver", [q.CG_F64, q.CG_MIXED, q.CG_HALF], ids=["f64", "mixed", "half"])
def test_kernel_timed_solve_matches(gpu, solver):
    """The event-timed solve (direct launches) gives the same solution and iteration count
    as the graph-replayed one, and positive per-kernel times."""
    lattice, n, Lt, M5 = 0, 4, 8, 0.05
    Do, Dq = _pair(lattice, n, Lt, M5)
    b = source(lattice, n, Lt, seed=81)
    d_b = q.DeviceArray.from_numpy(b)
    d_x1, d_x2 = q.DeviceArray(b.nbytes), q.DeviceArray(b.nbytes)
    r1 = Dq.solve_device(d_x1, d_b, tol=1e-12, solver=solver)
    r2, ms = Dq.solve_kernel_times(d_x2, d_b, tol=1e-12, solver=solver)
    assert r1.status == 0 and r2.status == 0 and r1.iterations == r2.iterations
    assert rel(d_x2.to_numpy(np.complex128), d_x1.to_numpy(np.complex128)) < 1e-14
    assert all(t > 0 for t in ms), ms
    if solver != q.CG_HALF:
        assert ms[2] == ms[1]
