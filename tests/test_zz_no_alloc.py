"""qed3_cg_noalloc, the FP64 CG that works in the library's own per-GPU array instead of
allocated memory, against qed3_cg (FP64) and the oracle's MatPoly::solve restatement:
calls in a row on lattices of different sizes (the array keeps the last call's contents)
and the size limit."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import qed3_amd as q
from qed3_amd._lib import CGOpts, CGReport, load
from conftest import rel
from pairs import links, oracle_setup, source

ERR_ARG, ERR_NOCONV, ERR_NAN = 1, 3, 4  # QED3_ERR_* of include/qed3.h


def last_error():
    return load().qed3_last_error().decode()


def _noalloc(Dq, d_x, d_b, tol, maxiter=100000, solver=q.CG_F64):
    opts = CGOpts(tol, maxiter, solver, 0.01, 0)
    rep = CGReport()
    code = load().qed3_cg_noalloc(Dq._h, d_x.ptr, d_b.ptr, C.byref(opts), C.byref(rep))
    return code, rep


def _case(lattice, n, Lt, M5, bc, seed):
    L, Do, at = oracle_setup(lattice, n, Lt, M5=M5, bc=bc)
    sp, tp = links(lattice, n, Lt, seed=seed)
    Do.update(sp, tp)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=M5, at=at, bc=bc)
    Dq.update(sp, tp)
    return Do, Dq


@pytest.mark.gpu
def test_noalloc_matches_cg_in_a_row(gpu):
    """Three calls in a row: a larger lattice, a smaller one, the larger one again; each
    solution equals qed3_cg's FP64 solution and the oracle's."""
    cases = [(0, 8, 8, 0.05, 0, 3), (1, 4, 4, 0.1, 1, 5), (0, 8, 8, 0.05, 0, 3)]
    made = {}
    for lattice, n, Lt, M5, bc, seed in cases:
        print(f"case lattice={lattice} n={n} Lt={Lt} M5={M5} bc={bc}", flush=True)
        key = (lattice, n, Lt, M5, bc, seed)
        if key not in made:
            made[key] = _case(lattice, n, Lt, M5, bc, seed)
        Do, Dq = made[key]
        b = source(lattice, n, Lt, seed=seed + 40)
        d_b = q.DeviceArray.from_numpy(b)
        d_x = q.DeviceArray(b.nbytes)
        d_x_cg = q.DeviceArray(b.nbytes)
        code, rep = _noalloc(Dq, d_x, d_b, 1e-11)
        assert code == 0, last_error()
        assert rep.status == 0 and rep.true_resid <= 1e-11, (rep.status, rep.true_resid)
        rep_cg = Dq.solve_device(d_x_cg, d_b, tol=1e-11, solver=q.CG_F64)
        x = d_x.to_numpy(np.complex128)
        x_cg = d_x_cg.to_numpy(np.complex128)
        assert abs(rep.iterations - rep_cg.iterations) <= 2, (rep.iterations, rep_cg.iterations)
        assert rel(x, x_cg) < 1e-10, rel(x, x_cg)
        x_ref, _, _ = Do.solve(b, tol=1e-13)
        assert rel(x, x_ref) < 1e-10, rel(x, x_ref)
        r = b - Do.apply(2, x)
        assert abs(np.linalg.norm(r) / np.linalg.norm(b) - rep.true_resid) < 1e-13


@pytest.mark.gpu
def test_noalloc_edge_cases(gpu):
    """Zero source, an iteration cap, a non-finite source, a solver other than FP64 and a
    lattice beyond the array's limit (V*Lt = 655,360 > 349,520 points) are reported like
    qed3_cg reports them; the limit names itself in the error text."""
    Do, Dq = _case(1, 4, 8, 0.05, 0, 9)
    b = source(1, 4, 8, seed=2)
    d_x = q.DeviceArray(b.nbytes)
    code, rep = _noalloc(Dq, d_x, q.DeviceArray.from_numpy(np.zeros_like(b)), 1e-10)
    assert code == 0 and rep.status == 0
    assert np.all(d_x.to_numpy(np.complex128) == 0)
    code, rep = _noalloc(Dq, d_x, q.DeviceArray.from_numpy(b), 1e-12, maxiter=3)
    assert code == ERR_NOCONV and rep.status == ERR_NOCONV and rep.iterations <= 3 + 8
    bn = b.copy()
    bn[5] = np.nan
    code, rep = _noalloc(Dq, d_x, q.DeviceArray.from_numpy(bn), 1e-10)
    assert code == ERR_NAN and rep.status == ERR_NAN
    code, _ = _noalloc(Dq, d_x, q.DeviceArray.from_numpy(b), 1e-10, solver=q.CG_HALF)
    assert code == ERR_ARG
    big = q.WilsonDirac(1, 32, 32, r=1.0, M5=0.1, at=0.05)
    big.update(*links(1, 32, 32, seed=1))
    V, _, _ = q.lattice_sizes(1, 32)
    nb = 2 * V * 32
    assert V * 32 == 655360
    d_bb = q.DeviceArray(nb * 16)
    d_xb = q.DeviceArray(nb * 16)
    code, _ = _noalloc(big, d_xb, d_bb, 1e-10)
    assert code == ERR_ARG
    assert "349520" in last_error(), last_error()
    # the entry point still works after the rejection
    code, rep = _noalloc(Dq, d_x, q.DeviceArray.from_numpy(b), 1e-11)
    assert code == 0 and rep.true_resid <= 1e-11
    x_ref, _, _ = Do.solve(b, tol=1e-13)
    assert rel(d_x.to_numpy(np.complex128), x_ref) < 1e-10
