"""GPU parity of the stencil operator (D, D^dag, D^dag D) against the oracle's CSR SpMV."""
import numpy as np
import pytest

import qed3_amd as q
from conftest import rel
from pairs import links, oracle_setup, source

CASES = [(lat, n, Lt) for lat in (0, 1) for n in (1, 2, 4) for Lt in (1, 2, 3, 8)]


@pytest.mark.gpu
@pytest.mark.parametrize("lattice,n,Lt", CASES)
def test_dirac_f64_matches_oracle(gpu, lattice, n, Lt):
    print(f"case lattice={lattice} n={n} Lt={Lt}", flush=True)
    L, Do, at = oracle_setup(lattice, n, Lt, M5=0.1)
    sp, tp = links(lattice, n, Lt, seed=7 + Lt, kappa=1.0)
    Do.update(sp, tp)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=0.1, at=at)
    Dq.update(sp, tp)
    v = source(lattice, n, Lt, seed=3)
    for op in (q.OP_D, q.OP_DDAG, q.OP_DDAGD):
        ref = Do.apply(op, v)
        got = Dq.apply(op, v, q.F64)
        assert rel(got, ref) < 1e-14, (op, rel(got, ref))


@pytest.mark.gpu
@pytest.mark.parametrize("lattice", [0, 1])
def test_dirac_f32_matches_oracle(gpu, lattice):
    n, Lt = 4, 4
    L, Do, at = oracle_setup(lattice, n, Lt, M5=0.05)
    sp, tp = links(lattice, n, Lt, seed=5, kappa=1.4)
    Do.update(sp, tp)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=0.05, at=at)
    Dq.update(sp, tp)
    v = source(lattice, n, Lt, seed=9)
    for op in (q.OP_D, q.OP_DDAG):
        ref = Do.apply(op, v)
        got = Dq.apply(op, v.astype(np.complex64), q.F32)
        assert rel(got, ref) < 2e-6, (op, rel(got, ref))


def round16(v):
    """v with real and imaginary parts rounded to binary16 (the F16 kernels' input)."""
    return v.real.astype(np.float16).astype(np.float64) + 1j * v.imag.astype(np.float16)


@pytest.mark.gpu
@pytest.mark.parametrize("lattice", [0, 1])
def test_dirac_f16_matches_oracle(gpu, lattice):
    """binary16 fields, 16-bit fixed-point link angles, FP32 arithmetic: the FP64 operator
    on the rounded input to binary16 accuracy (unit roundoff 2^-11; the angle step 2 pi/65536
    is finer); the stored angles follow gauge updates."""
    n, Lt = 4, 4
    L, Do, at = oracle_setup(lattice, n, Lt, M5=0.05)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=0.05, at=at)
    v = round16(source(lattice, n, Lt, seed=9))
    for seed in (5, 6):
        sp, tp = links(lattice, n, Lt, seed=seed, kappa=1.4)
        Do.update(sp, tp)
        Dq.update(sp, tp)
        for op in (q.OP_D, q.OP_DDAG, q.OP_DDAGD):
            ref = Do.apply(op, v)
            got = Dq.apply(op, v, q.F16)
            assert rel(got, ref) < 1e-3, (seed, op, rel(got, ref))
    # the rounding is that of binary16, not of something coarser
    assert rel(Dq.apply(q.OP_D, v, q.F16), Do.apply(q.OP_D, v)) > 1e-6


TILES = [dict(block=b, ts=t, remap=r, lds=l) for (b, t, r, l) in
         [(64, 1, 0, 0), (128, 3, 1, 0), (256, 8, 1, 0), (64, 1, 0, 1), (128, 2, 1, 1),
          (256, 5, 0, 1), (256, 32, 1, 1), (64, 64, 0, 0)]]


@pytest.mark.gpu
@pytest.mark.parametrize("lattice", [0, 1])
@pytest.mark.parametrize("tile", TILES, ids=lambda t: "b{block}t{ts}r{remap}l{lds}".format(**t))
def test_every_launch_shape_matches_oracle(gpu, lattice, tile):
    """Every autotuner candidate (block size, time tile, XCD order, LDS staging)."""
    n, Lt = 8, 7
    L, Do, at = oracle_setup(lattice, n, Lt, M5=0.1)
    sp, tp = links(lattice, n, Lt, seed=29, kappa=1.0)
    Do.update(sp, tp)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=0.1, at=at)
    Dq.update(sp, tp)
    Dq.set_tile(**tile)
    v = source(lattice, n, Lt, seed=5)
    for op in (q.OP_D, q.OP_DDAG):
        assert rel(Dq.apply(op, v, q.F64), Do.apply(op, v)) < 1e-14
        assert rel(Dq.apply(op, v.astype(np.complex64), q.F32), Do.apply(op, v)) < 2e-6
        v16 = round16(v)
        assert rel(Dq.apply(op, v16, q.F16), Do.apply(op, v16)) < 1e-3
    x, rep = Dq.solve(Do.apply(1, v), tol=1e-11, solver=q.CG_MIXED)
    assert rep.status == 0 and rep.true_resid <= 1e-11
    x, rep = Dq.solve(Do.apply(1, v), tol=1e-11, solver=q.CG_HALF)
    assert rep.status == 0 and rep.true_resid <= 1e-11


@pytest.mark.gpu
@pytest.mark.parametrize("lattice", [0, 1])
def test_periodic_bc_matches_oracle(gpu, lattice):
    """Periodic time boundary at every precision, including the binary16 operator and the
    half-storage CG (whose temporal-gauge wrap phase carries the boundary sign)."""
    n, Lt = 2, 4
    L, Do, at = oracle_setup(lattice, n, Lt, M5=0.2, bc=1)
    sp, tp = links(lattice, n, Lt, seed=2)
    Do.update(sp, tp)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=0.2, at=at, bc=q.PERIODIC)
    Dq.update(sp, tp)
    v = source(lattice, n, Lt, seed=4)
    for op in (q.OP_D, q.OP_DDAG, q.OP_DDAGD, q.OP_DDDAG):
        ref = Do.apply(op, v) if op != q.OP_DDDAG else Do.apply(0, Do.apply(1, v))
        assert rel(Dq.apply(op, v, q.F64), ref) < 1e-14, op
        assert rel(Dq.apply(op, v.astype(np.complex64), q.F32), ref) < 2e-6, op
    v16 = round16(v)
    for op in (q.OP_D, q.OP_DDAG):
        assert rel(Dq.apply(op, v16, q.F16), Do.apply(op, v16)) < 1e-3, op
    b = source(lattice, n, Lt, seed=8)
    x_ref, _, _ = Do.solve(b, tol=1e-13)
    for solver in (q.CG_F64, q.CG_MIXED, q.CG_HALF):
        x, rep = Dq.solve(b, tol=1e-11, solver=solver)
        assert rep.status == 0 and rep.true_resid <= 1e-11
        assert rel(x, x_ref) < 1e-10, solver


@pytest.mark.gpu
@pytest.mark.parametrize("lattice,n,Lt", [(0, 2, 3), (1, 4, 4), (0, 4, 1)])
def test_dddag_and_dot_match_oracle(gpu, lattice, n, Lt):
    """D D^dag (DiracPf::DDH, dirac_pf.h:85-92) at every precision and <x, y> (the role of
    MatPoly::dot, matpoly.h:129-144) against the oracle / numpy."""
    L, Do, at = oracle_setup(lattice, n, Lt, M5=0.1)
    sp, tp = links(lattice, n, Lt, seed=13)
    Do.update(sp, tp)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=0.1, at=at)
    Dq.update(sp, tp)
    v = source(lattice, n, Lt, seed=21)
    ref = Do.apply(0, Do.apply(1, v))
    assert rel(Dq.apply(q.OP_DDDAG, v, q.F64), ref) < 1e-14
    assert rel(Dq.apply(q.OP_DDDAG, v.astype(np.complex64), q.F32), ref) < 4e-6
    v16 = round16(v)
    assert rel(Dq.apply(q.OP_DDDAG, v16, q.F16), Do.apply(0, Do.apply(1, v16))) < 3e-3
    x, y = source(lattice, n, Lt, seed=31), source(lattice, n, Lt, seed=32)
    d_x, d_y = q.DeviceArray.from_numpy(x), q.DeviceArray.from_numpy(y)
    got = Dq.dot(d_x, d_y)
    want = np.vdot(x, y)  # sum conj(x) y
    assert abs(got - want) < 1e-13 * np.linalg.norm(x) * np.linalg.norm(y)
    assert abs(got.imag - want.imag) < 1e-12 * abs(want) + 1e-12 and abs(want.imag) > 1e-3
    assert abs(Dq.dot(d_x, d_x) - np.linalg.norm(x) ** 2) < 1e-12 * np.linalg.norm(x) ** 2


@pytest.mark.gpu
def test_adjoint_identity_large(gpu):
    """<u, D v> = <D^dag u, v> at a lattice too large for the dense oracle."""
    lattice, n, Lt = 0, 16, 16
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=0.05, at=0.05)
    sp, tp = links(lattice, n, Lt, seed=21)
    Dq.update(sp, tp)
    u = source(lattice, n, Lt, seed=1)
    v = source(lattice, n, Lt, seed=2)
    lhs = np.vdot(u, Dq.mult(v))
    rhs = np.vdot(Dq.adj(u), v)
    assert abs(lhs - rhs) / abs(lhs) < 1e-13


@pytest.mark.gpu
def test_imported_dat_geometry_gives_same_operator(gpu, tmp_path):
    """A context built from .dat tables (the reference's input format) applies the same D."""
    from qed3_amd import dat_io
    lattice, n, Lt = 1, 3, 4
    g = q.Geometry(n)
    dat_io.write_dats(g, tmp_path)
    h = dat_io.read_dats(tmp_path, n)
    sp, tp = links(lattice, n, Lt, seed=3)
    A = q.WilsonDirac(lattice, n, Lt, M5=0.1, at=0.2)
    B = q.WilsonDirac(lattice, n, Lt, M5=0.1, at=0.2, geometry=h)
    A.update(sp, tp)
    B.update(sp, tp)
    v = source(lattice, n, Lt, seed=1)
    assert rel(B.mult(v), A.mult(v)) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("lattice", [0, 1])
def test_f32_phases_over_the_whole_angle_range(gpu, lattice):
    """Uniform angles, shifted by multiples of 2 pi up to +-40 pi, give the FP64 operator
    to FP32 accuracy in the FP32 kernels (phases formed in FP64, then rounded)."""
    n, Lt = 4, 4
    L, Do, at = oracle_setup(lattice, n, Lt, M5=0.1)
    sp, tp = links(lattice, n, Lt, seed=17, kappa=0.0)
    rng = np.random.default_rng(5)
    sp = sp + 2 * np.pi * rng.integers(-20, 21, sp.shape)
    tp = tp + 2 * np.pi * rng.integers(-20, 21, tp.shape)
    Do.update(sp, tp)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=0.1, at=at)
    Dq.update(sp, tp)
    v = source(lattice, n, Lt, seed=4)
    for op in (q.OP_D, q.OP_DDAG, q.OP_DDAGD):
        assert rel(Dq.apply(op, v.astype(np.complex64), q.F32), Do.apply(op, v)) < 2e-6
