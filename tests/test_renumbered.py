"""The drop-in boundary on a caller's own numbering: geometry tables with sites, triangles and
links relabelled at random and link orientations reversed (tests/renumber.py) are imported
through qed3_geom_import + qed3_ctx_create_geom, the gauge field and spinors are given in
that numbering, and D, D^dagger, D^dagger D, D D^dagger, the CG solvers, the learned solver
and the pseudofermion force must equal the oracle built on the same tables, and the
relabelled operator of the original one (D'(P v) = P D v, x' = P x, F' = P F)."""
import numpy as np
import pytest

import oracle as O
import qed3_amd as q
from conftest import rel
from pairs import geometry, links, source
from renumber import renumber


def _oracle(g, lattice, Lt, M5, at, sp, tp, bc=0):
    L = O.Lattice(g, lattice)
    D = O.Dirac(L, Lt, r=1.0, M5=M5, at=at, bc=bc)
    D.update(sp, tp)
    return L, D


@pytest.mark.parametrize("lattice,n,Lt", [(0, 3, 3), (1, 3, 4), (0, 2, 1), (1, 2, 2)])
def test_oracle_relabelling(lattice, n, Lt):
    """The oracle on relabelled tables is the relabelled operator, its solution and force the
    relabelled ones; alpha left unswapped on reversed links (a planted mistake) breaks it."""
    g = geometry(n)
    at = 0.5 * O.Lattice(g, lattice).checks()["mean_ell"]
    sp, tp = links(lattice, n, Lt, seed=4)
    _, D = _oracle(g, lattice, Lt, 0.1, at, sp, tp)
    h, relabel = renumber(g, seed=n + 10 * Lt)
    P = relabel(lattice, Lt)
    Lh, Dh = _oracle(h, lattice, Lt, 0.1, at, P.spatial(sp), P.temporal(tp))
    c = Lh.checks()
    assert c["spin_dev"] < 1e-12 and c["deficit_fail"] == 0
    assert abs(c["mean_ell"] - 2 * at) < 1e-14
    v = source(lattice, n, Lt, seed=3)
    assert rel(D.apply(3, v), D.apply(0, D.apply(1, v))) < 1e-15
    for op in (0, 1, 2, 3):
        assert np.linalg.norm(D.apply(op, v)) > 0.1 * np.linalg.norm(v)
        assert rel(Dh.apply(op, P.spinor(v)), P.spinor(D.apply(op, v))) < 1e-13, op
    b = source(lattice, n, Lt, seed=5)
    x, _, _ = D.solve(b, tol=1e-13)
    xh, _, _ = Dh.solve(P.spinor(b), tol=1e-13)
    assert rel(xh, P.spinor(x)) < 1e-10
    fs, ft = D.force(x)
    fsh, fth = Dh.force(P.spinor(x))
    assert rel(fsh, P.spatial(fs)) < 1e-12
    if Lt > 1:
        assert rel(fth, P.temporal(ft)) < 1e-12
    hw, relabel_w = renumber(g, seed=n + 10 * Lt, wrong="alpha")
    _, Dw = _oracle(hw, lattice, Lt, 0.1, at, P.spatial(sp), P.temporal(tp))
    assert rel(Dw.apply(0, P.spinor(v)), P.spinor(D.apply(0, v))) > 1e-2


@pytest.mark.gpu
@pytest.mark.parametrize("lattice,n,Lt,bc", [(0, 3, 4, 0), (1, 3, 4, 0), (1, 4, 8, 1), (0, 4, 1, 0)])
def test_gpu_on_caller_numbering(gpu, lattice, n, Lt, bc):
    print(f"case lattice={lattice} n={n} Lt={Lt} bc={bc}", flush=True)
    g = geometry(n)
    at = 0.5 * O.Lattice(g, lattice).checks()["mean_ell"]
    M5 = 0.05
    sp, tp = links(lattice, n, Lt, seed=8)
    h, relabel = renumber(g, seed=3 + n)
    P = relabel(lattice, Lt)
    sph, tph = P.spatial(sp), P.temporal(tp)
    _, Do = _oracle(h, lattice, Lt, M5, at, sph, tph, bc=bc)
    Dh = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=M5, at=at, bc=bc, geometry=h)
    Dh.update(sph, tph)
    Dn = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=M5, at=at, bc=bc)
    Dn.update(sp, tp)
    v = source(lattice, n, Lt, seed=3)
    vh = P.spinor(v)
    v16 = (vh.real.astype(np.float16).astype(np.float64)
           + 1j * vh.imag.astype(np.float16).astype(np.float64))  # binary16-representable
    for op in (q.OP_D, q.OP_DDAG, q.OP_DDAGD, q.OP_DDDAG):
        ref = Do.apply(op, vh)
        assert np.linalg.norm(ref) > 0.1 * np.linalg.norm(vh)
        assert rel(Dh.apply(op, vh), ref) < 1e-13, op
        assert rel(Dh.apply(op, vh, q.F32), ref) < 4e-6, op
        tol16 = 1e-3 if op in (q.OP_D, q.OP_DDAG) else 3e-3
        assert rel(Dh.apply(op, v16, q.F16), Do.apply(op, v16)) < tol16, op
        assert rel(Dh.apply(op, vh), P.spinor(Dn.apply(op, v))) < 1e-13, op
    b = source(lattice, n, Lt, seed=5)
    bh = P.spinor(b)
    x_ref, _, _ = Do.solve(bh, tol=1e-13)
    for solver in (q.CG_F64, q.CG_MIXED, q.CG_HALF):
        xh, rep = Dh.solve(bh, tol=1e-11, solver=solver)
        assert rep.status == 0 and rep.true_resid <= 1e-11
        assert rel(xh, x_ref) < 1e-9, solver
    xn, _ = Dn.solve(b, tol=1e-12)
    assert rel(P.spinor(xn), x_ref) < 1e-9
    K = q.Precond(Dh, layers=3)
    xk, rep = K.solve(bh, tol=1e-11)
    assert rep.status == 0 and rel(xk, x_ref) < 1e-9
    fs_r, ft_r = Do.force(x_ref)
    fs, ft = Dh.force(x_ref)
    assert rel(fs, fs_r) < 1e-12
    if Lt > 1:
        assert rel(ft, ft_r) < 1e-12
    fsn, ftn = Dn.force(xn)
    assert rel(fs, P.spatial(fsn)) < 1e-8


@pytest.mark.gpu
def test_caller_stream_orders_work(gpu):
    """qed3_ctx_set_stream + qed3_apply_async on a torch stream: the copy that writes the
    input, the queued applies and the copy that reads the output run in stream order
    without a host wait in between; the results equal the oracle's."""
    import torch
    lattice, n, Lt = 1, 4, 8
    g = geometry(n)
    at = 0.5 * O.Lattice(g, lattice).checks()["mean_ell"]
    sp, tp = links(lattice, n, Lt, seed=12)
    _, Do = _oracle(g, lattice, Lt, 0.1, at, sp, tp)
    Dq = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=0.1, at=at)
    Dq.update(sp, tp)
    st = torch.cuda.Stream()
    Dq.set_stream(st.cuda_stream)
    vs = [source(lattice, n, Lt, seed=s) for s in (21, 22, 23)]
    host = [torch.from_numpy(v) for v in vs]
    dev_in = torch.empty(len(vs[0]), dtype=torch.complex128, device="cuda")
    dev_out = torch.empty_like(dev_in)
    tmp = torch.empty_like(dev_in)
    res = []
    with torch.cuda.stream(st):
        for op, h in zip((q.OP_D, q.OP_DDAGD, q.OP_DDDAG), host):
            dev_in.copy_(h.to("cuda", non_blocking=False))
            Dq.apply_async(op, q.F64, dev_out.data_ptr(), dev_in.data_ptr())
            Dq.apply_async(q.OP_DDAG, q.F64, tmp.data_ptr(), dev_out.data_ptr())
            res.append(tmp.clone())
    st.synchronize()
    for (op, v), r in zip(zip((0, 2, 3), vs), res):
        ref = Do.apply(1, Do.apply(op, v))
        assert rel(r.cpu().numpy(), ref) < 1e-13, op
    Dq.set_stream(None)
    assert rel(Dq.apply(q.OP_D, vs[0]), Do.apply(0, vs[0])) < 1e-13
