"""qed3_amd: MI355X-native fermion path of qed3 (Wilson-Dirac on the refined icosahedron,
CG on D^dagger D, pseudofermion force) behind the C ABI of include/qed3.h.

The Python layer is a thin host mirror of the reference's roles for tests and scripts:
  Geometry       -> the .dat tables the reference reads (s2n_simp.h)
  WilsonDirac    -> DiracExt + DiracPf (update / mult / adj / DHD / force)
  WilsonDirac.solve -> MatPoly::solve on D^dag D
All compute goes through libqed3.so; there is no fallback path.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import CGOpts, CGReport, Params, Qed3Error, check, load

SIMPLICIAL, DUAL = 0, 1
ANTIPERIODIC, PERIODIC = 0, 1
OP_D, OP_DDAG, OP_DDAGD, OP_DDDAG = 0, 1, 2, 3
F64, F32, F16 = 0, 1, 2
CG_F64, CG_MIXED, CG_HALF = 0, 1, 2
PCG = 3  # the learned method (time-exact preconditioner in flexible CG), as an HMC solver


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Geometry:
    """Refined icosahedron tables (sites, links, faces, dual, spin structure)."""

    def __init__(self, n_refine: int):
        lib = load()
        h = C.c_void_p()
        check(lib.qed3_geom_create(n_refine, C.byref(h)))
        try:
            V, E, F = C.c_int(), C.c_int(), C.c_int()
            check(lib.qed3_geom_counts(h, C.byref(V), C.byref(E), C.byref(F)))
            V, E, F = V.value, E.value, F.value
            self.n, self.V, self.E, self.F = n_refine, V, E, F
            self.sites = np.zeros((V, 3))
            self.faces = np.zeros((F, 3), np.int32)
            self.links = np.zeros((E, 2), np.int32)
            self.nn_offsets = np.zeros(V + 1, np.int32)
            self.nns = np.zeros(2 * E, np.int32)
            self.dual_sites = np.zeros((F, 3))
            self.dual_links = np.zeros((E, 2), np.int32)
            self.dual_nns = np.zeros((F, 3), np.int32)
            self.dual_face_offsets = np.zeros(V + 1, np.int32)
            self.dual_faces = np.zeros(3 * F, np.int32)
            self.alpha = np.zeros((E, 2))
            self.omega = np.zeros(E)
            self.dual_alpha = np.zeros((E, 2))
            self.dual_omega = np.zeros(E)
            check(lib.qed3_geom_export(
                h, _dp(self.sites), _ip(self.faces), _ip(self.links), _ip(self.nn_offsets),
                _ip(self.nns), _dp(self.dual_sites), _ip(self.dual_links), _ip(self.dual_nns),
                _ip(self.dual_face_offsets), _ip(self.dual_faces), _dp(self.alpha),
                _dp(self.omega), _dp(self.dual_alpha), _dp(self.dual_omega)))
        finally:
            lib.qed3_geom_destroy(h)

    def neighbours(self, x):
        return self.nns[self.nn_offsets[x]:self.nn_offsets[x + 1]]

    def dual_ring(self, x):
        return self.dual_faces[self.dual_face_offsets[x]:self.dual_face_offsets[x + 1]]


def import_geometry(g):
    """qed3_geom handle built from tables with Geometry's attributes (caller destroys)."""
    arrs = [np.ascontiguousarray(a, dt) for a, dt in (
        (g.sites, np.float64), (g.faces, np.int32), (g.links, np.int32),
        (g.nn_offsets, np.int32), (g.nns, np.int32), (g.dual_sites, np.float64),
        (g.dual_links, np.int32), (g.dual_nns, np.int32), (g.dual_face_offsets, np.int32),
        (g.dual_faces, np.int32), (g.alpha, np.float64), (g.omega, np.float64),
        (g.dual_alpha, np.float64), (g.dual_omega, np.float64))]
    ptrs = [(_dp if a.dtype == np.float64 else _ip)(a) for a in arrs]
    h = C.c_void_p()
    check(load().qed3_geom_import(int(g.n), len(g.sites), len(g.links), len(g.faces), *ptrs,
                                  C.byref(h)))
    return h


def lattice_sizes(lattice: int, n: int):
    """(V, E, F) of the simplicial or dual lattice at refinement n."""
    if lattice == SIMPLICIAL:
        return 10 * n * n + 2, 30 * n * n, 20 * n * n
    return 20 * n * n, 30 * n * n, 10 * n * n + 2


def random_gauge(lattice: int, n: int, Lt: int, seed: int, kappa: float):
    """Seeded angles (spatial[Lt][E], temporal[Lt][V]); kappa<=0: uniform [-pi,pi)."""
    V, E, _ = lattice_sizes(lattice, n)
    sp = np.zeros((Lt, E))
    tp = np.zeros((Lt, V))
    check(load().qed3_random_gauge(Lt, E, V, seed, float(kappa), _dp(sp), _dp(tp)))
    return sp, tp


def random_source(lattice: int, n: int, Lt: int, seed: int):
    """Seeded complex Gaussian field of 2*V*Lt components (reference layout)."""
    V, E, _ = lattice_sizes(lattice, n)
    xi = np.zeros(2 * V * Lt, np.complex128)
    check(load().qed3_random_source(Lt, E, V, seed, _dp(xi.view(np.float64))))
    return xi


class DeviceArray:
    """Raw device allocation owned by Python (hipMalloc through the C ABI)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        check(load().qed3_malloc(C.byref(self.ptr), self.nbytes))

    @classmethod
    def from_numpy(cls, a: np.ndarray):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        check(load().qed3_memcpy(d.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, 0))
        return d

    def to_numpy(self, dtype, count=None):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize if count is None else count, dtype)
        check(load().qed3_memcpy(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes, 1))
        return out

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(load().qed3_memcpy(self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, 0))

    def free(self):
        if self.ptr:
            load().qed3_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class WilsonDirac:
    """Wilson-Dirac operator on S^2 x time (DiracExt + DiracPf roles) on one GPU."""

    def __init__(self, lattice: int, n_refine: int, Lt: int, r: float = 1.0, M5: float = 0.0,
                 at: float = None, bc: int = ANTIPERIODIC, geometry=None):
        """geometry: optional tables (a Geometry or dat_io.read_dats() result) to use
        instead of the generated refined icosahedron."""
        lib = load()
        if at is None:
            at = 1.0
        self.params = Params(lattice, n_refine, Lt, bc, r, M5, at)
        self._h = C.c_void_p()
        if geometry is None:
            check(lib.qed3_ctx_create(C.byref(self.params), C.byref(self._h)))
        else:
            gh = import_geometry(geometry)
            try:
                check(lib.qed3_ctx_create_geom(C.byref(self.params), gh, C.byref(self._h)))
            finally:
                lib.qed3_geom_destroy(gh)
        Vc, Ec, Fc, ml = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        check(lib.qed3_ctx_sizes(self._h, C.byref(Vc), C.byref(Ec), C.byref(Fc), C.byref(ml)))
        self.V, self.E, self.F, self.mean_ell = Vc.value, Ec.value, Fc.value, ml.value
        self.Lt, self.lattice, self.n = Lt, lattice, n_refine
        self.N = 2 * self.V * Lt  # complex components

    def __del__(self):
        try:
            if self._h:
                load().qed3_ctx_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass

    def geometry(self):
        k, kt = np.zeros(self.E), np.zeros(self.V)
        lv, ell, da = np.zeros(self.E), np.zeros(self.E), np.zeros(self.V)
        check(load().qed3_ctx_geometry(self._h, _dp(k), _dp(kt), _dp(lv), _dp(ell), _dp(da)))
        return dict(kappa=k, kappa_t=kt, link_volume=lv, ell=ell, dual_area=da)

    def update(self, spatial, temporal):
        sp = np.ascontiguousarray(spatial, np.float64)
        tp = np.ascontiguousarray(temporal, np.float64)
        assert sp.size == self.Lt * self.E and tp.size == self.Lt * self.V
        check(load().qed3_set_gauge(self._h, sp.ctypes.data_as(C.c_void_p),
                                    tp.ctypes.data_as(C.c_void_p), 0))

    # device-pointer level
    def apply_device(self, op, prec, d_out, d_in):
        check(load().qed3_apply(self._h, op, prec, d_out.ptr, d_in.ptr))

    def apply_async(self, op, prec, out_ptr, in_ptr):
        """op(in) queued on the context's stream without waiting; raw device addresses
        (ints, e.g. torch's data_ptr()) or DeviceArrays."""
        ptr = lambda a: a.ptr if isinstance(a, DeviceArray) else C.c_void_p(int(a))
        check(load().qed3_apply_async(self._h, op, prec, ptr(out_ptr), ptr(in_ptr)))

    def set_stream(self, stream):
        """Queue this context's work on `stream` (a hipStream_t as an int, e.g. torch's
        Stream.cuda_stream; None or 0 selects the context's own stream)."""
        check(load().qed3_ctx_set_stream(self._h, C.c_void_p(int(stream or 0))))

    def apply(self, op, v: np.ndarray, prec=F64):
        """op(v) for a host complex vector (round trip through the device; F16 packs the
        vector into binary16 (re, im) pairs and returns complex128)."""
        if prec == F16:
            v = np.ascontiguousarray(v, np.complex128)
            assert v.size == self.N
            h = np.empty(2 * v.size, np.float16)
            h[0::2], h[1::2] = v.real, v.imag
            d_in = DeviceArray.from_numpy(h)
            d_out = DeviceArray(h.nbytes)
            self.apply_device(op, prec, d_out, d_in)
            o = d_out.to_numpy(np.float16).astype(np.float64)
            return o[0::2] + 1j * o[1::2]
        dt = np.complex128 if prec == F64 else np.complex64
        v = np.ascontiguousarray(v, dt)
        assert v.size == self.N
        d_in = DeviceArray.from_numpy(v)
        d_out = DeviceArray(v.nbytes)
        self.apply_device(op, prec, d_out, d_in)
        return d_out.to_numpy(dt)

    def mult(self, v, prec=F64):
        return self.apply(OP_D, v, prec)

    def adj(self, v, prec=F64):
        return self.apply(OP_DDAG, v, prec)

    def DHD(self, v, prec=F64):
        return self.apply(OP_DDAGD, v, prec)

    def solve_device(self, d_x, d_b, tol=1e-10, maxiter=100000, solver=CG_F64, delta=0.01,
                     check_every=0, raise_on_fail=True):
        opts = CGOpts(tol, maxiter, solver, delta, check_every)
        rep = CGReport()
        code = load().qed3_cg(self._h, d_x.ptr, d_b.ptr, C.byref(opts), C.byref(rep))
        if code != 0 and raise_on_fail:
            check(code)
        return rep

    def solve_kernel_times(self, d_x, d_b, tol=1e-10, maxiter=100000, solver=CG_HALF,
                           delta=0.01):
        """One solve with HIP-event timing of its stencil kernels: (report, [K1 ms,
        K2 ms, K2-with-x-update ms]) averaged over the iterations run."""
        opts = CGOpts(tol, maxiter, solver, delta, 0)
        rep = CGReport()
        ms = (C.c_double * 3)()
        check(load().qed3_cg_kernel_times(self._h, d_x.ptr, d_b.ptr, C.byref(opts), C.byref(rep), ms))
        return rep, [ms[0], ms[1], ms[2]]

    def solve(self, b: np.ndarray, **kw):
        """x with D^dag D x = b (host vectors); returns (x, report)."""
        b = np.ascontiguousarray(b, np.complex128)
        d_b = DeviceArray.from_numpy(b)
        d_x = DeviceArray(b.nbytes)
        rep = self.solve_device(d_x, d_b, **kw)
        return d_x.to_numpy(np.complex128), rep

    def dot(self, d_x, d_y):
        re, im = C.c_double(), C.c_double()
        check(load().qed3_dot(self._h, d_x.ptr, d_y.ptr, C.byref(re), C.byref(im)))
        return complex(re.value, im.value)

    def pf_action(self, d_phi, d_eta):
        """S_f = Re <phi, eta> of device fields (PseudoFermion::S)."""
        S = C.c_double()
        check(load().qed3_pf_action(self._h, d_phi.ptr, d_eta.ptr, C.byref(S)))
        return S.value

    def force(self, eta: np.ndarray):
        """Pseudofermion force (spatial[Lt][E], temporal[Lt][V]) from eta (host)."""
        d_eta = DeviceArray.from_numpy(np.ascontiguousarray(eta, np.complex128))
        d_fs = DeviceArray(8 * self.Lt * self.E)
        d_ft = DeviceArray(8 * self.Lt * self.V)
        check(load().qed3_pf_force(self._h, d_eta.ptr, d_fs.ptr, d_ft.ptr))
        return (d_fs.to_numpy(np.float64).reshape(self.Lt, self.E),
                d_ft.to_numpy(np.float64).reshape(self.Lt, self.V))

    def gauge_action(self, spatial, temporal, gsq, compact=True, force=False):
        sp = np.ascontiguousarray(spatial, np.float64)
        tp = np.ascontiguousarray(temporal, np.float64)
        S = C.c_double()
        fs = np.zeros((self.Lt, self.E)) if force else None
        ft = np.zeros((self.Lt, self.V)) if force else None
        check(load().qed3_gauge_action(self._h, gsq, int(compact), _dp(sp), _dp(tp), C.byref(S),
                                       _dp(fs) if force else None, _dp(ft) if force else None))
        return (S.value, fs, ft) if force else S.value

    def autotune(self):
        check(load().qed3_autotune(self._h))

    def set_tile(self, block=256, ts=4, remap=1, lds=0, ops=(OP_D, OP_DDAG), precs=(F64, F32)):
        arr = (C.c_int * 4)(block, ts, remap, lds)
        for op in ops:
            for prec in precs:
                check(load().qed3_set_tile(self._h, op, prec, arr))

    def tile(self, op=OP_D, prec=F32):
        out = (C.c_int * 4)()
        check(load().qed3_tile_info(self._h, op, prec, out))
        return dict(block=out[0], ts=out[1], remap=out[2], lds=out[3])

    def time_apply(self, op=OP_D, prec=F64, warmup=3, reps=20):
        ms = C.c_double()
        check(load().qed3_time_apply(self._h, op, prec, warmup, reps, C.byref(ms)))
        return ms.value


def batch_solve(ops, xs, bs, tol=1e-10, maxiter=100000, delta=0.01, check_every=0,
                raise_on_fail=True):
    """Batched half-storage CG over WilsonDirac systems of any shapes; xs/bs DeviceArrays."""
    n = len(ops)
    ctxs = (C.c_void_p * n)(*[o._h for o in ops])
    dx = (C.c_void_p * n)(*[x.ptr for x in xs])
    db = (C.c_void_p * n)(*[b.ptr for b in bs])
    opts = CGOpts(tol, maxiter, CG_MIXED, delta, check_every)
    reps = (CGReport * n)()
    code = load().qed3_batch_cg(ctxs, n, dx, db, C.byref(opts), reps)
    if code != 0 and raise_on_fail:
        check(code)
    return list(reps)


class Net:
    """Learned gauge-covariant preconditioner on the lattice of a WilsonDirac."""

    def __init__(self, D: WilsonDirac, layers=4, seed=1):
        self.D = D
        self._h = C.c_void_p()
        check(load().qed3_net_create(D._h, layers, seed, C.byref(self._h)))
        n = C.c_int()
        check(load().qed3_net_nparams(self._h, C.byref(n)))
        self.nparams = n.value
        self.layers = layers

    def __del__(self):
        try:
            if self._h:
                load().qed3_net_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass

    @property
    def params(self):
        p = np.zeros(self.nparams, np.complex128)
        check(load().qed3_net_get_params(self._h, _dp(p.view(np.float64))))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.complex128)
        assert p.size == self.nparams
        check(load().qed3_net_set_params(self._h, _dp(p.view(np.float64))))

    def apply(self, r):
        """N(r) for a host array of shape [B, N] (or [N]) complex."""
        r = np.ascontiguousarray(r, np.complex64)
        B = 1 if r.ndim == 1 else r.shape[0]
        d_in = DeviceArray.from_numpy(r)
        d_out = DeviceArray(r.nbytes)
        check(load().qed3_net_apply(self._h, d_out.ptr, d_in.ptr, B))
        return d_out.to_numpy(np.complex64).reshape(r.shape)

    def apply_device(self, d_out, d_in, batch=1):
        check(load().qed3_net_apply(self._h, d_out.ptr, d_in.ptr, batch))

    def apply_adj_device(self, d_out, d_in, batch=1):
        check(load().qed3_net_apply_adjoint(self._h, d_out.ptr, d_in.ptr, batch))

    def apply_adj(self, r):
        """N^dagger(r) for a host array of shape [B, N] (or [N]) complex."""
        r = np.ascontiguousarray(r, np.complex64)
        B = 1 if r.ndim == 1 else r.shape[0]
        d_in = DeviceArray.from_numpy(r)
        d_out = DeviceArray(r.nbytes)
        check(load().qed3_net_apply_adjoint(self._h, d_out.ptr, d_in.ptr, B))
        return d_out.to_numpy(np.complex64).reshape(r.shape)

    def loss_grad(self, r, e):
        r = np.ascontiguousarray(r, np.complex64)
        e = np.ascontiguousarray(e, np.complex64)
        B = r.shape[0]
        d_r, d_e = DeviceArray.from_numpy(r), DeviceArray.from_numpy(e)
        loss = C.c_double()
        g = np.zeros(self.nparams, np.complex128)
        check(load().qed3_net_loss_grad(self._h, d_r.ptr, d_e.ptr, B, C.byref(loss),
                                        _dp(g.view(np.float64))))
        return loss.value, g

    @property
    def skip(self):
        c = C.c_double()
        check(load().qed3_net_get_skip(self._h, C.byref(c)))
        return c.value

    @skip.setter
    def skip(self, c0):
        check(load().qed3_net_set_skip(self._h, float(c0)))

    def train_device(self, d_r, d_e, npairs, batch=4, epochs=10, lr=1e-2):
        hist = np.zeros(epochs)
        check(load().qed3_net_train(self._h, d_r.ptr, d_e.ptr, npairs, batch, epochs, lr,
                                    _dp(hist)))
        return hist

    def fcg(self, d_x, d_b, tol=1e-10, maxiter=20000, delta=0.01, raise_on_fail=True):
        opts = CGOpts(tol, maxiter, CG_MIXED, delta, 0)
        rep = CGReport()
        pa, fb = C.c_int(), C.c_int()
        code = load().qed3_fcg(self.D._h, self._h, d_x.ptr, d_b.ptr, C.byref(opts), C.byref(rep),
                               C.byref(pa), C.byref(fb))
        if code != 0 and raise_on_fail:
            check(code)
        return rep, pa.value, fb.value


class Precond:
    """Time-exact learned preconditioner K (Horner message-passing net over the exact inverse
    of the operator's time part) on a WilsonDirac with r = 1, and flexible CG with
    M = K K^dagger: the metric's learned method."""

    def __init__(self, D: WilsonDirac, layers=4):
        self.D = D
        self.layers = layers
        self._h = C.c_void_p()
        check(load().qed3_precond_create(D._h, layers, C.byref(self._h)))
        n = C.c_int()
        check(load().qed3_precond_nparams(self._h, C.byref(n)))
        self.nparams = n.value

    def __del__(self):
        try:
            if self._h:
                load().qed3_precond_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass

    @property
    def params(self):
        """[c_0 .. c_{L-1}, A0, A1, B0, B1] (complex)"""
        p = np.zeros(self.nparams, np.complex128)
        check(load().qed3_precond_get_params(self._h, _dp(p.view(np.float64))))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.complex128)
        assert p.size == self.nparams
        check(load().qed3_precond_set_params(self._h, _dp(p.view(np.float64))))

    def apply_device(self, d_out, d_in, adjoint=False):
        check(load().qed3_precond_apply(self._h, OP_DDAG if adjoint else OP_D, d_out.ptr, d_in.ptr))

    def apply_solver(self, v, adjoint=False, both=False):
        """K v (K^dagger v; with both, K K^dagger v) by the solver's kernels (binary16,
        temporal gauge)."""
        v = np.ascontiguousarray(v, np.complex128)
        d_in = DeviceArray.from_numpy(v)
        d_out = DeviceArray(v.nbytes)
        op = OP_DDDAG if both else (OP_DDAG if adjoint else OP_D)
        check(load().qed3_precond_apply_solver(self._h, op, d_out.ptr, d_in.ptr))
        return d_out.to_numpy(np.complex128)

    def apply(self, v, adjoint=False):
        """K v (or K^dagger v) for a host complex field."""
        v = np.ascontiguousarray(v, np.complex128)
        d_in = DeviceArray.from_numpy(v)
        d_out = DeviceArray(v.nbytes)
        self.apply_device(d_out, d_in, adjoint)
        return d_out.to_numpy(np.complex128)

    def loss_grad_device(self, d_v, d_e, batch):
        loss = C.c_double()
        g = np.zeros(self.nparams, np.complex128)
        check(load().qed3_precond_loss_grad(self._h, d_v.ptr, d_e.ptr, batch, C.byref(loss),
                                            _dp(g.view(np.float64))))
        return loss.value, g

    def loss_grad(self, v, e):
        """mean_b |K v_b - e_b|^2/|e_b|^2 and its gradient for host arrays [B, N]."""
        v = np.ascontiguousarray(v, np.complex128)
        e = np.ascontiguousarray(e, np.complex128)
        return self.loss_grad_device(DeviceArray.from_numpy(v), DeviceArray.from_numpy(e), v.shape[0])

    def train_device(self, d_v, d_e, npairs, batch=4, epochs=10, lr=1e-2):
        hist = np.zeros(epochs)
        check(load().qed3_precond_train(self._h, d_v.ptr, d_e.ptr, npairs, batch, epochs, lr,
                                        _dp(hist)))
        return hist

    def solve_device(self, d_x, d_b, tol=1e-10, maxiter=20000, delta=0.01, check_every=0,
                     raise_on_fail=True):
        opts = CGOpts(tol, maxiter, CG_HALF, delta, check_every)
        rep = CGReport()
        code = load().qed3_pcg(self.D._h, self._h, d_x.ptr, d_b.ptr, C.byref(opts), C.byref(rep))
        if code != 0 and raise_on_fail:
            check(code)
        return rep

    def solve(self, b, **kw):
        b = np.ascontiguousarray(b, np.complex128)
        d_b = DeviceArray.from_numpy(b)
        d_x = DeviceArray(b.nbytes)
        rep = self.solve_device(d_x, d_b, **kw)
        return d_x.to_numpy(np.complex128), rep

    def solve_kernel_times(self, d_x, d_b, tol=1e-10, maxiter=20000, delta=0.01):
        """One solve with HIP-event timing: (report, [K1, K2, 0 (no K2 updates x), first
        layer, K^dagger transport layer, K transport layer, K^dagger's last layer fused with
        K's first, K's last layer with the reductions] ms per launch)."""
        opts = CGOpts(tol, maxiter, CG_HALF, delta, 0)
        rep = CGReport()
        ms = (C.c_double * 8)()
        check(load().qed3_pcg_kernel_times(self.D._h, self._h, d_x.ptr, d_b.ptr, C.byref(opts),
                                           C.byref(rep), ms))
        return rep, list(ms)


class HMC:
    """HMC2 + ExplicitLeapfrogML over a WilsonDirac (host FP64 H and accept step)."""

    def __init__(self, D: WilsonDirac, gsq, compact=True, npf=1, tau=1.0, nsteps=20,
                 nsteps_inner=10, tol=1e-10, solver=CG_MIXED, seed=1):
        self.D = D
        self.p = _lib.HMCParams(gsq, int(compact), npf, tau, nsteps, nsteps_inner, tol, solver,
                                seed)
        self._h = C.c_void_p()
        check(load().qed3_hmc_create(D._h, C.byref(self.p), C.byref(self._h)))

    def __del__(self):
        try:
            if self._h:
                load().qed3_hmc_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass

    def set_overlap(self, ov, refresh=False):
        """Overlap fermions (hmc.cu with IS_OVERLAP): pseudofermions, solves and forces of
        `ov` (an Overlap on this handle's WilsonDirac); refresh recomputes its spectral
        bounds and Zolotarev coefficients after every gauge update as Overlap::update does.
        None returns to Wilson fermions."""
        check(load().qed3_hmc_set_overlap(self._h, ov._h if ov is not None else None, int(refresh)))
        self._ov = ov  # keeps the overlap handle alive while the HMC uses it

    def rng_save(self, path):
        check(load().qed3_hmc_rng_save(self._h, str(path).encode()))

    def rng_load(self, path):
        check(load().qed3_hmc_rng_load(self._h, str(path).encode()))

    def set_steps(self, tau, nsteps, nsteps_inner):
        """Trajectory length and leapfrog steps of the following trajectories."""
        check(load().qed3_hmc_set_steps(self._h, float(tau), int(nsteps), int(nsteps_inner)))

    def trajectory(self, spatial, temporal, no_reject=False):
        """Runs one trajectory; spatial/temporal (float64 arrays) are updated in place."""
        assert spatial.dtype == np.float64 and spatial.flags.c_contiguous
        assert temporal.dtype == np.float64 and temporal.flags.c_contiguous
        rep = _lib.HMCReport()
        check(load().qed3_hmc_trajectory(self._h, _dp(spatial), _dp(temporal), int(no_reject),
                                         C.byref(rep)))
        return {k: getattr(rep, k) for k, _ in rep._fields_}


# ---- checkpoint and resume (gauge_ext.h:225-274, rng.h:131-168, hmc.cu:423-447) ----------
def gauge_write(path, spatial, temporal):
    """Raw FP64 gauge file: spatial[Lt][E] then temporal[Lt][V]."""
    sp = np.ascontiguousarray(spatial, np.float64)
    tp = np.ascontiguousarray(temporal, np.float64)
    Lt, E = sp.shape
    check(load().qed3_gauge_write(str(path).encode(), Lt, E, tp.shape[1], _dp(sp), _dp(tp)))


def gauge_read(path, Lt, E, V):
    sp = np.zeros((Lt, E))
    tp = np.zeros((Lt, V))
    check(load().qed3_gauge_read(str(path).encode(), Lt, E, V, _dp(sp), _dp(tp)))
    return sp, tp


class Rng:
    """ParallelRngExt streams on the host (master, Lt*E link streams, Lt*V site streams)."""

    MASTER, LINK, SITE = 0, 1, 2

    def __init__(self, Lt, E, V, seed):
        self._h = C.c_void_p()
        check(load().qed3_rng_create(Lt, E, V, seed, C.byref(self._h)))

    def __del__(self):
        try:
            if self._h:
                load().qed3_rng_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass

    def gaussian(self, which=0, s=0, i=0):
        out = C.c_double()
        check(load().qed3_rng_gaussian(self._h, which, s, i, C.byref(out)))
        return out.value

    def save(self, path):
        check(load().qed3_rng_save(self._h, str(path).encode()))

    def load(self, path):
        check(load().qed3_rng_load(self._h, str(path).encode()))


def checkpoint_paths(directory, k):
    """The reference's checkpoint names for trajectory k (hmc.cu:431-432)."""
    import os
    return (os.path.join(directory, f"ckpoint_lat.{k}"), os.path.join(directory, f"ckpoint_rng.{k}"))


def resume_point(directory, k_ckpoint=10, kmax=100000):
    """Largest k = k_ckpoint, 2 k_ckpoint, ... with both checkpoint files present and every
    earlier multiple present too (the scan of hmc.cu:427-439 stops at the first gap); 0 when
    there is none, i.e. start from scratch."""
    import os
    k = k_ckpoint
    while k <= kmax:
        lat, rng = checkpoint_paths(directory, k)
        if not (os.path.exists(lat) and os.path.exists(rng)):
            break
        k += k_ckpoint
    return k - k_ckpoint


# ---- multi-shift CG and the overlap operator (overlap.h:7-465) -----------------------------
def zolotarev(n, k):
    """Coefficients of eps(x) = C x (1 + sum A_m/(x^2 - k^2/cp_m)): dict with A, c, cp
    (index 0 unused), M, lambda_inv, E, C, Delta and the shifts sigma_m = -k^2/cp_m."""
    size = n // 2 + 1
    A, c, cp, consts = np.zeros(size), np.zeros(size), np.zeros(size), np.zeros(5)
    sz = C.c_int()
    check(load().qed3_zolotarev(n, k, C.byref(sz), _dp(A), _dp(c), _dp(cp), _dp(consts)))
    out = dict(A=A, c=c, cp=cp, M=consts[0], lambda_inv=consts[1], E=consts[2], C=consts[3],
               Delta=consts[4], k=k, n=n)
    out["shifts"] = -k * k / cp[1:]
    return out


def multishift_solve(D, b, shifts, tol=1e-10, maxiter=100000):
    """x_i with (D^dag D + shifts[i]) x_i = b (host arrays); returns (list of x_i, reports)."""
    b = np.ascontiguousarray(b, np.complex128)
    d_b = DeviceArray.from_numpy(b)
    d_x = [DeviceArray(b.nbytes) for _ in shifts]
    ptrs = (C.c_void_p * len(shifts))(*[d.ptr for d in d_x])
    sh = np.ascontiguousarray(shifts, np.float64)
    reps = (CGReport * len(shifts))()
    opts = CGOpts(tol, maxiter, CG_F64, 0.1, 0)
    check(load().qed3_multishift_cg(D._h, len(shifts), _dp(sh), ptrs, d_b.ptr, C.byref(opts), reps))
    return [d.to_numpy(np.complex128) for d in d_x], list(reps)


def spectral_bounds(D, tol=1e-8, maxiter=2000):
    """(mu_min, mu_max) of D^dag D."""
    lo, hi = C.c_double(), C.c_double()
    check(load().qed3_spectral_bounds(D._h, tol, maxiter, C.byref(lo), C.byref(hi)))
    return lo.value, hi.value


class Overlap:
    """D_ov = 1 + X eps(X^dag X), X = D/lambda_max, Zolotarev degree n (Overlap of
    overlap.h:137-361).  Without explicit bounds they come from the spectrum of D^dag D
    with the reference's margins: lambda_min = sqrt((1 - 100 tol) mu_min),
    lambda_max = sqrt((1 + 100 tol) mu_max) (overlap.h:298-301)."""

    def __init__(self, D, n=21, lambda_min=None, lambda_max=None, tol=1e-10, bound_tol=1e-8):
        self.D = D
        if lambda_min is None or lambda_max is None:
            mu_min, mu_max = spectral_bounds(D, bound_tol)
            lambda_min = np.sqrt((1 - 100 * bound_tol) * mu_min)
            lambda_max = np.sqrt((1 + 100 * bound_tol) * mu_max)
        self.lambda_min, self.lambda_max, self.n = lambda_min, lambda_max, n
        self._h = C.c_void_p()
        check(load().qed3_overlap_create(D._h, n, lambda_min, lambda_max, tol, C.byref(self._h)))

    def __del__(self):
        try:
            if self._h:
                load().qed3_overlap_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass

    def apply(self, v, dag=False):
        v = np.ascontiguousarray(v, np.complex128)
        d_in = DeviceArray.from_numpy(v)
        d_out = DeviceArray(v.nbytes)
        check(load().qed3_overlap_apply(self._h, OP_DDAG if dag else OP_D, d_out.ptr, d_in.ptr))
        return d_out.to_numpy(np.complex128)

    def force(self, eta):
        """Overlap pseudofermion force -2 Re <eta, dD_ov/dtheta eta> (spatial[Lt][E],
        temporal[Lt][V]) with lambda_max and the Zolotarev coefficients fixed
        (Overlap::grad_deviceAsyncLaunch, overlap.h:393-465)."""
        D = self.D
        d_eta = DeviceArray.from_numpy(np.ascontiguousarray(eta, np.complex128))
        d_fs = DeviceArray(8 * D.Lt * D.E)
        d_ft = DeviceArray(8 * D.Lt * D.V)
        check(load().qed3_overlap_force(self._h, d_eta.ptr, d_fs.ptr, d_ft.ptr))
        return (d_fs.to_numpy(np.float64).reshape(D.Lt, D.E),
                d_ft.to_numpy(np.float64).reshape(D.Lt, D.V))

    def solve(self, b, tol=1e-10, maxiter=10000):
        """x = (D_ov + D_ov^dag)^{-1} b by FP64 CG (PseudoFermion<Overlap>::update_eta);
        returns (x, report)."""
        b = np.ascontiguousarray(b, np.complex128)
        d_b = DeviceArray.from_numpy(b)
        d_x = DeviceArray(b.nbytes)
        rep = CGReport()
        check(load().qed3_overlap_solve(self._h, d_x.ptr, d_b.ptr, tol, maxiter, C.byref(rep)))
        return d_x.to_numpy(np.complex128), rep
