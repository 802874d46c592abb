"""ORACLE (test infrastructure only): ctypes binding of oracle/liboracle.so, the CPU FP64
restatement of the reference path (see oracle.cpp for the file:line each part follows).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
and only as the checker / CPU baseline.
"""
import ctypes as C
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
# QED3_ORACLE_LIB selects another build of oracle.cpp (the sanitizer build of the CPU tests)
LIB = pathlib.Path(os.environ.get("QED3_ORACLE_LIB") or HERE / "liboracle.so")

_dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
_ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def use_library(path):
    """Load the oracle from another build of oracle.cpp (e.g. the -march=native build the
    CPU baseline times); must be called before the first load()."""
    global LIB
    if _lib is not None:
        raise RuntimeError("oracle library already loaded")
    LIB = pathlib.Path(path)


def build_native(out="/tmp/qed3_oracle_native/liboracle.so"):
    """-O3 -march=native build of oracle.cpp for the host it runs on; returns its path."""
    subprocess.run(["make", "-s", "-C", str(HERE), "native", f"NATIVE_OUT={out}"], check=True)
    return out


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        sig = {
            "oracle_lattice_create": (vp, [C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, ip, ip, ip,
                                           ip, ip, ip, ip, ip, dp, dp, dp, dp]),
            "oracle_lattice_destroy": (None, [vp]),
            "oracle_lattice_checks": (None, [vp, dp]),
            "oracle_lattice_kappa": (None, [vp, dp, dp, dp, dp]),
            "oracle_dirac_create": (vp, [vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]),
            "oracle_dirac_destroy": (None, [vp]),
            "oracle_dirac_nnz": (C.c_long, [vp]),
            "oracle_dirac_update": (None, [vp, dp, dp]),
            "oracle_apply": (None, [vp, C.c_int, dp, dp]),
            "oracle_cg": (C.c_int, [vp, dp, dp, C.c_double, C.c_int, dp]),
            "oracle_dense": (None, [vp, dp]),
            "oracle_csr": (None, [vp, C.POINTER(C.c_long), ip, dp]),
            "oracle_horner_apply": (None, [vp, C.c_int, dp, C.c_int, dp, dp]),
            "oracle_horner_loss_grad": (C.c_double, [vp, C.c_int, dp, C.c_int, dp, dp, dp]),
            "oracle_force": (None, [vp, dp, C.c_int, dp, dp]),
            "oracle_gauge_action_create": (vp, [vp, C.c_int, C.c_double, C.c_double, C.c_int]),
            "oracle_gauge_action_destroy": (None, [vp]),
            "oracle_gauge_action": (C.c_double, [vp, dp, dp]),
            "oracle_gauge_force": (None, [vp, dp, dp, dp, dp]),
            "oracle_mt64": (C.c_uint64, [C.c_uint64, C.c_long]),
            "oracle_mpnn_forward": (None, [vp, C.c_int, C.c_int, C.c_int, dp, C.c_double, dp, dp]),
            "oracle_mpnn_loss_grad": (C.c_double, [vp, C.c_int, C.c_int, C.c_int, dp, C.c_double, dp, dp,
                                                                 dp]),
            "oracle_hmc_trajectory": (C.c_double, [vp, vp, C.c_int, C.c_int, C.c_double, C.c_int,
                                                   C.c_int, C.c_double, dp, dp, dp, ip,
                                                   C.POINTER(C.c_long)]),
        }
        for k, (res, args) in sig.items():
            f = getattr(lib, k)
            f.restype, f.argtypes = res, args
        _lib = lib
    return _lib


class Lattice:
    """S2Simp (lattice=0) or S2Trivalent (lattice=1) built from geometry tables."""

    def __init__(self, geom, lattice: int):
        lib = load()
        g = geom
        self._keep = [np.ascontiguousarray(a) for a in (
            g.sites, g.dual_sites, g.faces, g.links, g.nn_offsets, g.nns, g.dual_links,
            g.dual_nns, g.dual_face_offsets, g.dual_faces, g.alpha, g.omega, g.dual_alpha,
            g.dual_omega)]
        (sites, dsites, faces, links, nno, nns, dlinks, dnns, dfo, dfaces, al, om, dal,
         dom) = self._keep
        self.h = lib.oracle_lattice_create(
            lattice, g.V, g.E, g.F, _dp(sites), _dp(dsites), _ip(faces), _ip(links), _ip(nno),
            _ip(nns), _ip(dlinks), _ip(dnns), _ip(dfo), _ip(dfaces), _dp(al), _dp(om), _dp(dal),
            _dp(dom))
        self.lattice = lattice
        if lattice == 0:
            self.V, self.E, self.F = g.V, g.E, g.F
        else:
            self.V, self.E, self.F = g.F, g.E, g.V

    def checks(self):
        out = np.zeros(7)
        load().oracle_lattice_checks(self.h, _dp(out))
        return dict(sum_face_area=out[0], sum_link_volume=out[1], sum_dual_area=out[2],
                    max_ell_diff=out[3], spin_dev=out[4], deficit_fail=int(out[5]),
                    mean_ell=out[6])

    def derived(self):
        k, ell, lv, da = np.zeros(self.E), np.zeros(self.E), np.zeros(self.E), np.zeros(self.V)
        load().oracle_lattice_kappa(self.h, _dp(k), _dp(ell), _dp(lv), _dp(da))
        return dict(kappa=k, ell=ell, link_volume=lv, dual_area=da)

    def __del__(self):
        try:
            load().oracle_lattice_destroy(self.h)
        except Exception:
            pass


class Dirac:
    """DiracExt COO -> CSR / CSR^H with the host SpMV and MatPoly::solve."""

    def __init__(self, lat: Lattice, Lt: int, r=1.0, M5=0.0, at=1.0, bc=0):
        self.lat, self.Lt = lat, Lt
        self.h = load().oracle_dirac_create(lat.h, Lt, bc, r, M5, at)
        self.N = 2 * lat.V * Lt

    def __del__(self):
        try:
            load().oracle_dirac_destroy(self.h)
        except Exception:
            pass

    def nnz(self):
        return load().oracle_dirac_nnz(self.h)

    def update(self, spatial, temporal):
        self._sp = np.ascontiguousarray(spatial, np.float64)
        self._tp = np.ascontiguousarray(temporal, np.float64)
        load().oracle_dirac_update(self.h, _dp(self._sp), _dp(self._tp))

    def apply(self, op, v):
        """op 0 = D, 1 = D^dagger, 2 = D^dagger D, 3 = D D^dagger."""
        if op not in (0, 1, 2, 3):
            raise ValueError(f"unknown operator {op}")
        v = np.ascontiguousarray(v, np.complex128)
        out = np.zeros_like(v)
        load().oracle_apply(self.h, op, _dp(out.view(np.float64)), _dp(v.view(np.float64)))
        return out

    def solve(self, b, tol=1e-13, maxiter=1000000):
        b = np.ascontiguousarray(b, np.complex128)
        x = np.zeros_like(b)
        mu = C.c_double()
        it = load().oracle_cg(self.h, _dp(x.view(np.float64)), _dp(b.view(np.float64)), tol,
                              maxiter, C.byref(mu))
        return x, it, mu.value

    def csr(self):
        """scipy.sparse CSR matrix of D in the reference's row order."""
        import scipy.sparse as sps
        nnz = self.nnz()
        rows = np.zeros(self.N + 1, np.int64)
        cols = np.zeros(nnz, np.int32)
        vals = np.zeros(nnz, np.complex128)
        load().oracle_csr(self.h, rows.ctypes.data_as(C.POINTER(C.c_long)), _ip(cols),
                          _dp(vals.view(np.float64)))
        return sps.csr_matrix((vals, cols, rows), shape=(self.N, self.N))

    def dense(self):
        out = np.zeros((self.N, self.N), np.complex128)
        load().oracle_dense(self.h, _dp(out.view(np.float64)))
        return out

    def force(self, eta, full=False):
        eta = np.ascontiguousarray(eta, np.complex128)
        fs = np.zeros((self.Lt, self.lat.E))
        ft = np.zeros((self.Lt, self.lat.V))
        load().oracle_force(self.h, _dp(eta.view(np.float64)), int(full), _dp(fs), _dp(ft))
        return fs, ft


class GaugeAction:
    def __init__(self, lat: Lattice, Lt: int, gsq: float, at: float, compact: bool):
        self.h = load().oracle_gauge_action_create(lat.h, Lt, gsq, at, int(compact))
        self.lat, self.Lt = lat, Lt

    def __del__(self):
        try:
            load().oracle_gauge_action_destroy(self.h)
        except Exception:
            pass

    def S(self, sp, tp):
        sp = np.ascontiguousarray(sp, np.float64)
        tp = np.ascontiguousarray(tp, np.float64)
        return load().oracle_gauge_action(self.h, _dp(sp), _dp(tp))

    def force(self, sp, tp):
        sp = np.ascontiguousarray(sp, np.float64)
        tp = np.ascontiguousarray(tp, np.float64)
        fs = np.zeros((self.Lt, self.lat.E))
        ft = np.zeros((self.Lt, self.lat.V))
        load().oracle_gauge_force(self.h, _dp(sp), _dp(tp), _dp(fs), _dp(ft))
        return fs, ft


def mpnn_forward(dirac: "Dirac", C_, L_, params, r, skip=0.0):
    """Net output N(r) + skip r for fields r (shape [B, N] complex) with the Dirac's links."""
    r = np.ascontiguousarray(r, np.complex128)
    B = r.shape[0]
    p = np.ascontiguousarray(params, np.complex128)
    out = np.zeros_like(r)
    load().oracle_mpnn_forward(dirac.h, C_, L_, B, _dp(p.view(np.float64)), float(skip),
                               _dp(r.view(np.float64)), _dp(out.view(np.float64)))
    return out


def mpnn_loss_grad(dirac: "Dirac", C_, L_, params, r, e, skip=0.0):
    r = np.ascontiguousarray(r, np.complex128)
    e = np.ascontiguousarray(e, np.complex128)
    B = r.shape[0]
    p = np.ascontiguousarray(params, np.complex128)
    g = np.zeros_like(p)
    loss = load().oracle_mpnn_loss_grad(dirac.h, C_, L_, B, _dp(p.view(np.float64)), float(skip),
                                        _dp(r.view(np.float64)), _dp(e.view(np.float64)),
                                        _dp(g.view(np.float64)))
    return loss, g


def horner_apply(dirac: "Dirac", layers, params, v, adjoint=False):
    """K v (or K^dagger v) of the time-exact Horner preconditioner, FP64 (dense LU per site)."""
    v = np.ascontiguousarray(v, np.complex128)
    p = np.ascontiguousarray(params, np.complex128)
    out = np.zeros_like(v)
    load().oracle_horner_apply(dirac.h, layers, _dp(p.view(np.float64)), int(adjoint),
                               _dp(v.view(np.float64)), _dp(out.view(np.float64)))
    return out


def horner_loss_grad(dirac: "Dirac", layers, params, v, e):
    """(loss, 2 dloss/dconj(p)) of mean_b |K v_b - e_b|^2/|e_b|^2 for arrays [B, N]."""
    v = np.ascontiguousarray(v, np.complex128)
    e = np.ascontiguousarray(e, np.complex128)
    p = np.ascontiguousarray(params, np.complex128)
    g = np.zeros_like(p)
    loss = load().oracle_horner_loss_grad(dirac.h, layers, _dp(p.view(np.float64)), v.shape[0],
                                          _dp(v.view(np.float64)), _dp(e.view(np.float64)),
                                          _dp(g.view(np.float64)))
    return loss, g


def mt64(seed: int, k: int) -> int:
    return int(load().oracle_mt64(seed, k))


def hmc_trajectory(dirac: Dirac, action: GaugeAction, seed, npf, tau, nsteps, nsteps_inner,
                   tol, sp, tp):
    u = np.ascontiguousarray(sp, np.float64).copy()
    t = np.ascontiguousarray(tp, np.float64).copy()
    hh = np.zeros(2)
    acc = C.c_int()
    ncg = C.c_long()
    dH = load().oracle_hmc_trajectory(dirac.h, action.h, seed, npf, tau, nsteps, nsteps_inner,
                                      tol, _dp(u), _dp(t), _dp(hh), C.byref(acc), C.byref(ncg))
    return dict(dH=dH, h0=hh[0], h1=hh[1], accept=bool(acc.value), ncg=ncg.value, u=u, t=t)
