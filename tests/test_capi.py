"""The C ABI library loads and exports every function include/qed3.h declares."""
import ctypes as C
import pathlib
import re
import subprocess

import qed3_amd
from qed3_amd import _lib

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(qed3_\w+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported():
    lib = _lib.load()
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (qed3_\w+)", out))
    assert names <= exported


def test_python_binding_covers_the_header():
    assert declared() == set(_lib.SIGNATURES)


def test_argument_errors_without_device():
    lib = _lib.load()
    assert lib.qed3_geom_create(0, C.byref(C.c_void_p())) == 1
    assert lib.qed3_ctx_create(None, None) == 1
    p = _lib.Params(7, 2, 4, 0, 1.0, 0.1, 0.2)
    assert lib.qed3_ctx_create(C.byref(p), C.byref(C.c_void_p())) == 1
    assert b"lattice" in lib.qed3_last_error()
    assert lib.qed3_memcpy(None, None, 0, 9) == 1


def test_argument_errors_of_the_host_entry_points():
    """Entry points that need no device reject bad arguments with QED3_ERR_ARG (1)."""
    lib = _lib.load()
    r = C.c_void_p()
    assert lib.qed3_rng_create(0, 5, 3, 1, C.byref(r)) == 1
    assert lib.qed3_rng_create(2, 5, 3, 1, C.byref(r)) == 0
    out = C.c_double()
    assert lib.qed3_rng_gaussian(r, 1, 2, 0, C.byref(out)) == 1   # slice out of range
    assert lib.qed3_rng_gaussian(r, 2, 0, 3, C.byref(out)) == 1   # site out of range
    assert lib.qed3_rng_gaussian(r, 7, 0, 0, C.byref(out)) == 1   # no such stream kind
    assert lib.qed3_rng_gaussian(r, 1, 1, 4, C.byref(out)) == 0
    assert lib.qed3_rng_load(r, b"/nonexistent/rng") == 1
    lib.qed3_rng_destroy(r)
    assert lib.qed3_gauge_write(b"/tmp/x", 0, 1, 1, None, None) == 1
    assert lib.qed3_gauge_read(b"/nonexistent/lat", 1, 1, 1, None, None) == 1
    size = C.c_int()
    buf = (C.c_double * 16)()
    assert lib.qed3_zolotarev(21, 0.0, C.byref(size), buf, buf, buf, buf) == 1
    assert lib.qed3_zolotarev(0, 0.1, C.byref(size), buf, buf, buf, buf) == 1
    g = C.c_void_p()
    assert lib.qed3_geom_create(1, C.byref(g)) == 0
    p = _lib.Params(5, 1, 2, 0, 1.0, 0.1, 0.2)
    assert lib.qed3_plaquette_correlator(g, C.byref(p), buf, buf) == 1
    lib.qed3_geom_destroy(g)
    assert lib.qed3_overlap_apply(None, 0, None, None) == 1
    assert lib.qed3_overlap_force(None, None, None, None) == 1
    assert lib.qed3_overlap_solve(None, None, None, 1e-10, 10, None) == 1
    assert lib.qed3_hmc_set_overlap(None, None, 0) == 1
    assert lib.qed3_multishift_cg(None, 1, None, None, None, None, None) == 1


def test_seeded_inputs_are_reproducible():
    sp1, tp1 = qed3_amd.random_gauge(0, 2, 3, seed=5, kappa=1.4)
    sp2, tp2 = qed3_amd.random_gauge(0, 2, 3, seed=5, kappa=1.4)
    assert (sp1 == sp2).all() and (tp1 == tp2).all()
    assert abs(sp1).max() <= 3.141592653589793
    xi = qed3_amd.random_source(0, 2, 3, seed=5)
    assert abs(abs(xi).mean() - 0.886) < 0.1  # E|z| for unit complex Gaussian = sqrt(pi)/2
