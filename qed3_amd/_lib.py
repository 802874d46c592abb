"""ctypes binding of libqed3.so (include/qed3.h)."""
import ctypes as C
import os
import pathlib

_HERE = pathlib.Path(__file__).resolve().parent
# QED3_LIB selects another build of the same sources (the sanitizer build of the CPU tests)
LIB_PATH = pathlib.Path(os.environ.get("QED3_LIB") or _HERE / "libqed3.so")

c_int_p = C.POINTER(C.c_int)
c_double_p = C.POINTER(C.c_double)


class Params(C.Structure):
    _fields_ = [("lattice", C.c_int), ("n_refine", C.c_int), ("Lt", C.c_int),
                ("bc_time", C.c_int), ("r", C.c_double), ("M5", C.c_double),
                ("at", C.c_double)]


class CGOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("maxiter", C.c_int), ("solver", C.c_int),
                ("delta", C.c_double), ("check_every", C.c_int)]


class CGReport(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("reliable_updates", C.c_int),
                ("true_resid", C.c_double), ("seconds", C.c_double)]


class HMCParams(C.Structure):
    _fields_ = [("gsq", C.c_double), ("compact", C.c_int), ("npf", C.c_int), ("tau", C.c_double),
                ("nsteps", C.c_int), ("nsteps_inner", C.c_int), ("tol", C.c_double),
                ("solver", C.c_int), ("seed", C.c_int)]


class HMCReport(C.Structure):
    _fields_ = [("dH", C.c_double), ("h0", C.c_double), ("h1", C.c_double), ("accept", C.c_int),
                ("cg_solves", C.c_int), ("cg_iterations", C.c_longlong),
                ("max_resid", C.c_double), ("seconds", C.c_double)]


# name -> (restype, argtypes); every symbol declared in include/qed3.h
SIGNATURES = {
    "qed3_last_error": (C.c_char_p, []),
    "qed3_malloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "qed3_free": (C.c_int, [C.c_void_p]),
    "qed3_memcpy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    "qed3_device_synchronize": (C.c_int, []),
    "qed3_geom_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "qed3_geom_destroy": (None, [C.c_void_p]),
    "qed3_geom_counts": (C.c_int, [C.c_void_p, c_int_p, c_int_p, c_int_p]),
    "qed3_geom_export": (C.c_int, [C.c_void_p, c_double_p, c_int_p, c_int_p, c_int_p, c_int_p,
                                   c_double_p, c_int_p, c_int_p, c_int_p, c_int_p, c_double_p,
                                   c_double_p, c_double_p, c_double_p]),
    "qed3_geom_import": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, c_double_p, c_int_p,
                                   c_int_p, c_int_p, c_int_p, c_double_p, c_int_p, c_int_p,
                                   c_int_p, c_int_p, c_double_p, c_double_p, c_double_p,
                                   c_double_p, C.POINTER(C.c_void_p)]),
    "qed3_ctx_create_geom": (C.c_int, [C.POINTER(Params), C.c_void_p, C.POINTER(C.c_void_p)]),
    "qed3_random_gauge": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, c_double_p,
                                    c_double_p]),
    "qed3_random_source": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, c_double_p]),
    "qed3_ctx_create": (C.c_int, [C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "qed3_ctx_destroy": (None, [C.c_void_p]),
    "qed3_ctx_sizes": (C.c_int, [C.c_void_p, c_int_p, c_int_p, c_int_p, c_double_p]),
    "qed3_ctx_geometry": (C.c_int, [C.c_void_p, c_double_p, c_double_p, c_double_p, c_double_p,
                                    c_double_p]),
    "qed3_set_gauge": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "qed3_apply": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "qed3_apply_async": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "qed3_ctx_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "qed3_cg": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(CGOpts),
                          C.POINTER(CGReport)]),
    "qed3_cg_kernel_times": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(CGOpts),
                                       C.POINTER(CGReport), C.POINTER(C.c_double)]),
    "qed3_cg_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(CGOpts),
                                  C.POINTER(CGReport)]),
    "qed3_dot": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, c_double_p, c_double_p]),
    "qed3_pf_action": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, c_double_p]),
    "qed3_batch_cg": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p),
                                C.POINTER(C.c_void_p), C.POINTER(CGOpts), C.POINTER(CGReport)]),
    "qed3_net_create": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "qed3_net_destroy": (None, [C.c_void_p]),
    "qed3_net_nparams": (C.c_int, [C.c_void_p, c_int_p]),
    "qed3_net_get_params": (C.c_int, [C.c_void_p, c_double_p]),
    "qed3_net_set_params": (C.c_int, [C.c_void_p, c_double_p]),
    "qed3_net_apply_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "qed3_net_get_skip": (C.c_int, [C.c_void_p, c_double_p]),
    "qed3_net_set_skip": (C.c_int, [C.c_void_p, C.c_double]),
    "qed3_net_apply": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "qed3_net_loss_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, c_double_p,
                                     c_double_p]),
    "qed3_net_train": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                 C.c_double, c_double_p]),
    "qed3_precond_create": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "qed3_precond_destroy": (None, [C.c_void_p]),
    "qed3_precond_nparams": (C.c_int, [C.c_void_p, c_int_p]),
    "qed3_precond_get_params": (C.c_int, [C.c_void_p, c_double_p]),
    "qed3_precond_set_params": (C.c_int, [C.c_void_p, c_double_p]),
    "qed3_precond_apply": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "qed3_precond_apply_solver": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "qed3_precond_loss_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, c_double_p,
                                         c_double_p]),
    "qed3_precond_train": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                     C.c_double, c_double_p]),
    "qed3_pcg": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(CGOpts),
                           C.POINTER(CGReport)]),
    "qed3_pcg_kernel_times": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.POINTER(CGOpts), C.POINTER(CGReport), c_double_p]),
    "qed3_fcg": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(CGOpts),
                           C.POINTER(CGReport), c_int_p, c_int_p]),
    "qed3_pf_force": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "qed3_gauge_action": (C.c_int, [C.c_void_p, C.c_double, C.c_int, c_double_p, c_double_p,
                                    c_double_p, c_double_p, c_double_p]),
    "qed3_hmc_create": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "qed3_hmc_destroy": (None, [C.c_void_p]),
    "qed3_hmc_set_overlap": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "qed3_hmc_trajectory": (C.c_int, [C.c_void_p, c_double_p, c_double_p, C.c_int, C.c_void_p]),
    "qed3_hmc_set_steps": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_int]),
    "qed3_hmc_rng_save": (C.c_int, [C.c_void_p, C.c_char_p]),
    "qed3_hmc_rng_load": (C.c_int, [C.c_void_p, C.c_char_p]),
    "qed3_plaquette_correlator": (C.c_int, [C.c_void_p, C.c_void_p, c_double_p, c_double_p]),
    "qed3_zolotarev": (C.c_int, [C.c_int, C.c_double, c_int_p, c_double_p, c_double_p, c_double_p,
                                 c_double_p]),
    "qed3_multishift_cg": (C.c_int, [C.c_void_p, C.c_int, c_double_p, C.POINTER(C.c_void_p),
                                     C.c_void_p, C.c_void_p, C.c_void_p]),
    "qed3_spectral_bounds": (C.c_int, [C.c_void_p, C.c_double, C.c_int, c_double_p, c_double_p]),
    "qed3_overlap_create": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.POINTER(C.c_void_p)]),
    "qed3_overlap_destroy": (None, [C.c_void_p]),
    "qed3_overlap_apply": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "qed3_overlap_force": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "qed3_overlap_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int,
                                     C.c_void_p]),
    "qed3_gauge_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, c_double_p, c_double_p]),
    "qed3_gauge_read": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, c_double_p, c_double_p]),
    "qed3_rng_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "qed3_rng_destroy": (None, [C.c_void_p]),
    "qed3_rng_save": (C.c_int, [C.c_void_p, C.c_char_p]),
    "qed3_rng_load": (C.c_int, [C.c_void_p, C.c_char_p]),
    "qed3_rng_gaussian": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, c_double_p]),
    "qed3_autotune": (C.c_int, [C.c_void_p]),
    "qed3_time_apply": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, c_double_p]),
    "qed3_time_copy": (C.c_int, [C.c_size_t, C.c_int, c_double_p]),
    "qed3_tile_info": (C.c_int, [C.c_void_p, C.c_int, C.c_int, c_int_p]),
    "qed3_set_tile": (C.c_int, [C.c_void_p, C.c_int, C.c_int, c_int_p]),
}

_lib = None


def load():
    """Load libqed3.so; raises if it has not been built (no fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m qed3_amd.build`")
    lib = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class Qed3Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qed3 error {code}: {msg}")
        self.code = code


def check(code):
    if code != 0:
        raise Qed3Error(code, load().qed3_last_error().decode())
    return code
