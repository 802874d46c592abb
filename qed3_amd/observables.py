"""Gauge observables and their error analysis (host side).

plaquette_correlator: the plaquette-angle correlator the reference's pure-gauge driver
writes every `interval` trajectories (src/both_3d/gauge.cu:313-343, files
plaq_ss_t_<k>.dat, one value per time separation per line).
jackknife: the binned jackknife of src/both_3d/analyze_corr.cpp:100-140 over such files.
"""
import ctypes as C
import os

import numpy as np

from . import _dp, _lib
from ._lib import check, load


def plaquette_correlator(lattice, n, Lt, spatial, at=1.0):
    """corr[t] = sum_{s,f} [P_f(s) P_f(s+t) + P_f(s) P_f(s-t)] / vol_f^2 / (2 Lt F)."""
    sp = np.ascontiguousarray(spatial, np.float64)
    lib = load()
    g = C.c_void_p()
    check(lib.qed3_geom_create(int(n), C.byref(g)))
    try:
        p = _lib.Params(lattice, n, Lt, 0, 1.0, 0.0, at)
        corr = np.zeros(Lt)
        check(lib.qed3_plaquette_correlator(g, C.byref(p), _dp(sp), _dp(corr)))
    finally:
        lib.qed3_geom_destroy(g)
    return corr


def write_correlator(path, corr):
    """One value per line, as the reference's driver writes them."""
    with open(path, "w") as f:
        for v in corr:
            f.write(repr(float(v)) + "\n")


def read_correlator(path):
    """Corr::readfile (analyze_corr.cpp:33-44): the first number of every line."""
    with open(path) as f:
        return np.array([float(line.split()[0]) for line in f if line.strip()])


def jackknife(bin_averages):
    """Estimate and error from bin averages [nbins, ...] (Jackknife::do_it,
    analyze_corr.cpp:110-138): leave-one-out means, their mean, and the variance
    (nbins - 1)/nbins * sum (jack_i - est)^2."""
    b = np.asarray(bin_averages, np.float64)
    nb = b.shape[0]
    if nb < 2:
        raise ValueError("jackknife needs at least two bins")
    jack = (b.sum(axis=0, keepdims=True) - b) / (nb - 1.0)
    est = jack.mean(axis=0)
    var = ((jack - est) ** 2).sum(axis=0) / nb * (nb - 1.0)
    return est, np.sqrt(var)


def bin_series(series, binsize):
    """Consecutive bins of `binsize` measurements [nmeas, ...] -> [nbins, ...] averages
    (trailing measurements that do not fill a bin are dropped)."""
    s = np.asarray(series, np.float64)
    nb = s.shape[0] // binsize
    return s[: nb * binsize].reshape((nb, binsize) + s.shape[1:]).mean(axis=1)


def analyze_runs(path, header, conf_min, conf_max, interval, binsize, prefix_max):
    """analyze_corr.cpp main: files <path><prefix>/<header><conf>.dat for prefix 1..prefix_max;
    conf_max is lowered to the smallest last existing conf over the runs; every run is
    binned separately and all bins enter one jackknife.  Returns (est, err)."""
    last = conf_max
    for prefix in range(1, prefix_max + 1):
        tmp = 0
        for conf in range(conf_min, conf_max, interval):
            if not os.path.exists(f"{path}{prefix}/{header}{conf}.dat"):
                break
            tmp = conf
        last = min(last, tmp)
    conf_max = last
    nbins = (conf_max - conf_min) // interval // binsize
    bins = []
    for prefix in range(1, prefix_max + 1):
        for ib in range(nbins):
            acc = None
            for conf in range(conf_min + ib * interval * binsize,
                              conf_min + (ib + 1) * interval * binsize, interval):
                c = read_correlator(f"{path}{prefix}/{header}{conf}.dat")
                acc = c if acc is None else acc + c
            bins.append(acc / binsize)
    return jackknife(np.array(bins))
