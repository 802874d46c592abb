"""Paired product (libqed3) / oracle set-ups on identical tables, links and sources."""
import numpy as np

import oracle as O
import qed3_amd as q

_geo = {}


def geometry(n):
    if n not in _geo:
        _geo[n] = q.Geometry(n)
    return _geo[n]


def oracle_setup(lattice, n, Lt, M5=0.1, r=1.0, bc=0, at=None):
    g = geometry(n)
    L = O.Lattice(g, lattice)
    mean_ell = L.checks()["mean_ell"]
    if at is None:
        at = 0.5 * mean_ell
    D = O.Dirac(L, Lt, r=r, M5=M5, at=at, bc=bc)
    return L, D, at


def links(lattice, n, Lt, seed=11, kappa=1.4):
    return q.random_gauge(lattice, n, Lt, seed, kappa)


def source(lattice, n, Lt, seed=101):
    return q.random_source(lattice, n, Lt, seed)


def wilson_hop_params(C, r):
    """L=2 net parameters whose output is (D - c_x) r: the input head lifts r into channel 0
    and the output head weighs the feature maps as the hops of dirac_ext.h:121-213 do:
    spatial -r F1 + F2 (the gamma term), forward time (-r + sigma3), backward (-r - sigma3)."""
    v = np.zeros((2, 5, C), complex)
    u = np.zeros((2, 5, C), complex)
    v[:, 0, 0] = 1.0
    u[:, 1, 0] = -r
    u[:, 2, 0] = 1.0
    u[0, 3, 0], u[1, 3, 0] = 1.0 - r, -1.0 - r
    u[0, 4, 0], u[1, 4, 0] = -1.0 - r, 1.0 - r
    return np.concatenate([v.ravel(), u.ravel()])
