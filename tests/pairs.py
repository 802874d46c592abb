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
