"""Geometry tables in the reference's .dat formats.

Writes / reads the files the reference loads (src/both_3d/s2n_simp.h:43-200,
dirac_simp.h:12-59, dirac_dual.h:15-61), named and formatted as its generators write them
(integrator/get_geometry.cc:355-435, integrator/test_sorting_simp.cc:445-470,
integrator/test_sorting.cc): coordinates in scientific notation with 25 digits, indices
as integers, alpha/omega with 15 digits.  read_dats() accepts externally produced tables
(e.g. QFE's) and returns an object usable wherever a qed3_amd.Geometry is.
"""
import pathlib
from types import SimpleNamespace

import numpy as np


def _names(n):
    s = f"n{n}_singlepatch.dat"
    return dict(pts=f"pts_{s}", pts_dual=f"pts_dual_{s}", nns=f"nns_{s}",
                nns_dual=f"nns_dual_{s}", links=f"links_{s}", dual_links=f"dual_links_{s}",
                face=f"face_{s}", face_dual=f"face_dual_n{n}.dat", omega=f"omega_{s}",
                alpha=f"alpha_{s}", omega_dual=f"omega_dual_{s}", alpha_dual=f"alpha_dual_{s}")


def write_dats(g, directory):
    """Write the tables of a qed3_amd.Geometry into `directory`."""
    d = pathlib.Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    nm = _names(g.n)

    def rows(path, rows_, fmt):
        with open(d / path, "w") as f:
            for r in rows_:
                f.write(" ".join(fmt(v) for v in r) + " \n")

    real = lambda v: f"{v:50.25e}"  # noqa: E731
    idx = lambda v: f"{int(v):50d}"  # noqa: E731
    rows(nm["pts"], g.sites, real)
    rows(nm["pts_dual"], g.dual_sites, real)
    rows(nm["nns"], [g.nns[g.nn_offsets[x]:g.nn_offsets[x + 1]] for x in range(g.V)], idx)
    rows(nm["nns_dual"], g.dual_nns, idx)
    rows(nm["links"], g.links, idx)
    rows(nm["dual_links"], g.dual_links, idx)
    rows(nm["face"], g.faces, idx)
    rows(nm["face_dual"],
         [g.dual_faces[g.dual_face_offsets[x]:g.dual_face_offsets[x + 1]] for x in range(g.V)], idx)

    def spin(path_o, path_a, links, alpha, omega):
        with open(d / path_o, "w") as f:
            for (i, j), w in zip(links, omega):
                f.write(f"{int(i):25d} {int(j):25d} {w:25.15e}\n")
        with open(d / path_a, "w") as f:
            for (i, j), (a0, a1) in zip(links, alpha):
                f.write(f"{int(i):25d} {int(j):25d} {a0:25.15e}\n")
                f.write(f"{int(j):25d} {int(i):25d} {a1:25.15e}\n")

    spin(nm["omega"], nm["alpha"], g.links, g.alpha, g.omega)
    spin(nm["omega_dual"], nm["alpha_dual"], g.dual_links, g.dual_alpha, g.dual_omega)


def read_dats(directory, n):
    """Read the tables written by write_dats() or by the reference's generators."""
    d = pathlib.Path(directory)
    nm = _names(n)

    def lines(p):
        return [l.split() for l in (d / nm[p]).read_text().splitlines() if l.strip()]

    g = SimpleNamespace(n=n)
    g.sites = np.array([[float(v) for v in r[:3]] for r in lines("pts")])
    g.dual_sites = np.array([[float(v) for v in r[:3]] for r in lines("pts_dual")])
    g.links = np.array([[int(float(v)) for v in r[:2]] for r in lines("links")], np.int32)
    g.dual_links = np.array([[int(float(v)) for v in r[:2]] for r in lines("dual_links")], np.int32)
    g.faces = np.array([[int(float(v)) for v in r[:3]] for r in lines("face")], np.int32)
    g.dual_nns = np.array([[int(float(v)) for v in r[:3]] for r in lines("nns_dual")], np.int32)
    nns = [[int(float(v)) for v in r] for r in lines("nns")]
    rings = [[int(float(v)) for v in r] for r in lines("face_dual")]
    g.V, g.E, g.F = len(g.sites), len(g.links), len(g.faces)
    g.nn_offsets = np.concatenate([[0], np.cumsum([len(r) for r in nns])]).astype(np.int32)
    g.nns = np.array([v for r in nns for v in r], np.int32)
    g.dual_face_offsets = np.concatenate([[0], np.cumsum([len(r) for r in rings])]).astype(np.int32)
    g.dual_faces = np.array([v for r in rings for v in r], np.int32)

    def spin(po, pa, links):
        om = {(int(r[0]), int(r[1])): float(r[2]) for r in lines(po)}
        al = {(int(r[0]), int(r[1])): float(r[2]) for r in lines(pa)}
        omega = np.array([om[(i, j)] if (i, j) in om else -om[(j, i)] for i, j in links])
        alpha = np.array([[al[(i, j)], al[(j, i)]] for i, j in links])
        return alpha, omega

    g.alpha, g.omega = spin("omega", "alpha", g.links)
    g.dual_alpha, g.dual_omega = spin("omega_dual", "alpha_dual", g.dual_links)
    return g
