"""Plaquette correlator (gauge.cu:313-343) and jackknife (analyze_corr.cpp:100-180)."""
import numpy as np
import pytest

import qed3_amd as q
from qed3_amd import observables as ob
from pairs import geometry


def tri_area(a, b, c):
    """Solid angle of a spherical triangle (Van Oosterom & Strackee, IEEE TBME 30 (1983) 125)."""
    num = np.abs(np.einsum("ij,ij->i", a, np.cross(b, c)))
    den = 1 + np.einsum("ij,ij->i", a, b) + np.einsum("ij,ij->i", b, c) + np.einsum("ij,ij->i", c, a)
    return 2 * np.arctan2(num, den)


def faces_and_areas(g, lattice):
    """Oriented site rings of the faces and their areas, from the raw tables."""
    if lattice == 0:
        rings = [list(f) for f in g.faces]
        pts = g.sites
        areas = tri_area(pts[g.faces[:, 0]], pts[g.faces[:, 1]], pts[g.faces[:, 2]])
        links = g.links
    else:
        rings = [list(g.dual_ring(x)) for x in range(g.V)]
        pts = g.dual_sites
        areas = []
        for x, r in enumerate(rings):
            c = np.repeat(g.sites[x][None, :], len(r), 0)
            areas.append(tri_area(c, pts[r], pts[np.roll(r, -1)]).sum())
        areas = np.array(areas)
        links = g.dual_links
    return rings, areas, {(int(a), int(b)): i for i, (a, b) in enumerate(links)}


def correlator_numpy(g, lattice, sp):
    rings, areas, lid = faces_and_areas(g, lattice)
    Lt = sp.shape[0]
    P = np.zeros((Lt, len(rings)))
    for f, r in enumerate(rings):
        for a, b in zip(r, r[1:] + r[:1]):
            if a < b:
                P[:, f] += sp[:, lid[(a, b)]]
            else:
                P[:, f] -= sp[:, lid[(b, a)]]
    w = 1 / areas ** 2
    return np.array([sum((P[s] * P[(s + t) % Lt] * w).sum() + (P[s] * P[(s - t) % Lt] * w).sum()
                         for s in range(Lt)) / (2 * Lt * len(rings)) for t in range(Lt)])


@pytest.mark.parametrize("lattice", [0, 1])
def test_plaquette_correlator_matches_restatement(lattice):
    n, Lt = 3, 6
    g = geometry(n)
    E = g.E
    sp = np.random.default_rng(4).uniform(-np.pi, np.pi, (Lt, E))
    got = ob.plaquette_correlator(lattice, n, Lt, sp)
    ref = correlator_numpy(g, lattice, sp)
    assert np.allclose(got, ref, rtol=0, atol=1e-13 * np.abs(ref).max())
    assert np.allclose(got[1:], got[1:][::-1], rtol=1e-12)  # C(t) = C(Lt - t)


def test_plaquette_correlator_single_link():
    """One link excited at one slice: two faces carry +-theta, so C(0) = theta^2 (1/A1^2 +
    1/A2^2)/(Lt F) and C(t != 0) = 0."""
    n, Lt, th = 2, 4, 0.7
    g = geometry(n)
    sp = np.zeros((Lt, g.E))
    sp[1, 5] = th
    got = ob.plaquette_correlator(0, n, Lt, sp)
    a, b = g.links[5]
    fa = [f for f in range(g.F) if a in g.faces[f] and b in g.faces[f]]
    A = tri_area(*(g.sites[g.faces[fa][:, k]] for k in range(3)))
    assert len(fa) == 2
    assert np.isclose(got[0], th ** 2 * np.sum(1 / A ** 2) / (Lt * g.F), rtol=1e-12)
    assert np.allclose(got[1:], 0, atol=1e-15)


def test_jackknife_of_a_mean_is_the_standard_error():
    x = np.random.default_rng(1).normal(size=(40, 3))
    est, err = ob.jackknife(x)
    assert np.allclose(est, x.mean(0), rtol=1e-13)
    assert np.allclose(err, x.std(0, ddof=1) / np.sqrt(40), rtol=1e-12)
    with pytest.raises(ValueError):
        ob.jackknife(x[:1])


def test_bin_series():
    s = np.arange(10.0)
    assert np.array_equal(ob.bin_series(s, 3), [1.0, 4.0, 7.0])


def test_analyze_runs_over_files(tmp_path):
    """Two runs of correlator files; the second stops earlier, which lowers conf_max."""
    rng = np.random.default_rng(2)
    data = {}
    for prefix, last in ((1, 90), (2, 70)):
        d = tmp_path / f"run{prefix}"
        d.mkdir()
        for conf in range(10, last + 1, 10):
            c = rng.normal(size=4)
            data[(prefix, conf)] = c
            ob.write_correlator(d / f"plaq_ss_t_{conf}.dat", c)
    est, err = ob.analyze_runs(str(tmp_path / "run"), "plaq_ss_t_", 10, 100, 10, 2, 2)
    # conf_max -> 70, nbins = (70-10)/10/2 = 3 per run, bins of confs (10,20),(30,40),(50,60)
    bins = [(data[(p, c0)] + data[(p, c0 + 10)]) / 2 for p in (1, 2) for c0 in (10, 30, 50)]
    e2, r2 = ob.jackknife(np.array(bins))
    assert np.allclose(est, e2) and np.allclose(err, r2)
    assert np.allclose(ob.read_correlator(tmp_path / "run1" / "plaq_ss_t_10.dat"), data[(1, 10)])
