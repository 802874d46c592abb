"""Geometry tables: counts, the reference's load-time invariants and the n=2 QFE dump."""
import json
import pathlib

import numpy as np
import pytest

import oracle as O
import qed3_amd as q

GOLD = pathlib.Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 16])
def test_counts_and_euler(n):
    g = q.Geometry(n)
    assert (g.V, g.E, g.F) == (10 * n * n + 2, 30 * n * n, 20 * n * n)
    assert g.V - g.E + g.F == 2
    deg = np.diff(g.nn_offsets)
    assert np.sum(deg == 5) == 12 and np.all((deg == 5) | (deg == 6))
    assert np.all(g.links[:, 0] < g.links[:, 1])
    assert np.allclose(np.linalg.norm(g.sites, axis=1), 1.0, atol=1e-15)
    assert np.allclose(np.linalg.norm(g.dual_sites, axis=1), 1.0, atol=1e-15)
    # faces ccw from outside
    a, b, c = (g.sites[g.faces[:, k]] for k in range(3))
    assert np.all(np.einsum("ij,ij->i", np.cross(b - a, c - a), a + b + c) > 0)


def test_n2_matches_qfe_dump():
    """Sites and neighbour lists of QfeLatticeS2(5, 2) (obsolete/test_geometry.log)."""
    d = json.loads((GOLD / "qfe_n2_geometry.json").read_text())
    R = np.array(d["sites"])
    g = q.Geometry(2)
    dist = np.linalg.norm(R[:, None, :] - g.sites[None, :, :], axis=2)
    m = dist.argmin(axis=1)
    assert dist.min(axis=1).max() < 1e-6          # the dump has 6 significant digits
    assert len(set(m.tolist())) == 42             # a bijection
    for i, nb in enumerate(d["neighbours"]):
        assert set(m[j] for j in nb) == set(g.neighbours(m[i]).tolist())


@pytest.mark.parametrize("n", [1, 2, 3, 4, 8, 12])
@pytest.mark.parametrize("lattice", [0, 1])
def test_reference_invariants(n, lattice):
    """s2n_simp.h:256,358,400 sums; dirac_dual.h:92-101 spin structure; deficits."""
    g = q.Geometry(n)
    c = O.Lattice(g, lattice).checks()
    assert abs(c["sum_face_area"] - 4 * np.pi) < 1e-10
    assert abs(c["sum_link_volume"] - 4 * np.pi) < 1e-10
    assert abs(c["sum_dual_area"] - 4 * np.pi) < 1e-10
    assert c["max_ell_diff"] < 1e-14
    assert c["spin_dev"] < 1e-12
    assert c["deficit_fail"] == 0


def test_deficit_check_detects_missing_branch_cut():
    """Undoing the phi0 = pi/48 cut (omega += 2pi on the cut links) breaks the deficit range."""
    g = q.Geometry(4)
    th = 1.0 / 4
    rot = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]])
    x = g.sites @ rot.T
    phi = np.arctan2(x[:, 1], x[:, 0])
    p1, p2 = phi[g.links[:, 0]], phi[g.links[:, 1]]
    w = lambda a: (a + np.pi) % (2 * np.pi) - np.pi  # noqa: E731
    phi0 = np.pi / 48
    cut = np.abs(w(p1 - phi0)) + np.abs(w(p2 - phi0)) - np.abs(w(p1 - p2)) < 1e-6
    assert cut.sum() > 0
    g.omega = g.omega + 2 * np.pi * cut
    c = O.Lattice(g, 0).checks()
    assert c["deficit_fail"] > 0
    assert c["spin_dev"] < 1e-12  # consistency mod 2pi cannot see it; the deficit can


def test_spin_check_detects_wrong_connection_sign():
    g = q.Geometry(3)
    g.omega = -g.omega
    assert O.Lattice(g, 0).checks()["spin_dev"] > 1e-3


def test_hilbert_order_is_local():
    """Neighbour ids stay close: the median |x - y| is a small fraction of V."""
    g = q.Geometry(32)
    d = np.abs(g.links[:, 0] - g.links[:, 1])
    assert np.median(d) < 0.01 * g.V


@pytest.mark.gpu
def test_derived_quantities_agree_with_oracle(gpu):
    """kappa, ell, link volumes, dual areas: product tables vs the s2n_simp.h restatement."""
    g = q.Geometry(3)
    for lattice in (0, 1):
        L = O.Lattice(g, lattice).derived()
        # the product computes the same quantities on its own (host part of libqed3)
        import ctypes as C
        from qed3_amd import _lib
        lib = _lib.load()
        p = _lib.Params(lattice, 3, 4, 0, 1.0, 0.1, 0.2)
        h = C.c_void_p()
        rc = lib.qed3_ctx_create(C.byref(p), C.byref(h))
        if rc != 0:
            pytest.skip("context creation needs a HIP device")
        V, E = (g.V, g.E) if lattice == 0 else (g.F, g.E)
        k, kt, lv, ell, da = (np.zeros(E), np.zeros(V), np.zeros(E), np.zeros(E), np.zeros(V))
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        lib.qed3_ctx_geometry(h, dp(k), dp(kt), dp(lv), dp(ell), dp(da))
        lib.qed3_ctx_destroy(h)
        assert np.allclose(k, L["kappa"], rtol=1e-13)
        assert np.allclose(ell, L["ell"], rtol=1e-13)
        assert np.allclose(lv, L["link_volume"], rtol=1e-13)
        assert np.allclose(da, L["dual_area"], rtol=1e-13)


def test_dat_roundtrip(tmp_path):
    """Tables written in the reference's .dat formats read back unchanged."""
    from qed3_amd import dat_io
    g = q.Geometry(3)
    dat_io.write_dats(g, tmp_path)
    assert (tmp_path / "pts_n3_singlepatch.dat").exists()
    assert (tmp_path / "face_dual_n3.dat").exists()
    h = dat_io.read_dats(tmp_path, 3)
    for name in ("sites", "dual_sites", "alpha", "omega", "dual_alpha", "dual_omega"):
        assert np.allclose(getattr(h, name), getattr(g, name), rtol=1e-14, atol=1e-15), name
    for name in ("links", "dual_links", "faces", "dual_nns", "nn_offsets", "nns",
                 "dual_face_offsets", "dual_faces"):
        assert np.array_equal(getattr(h, name), getattr(g, name)), name
    # the oracle's S2Simp restatement accepts the re-read tables
    c = O.Lattice(h, 0).checks()
    assert abs(c["sum_face_area"] - 4 * np.pi) < 1e-10 and c["deficit_fail"] == 0


def test_geometry_import_validates():
    import ctypes as C
    from qed3_amd import _lib
    g = q.Geometry(1)
    bad = np.array(g.links).copy()
    bad[0, 0] = 9999
    g.links = bad
    with pytest.raises(_lib.Qed3Error):
        q.import_geometry(g)
