"""Geometry tables under a random relabelling, as a caller with its own numbering (the
reference's QFE .dat order, s2n_simp.h:43-200) would hand them over: sites, triangles and
links renumbered, link orientations (simplicial and dual) reversed at random with alpha
swapped and omega negated, triangle vertices and dual rings rotated cyclically (orientation
kept), neighbour lists shuffled.  `Relabel` maps fields of the original numbering to the
new one."""
import types

import numpy as np


class Relabel:
    """Maps of one lattice (0 simplicial, 1 dual) between the original and the new tables."""

    def __init__(self, site_new, link_new, link_sign, Lt):
        self.site_new, self.link_new, self.link_sign, self.Lt = site_new, link_new, link_sign, Lt

    def spinor(self, v):
        """v[s, x, a] (reference layout i = 2 V s + 2 x + a) in the new site numbering."""
        V = len(self.site_new)
        v = np.asarray(v).reshape(self.Lt, V, 2)
        out = np.empty_like(v)
        out[:, self.site_new, :] = v
        return out.reshape(-1)

    def spatial(self, sp):
        """Link angles or link forces [Lt][E]: moved to the new link and negated where the
        orientation was reversed."""
        out = np.empty_like(sp)
        out[:, self.link_new] = sp * self.link_sign[None, :]
        return out

    def temporal(self, tp):
        out = np.empty_like(tp)
        out[:, self.site_new] = tp
        return out


def renumber(g, seed, wrong=None):
    """(tables, relabel(lattice, Lt)) for geometry g.  wrong="alpha" keeps alpha unswapped
    on reversed links (a planted mistake the parity checks must catch)."""
    rng = np.random.default_rng(seed)
    V, E, F = g.V, g.E, g.F
    sig = rng.permutation(V)  # new id of each site
    phi = rng.permutation(F)  # new id of each triangle (= dual site)
    lam = rng.permutation(E)  # new position of each link (simplicial and dual share it)
    flip = rng.random(E) < 0.5
    dflip = rng.random(E) < 0.5
    h = types.SimpleNamespace(n=g.n, V=V, E=E, F=F)
    h.sites = np.empty_like(g.sites)
    h.sites[sig] = g.sites
    h.dual_sites = np.empty_like(g.dual_sites)
    h.dual_sites[phi] = g.dual_sites
    h.faces = np.empty_like(g.faces)
    rot = rng.integers(0, 3, F)
    for f in range(F):
        h.faces[phi[f]] = sig[np.roll(g.faces[f], rot[f])]

    def links_of(lk, al, om, idmap, fl):
        nl = np.empty_like(lk)
        na = np.empty_like(al)
        no = np.empty_like(om)
        ends = idmap[lk]
        nl[lam] = np.where(fl[:, None], ends[:, ::-1], ends)
        if wrong == "alpha":
            na[lam] = al
        else:
            na[lam] = np.where(fl[:, None], al[:, ::-1], al)
        no[lam] = np.where(fl, -om, om)
        return nl, na, no

    h.links, h.alpha, h.omega = links_of(g.links, g.alpha, g.omega, sig, flip)
    h.dual_links, h.dual_alpha, h.dual_omega = links_of(g.dual_links, g.dual_alpha, g.dual_omega, phi, dflip)
    h.dual_nns = np.empty_like(g.dual_nns)
    for f in range(F):
        h.dual_nns[phi[f]] = rng.permutation(phi[g.dual_nns[f]])
    nbrs = [None] * V
    rings = [None] * V
    for x in range(V):
        nbrs[sig[x]] = rng.permutation(sig[g.neighbours(x)])
        ring = phi[g.dual_ring(x)]
        rings[sig[x]] = np.roll(ring, int(rng.integers(0, len(ring))))
    h.nn_offsets = np.concatenate([[0], np.cumsum([len(a) for a in nbrs])]).astype(np.int32)
    h.nns = np.concatenate(nbrs).astype(np.int32)
    h.dual_face_offsets = np.concatenate([[0], np.cumsum([len(a) for a in rings])]).astype(np.int32)
    h.dual_faces = np.concatenate(rings).astype(np.int32)
    h.neighbours = lambda x: h.nns[h.nn_offsets[x]:h.nn_offsets[x + 1]]
    h.dual_ring = lambda x: h.dual_faces[h.dual_face_offsets[x]:h.dual_face_offsets[x + 1]]

    def relabel(lattice, Lt):
        if lattice == 0:
            return Relabel(sig, lam, np.where(flip, -1.0, 1.0), Lt)
        return Relabel(phi, lam, np.where(dflip, -1.0, 1.0), Lt)

    return h, relabel
