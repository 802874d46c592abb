// Refined-icosahedron geometry of S^2 (simplicial lattice and its trivalent dual),
// spin-connection tables, and the per-site hop tables of the Wilson-Dirac operator.
//
// This replaces the reference's offline table generator (integrator/get_geometry.cc,
// integrator/test_sorting_simp.cc, integrator/test_sorting.cc, which need the external
// newQFE package and GSL) and the table loaders/derived quantities of
// src/both_3d/s2n_simp.h, s2n_dual.h, dirac_simp.h, dirac_dual.h, dirac_ext.h.
//
// Construction (ours): the icosahedron with vertex 0 at the north pole and the upper
// ring at z = 1/sqrt(5), phi = 0, 72, ... deg (the orientation of the reference's only
// geometry dump, obsolete/test_geometry.log); each face is subdivided by the integer
// barycentric grid (i,j,k)/n and projected radially, giving 10n^2+2 sites, 30n^2 links
// and 20n^2 faces.  Sites are numbered along a 3D Hilbert curve of their coordinates
// so that stencil neighbours sit close in memory; links are (i<j) pairs sorted
// lexicographically; faces are counter-clockwise seen from outside and sorted by the
// Hilbert key of their centroid.
#pragma once
#include <array>
#include <cstdint>
#include <vector>

namespace qed3 {

struct V3 {
  double x = 0, y = 0, z = 0;
  V3() = default;
  V3(double a, double b, double c) : x(a), y(b), z(c) {}
};

// Raw tables in the roles of the reference's .dat files
// (pts, pts_dual, links, dual_links, face, face_dual, nns, nns_dual, alpha, omega).
struct Geometry {
  int n = 0;
  std::vector<V3> sites;                        // V = 10n^2+2 unit vectors
  std::vector<std::array<int, 3>> faces;        // F = 20n^2, ccw from outside
  std::vector<std::array<int, 2>> links;        // E = 30n^2, (i<j), lexicographic
  std::vector<std::vector<int>> nns;            // per site, ascending
  std::vector<V3> dual_sites;                   // F normalised circumcentres (id = face id)
  std::vector<std::array<int, 2>> dual_links;   // E: (min,max) faces sharing links[k]
  std::vector<std::array<int, 3>> dual_nns;     // F: neighbouring faces, ascending
  std::vector<std::vector<int>> dual_faces;     // V: ring of faces around each site, ccw
  // Spin structure of the simplicial lattice (computed in the frame rotated by
  // R_y(1/n), as the reference generator does, test_sorting_simp.cc:69-86):
  std::vector<double> alpha_fwd;   // per link k=(i,j): alpha_ij (tangent angle at i toward j)
  std::vector<double> alpha_bwd;   // per link: alpha_ji in [0,2pi)
  std::vector<double> omega;       // per link: omega_ij (alpha_ij = alpha_ji + pi + omega_ij mod 2pi)
  // Spin structure of the dual lattice (unrotated frame, test_sorting.cc):
  std::vector<double> dual_alpha_fwd, dual_alpha_bwd, dual_omega;  // per dual link k
};

Geometry make_geometry(int n);

// Per-lattice operator tables (time independent), lattice = 0 simplicial, 1 dual.
// Slot j of site x is its j-th neighbour (ascending id order); sites with fewer than
// max_nn neighbours are padded with a zero-weight self slot.
struct HopTables {
  int lattice = 0;
  int V = 0, E = 0, F = 0, max_nn = 0;
  double mean_ell = 0, r = 1, M5 = 0, at = 1;
  int Lt = 1;
  std::vector<std::array<int, 2>> links;   // E links of this lattice
  std::vector<int> nn_count;               // V
  std::vector<int> nbr;                    // [max_nn][V]
  std::vector<int> lnk;                    // [max_nn][V]  link id, or ~id when traversed y->x
  // 2x2 hop matrix of slot x->y:  H = 1/2 kappa (-r + gamma_xy) Omega_xy
  //   = [[-r(C - iS), a - ib], [a + ib, -r(C + iS)]]
  // with C = kappa/2 cos(omega_xy/2), S = kappa/2 sin(omega_xy/2),
  //      a = C cos(alpha_xy) + S sin(alpha_xy), b = C sin(alpha_xy) - S cos(alpha_xy).
  // The D^dagger table uses H_yx^dagger, which has the same form with (C,S,a_yx,b_yx).
  std::vector<double> hopD;                // [max_nn][V][4] (C,S,a,b)
  std::vector<double> hopDdag;             // [max_nn][V][4]
  std::vector<double> slot_alpha;          // [max_nn][V] alpha_xy of slot x->y
  std::vector<double> slot_alpha_nb;       // [max_nn][V] alpha_yx (the angle at the neighbour)
  std::vector<double> diag;                // c_x = sum_y r kappa_xy/2 + r kappa_t(x) + M5
  std::vector<double> kappa_t;             // V
  std::vector<double> kappa;               // E
  std::vector<double> ell, link_volume, dual_area;
  // faces of this lattice (simplicial triangles or dual hexagons/pentagons) as
  // oriented site rings, with the per-face area (for the gauge action)
  std::vector<std::vector<int>> face_rings;
  std::vector<double> face_area;
};

HopTables make_hop_tables(const Geometry& g, int lattice, int Lt, double r, double M5,
                          double at);

// Spherical helpers (unit vectors).
double geodesic_length(const V3& a, const V3& b);
double spherical_area(const V3& a, const V3& b, const V3& c);  // L'Huilier

}  // namespace qed3
