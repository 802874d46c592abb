// Refined-icosahedron geometry, spin structure and hop tables (host C++).
// See geometry.hpp for the construction and the reference files each part stands in for.
#include "geometry.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <stdexcept>
#include <string>

namespace qed3 {
namespace {

inline V3 add(const V3& a, const V3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(double s, const V3& a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3& a, const V3& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(const V3& a) { return std::sqrt(dot(a, a)); }
inline V3 normalized(const V3& a) { return mul(1.0 / norm(a), a); }

// Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707 (2004) 381:
// axes -> transposed Hilbert index, then bit interleave.
uint64_t hilbert_key(const V3& p) {
  const int b = 20;
  uint32_t X[3];
  const double c[3] = {p.x, p.y, p.z};
  for (int i = 0; i < 3; i++) {
    double u = (c[i] + 1.0) * 0.5;
    u = std::min(std::max(u, 0.0), 1.0);
    X[i] = static_cast<uint32_t>(std::min(u * double(1u << b), double((1u << b) - 1)));
  }
  const uint32_t M = 1u << (b - 1);
  for (uint32_t Q = M; Q > 1; Q >>= 1) {
    const uint32_t P = Q - 1;
    for (int i = 0; i < 3; i++) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        const uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  for (int i = 1; i < 3; i++) X[i] ^= X[i - 1];
  uint32_t t = 0;
  for (uint32_t Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  for (int i = 0; i < 3; i++) X[i] ^= t;
  uint64_t key = 0;
  for (int bit = b - 1; bit >= 0; bit--)
    for (int i = 0; i < 3; i++) key = (key << 1) | ((X[i] >> bit) & 1u);
  return key;
}

// Icosahedron: north pole, upper ring (z=1/sqrt5, phi=72k deg), lower ring
// (z=-1/sqrt5, phi=36+72k deg), south pole.
void icosahedron(std::vector<V3>& v, std::vector<std::array<int, 3>>& f) {
  const double z1 = 1.0 / std::sqrt(5.0), rho = 2.0 / std::sqrt(5.0);
  v.clear();
  v.push_back({0, 0, 1});
  for (int k = 0; k < 5; k++) {
    const double ph = 2.0 * M_PI * k / 5.0;
    v.push_back({rho * std::cos(ph), rho * std::sin(ph), z1});
  }
  for (int k = 0; k < 5; k++) {
    const double ph = 2.0 * M_PI * k / 5.0 + M_PI / 5.0;
    v.push_back({rho * std::cos(ph), rho * std::sin(ph), -z1});
  }
  v.push_back({0, 0, -1});
  f.clear();
  for (int k = 0; k < 5; k++) {
    const int u0 = 1 + k, u1 = 1 + (k + 1) % 5;
    const int l0 = 6 + k, l1 = 6 + (k + 1) % 5;  // l0 sits between u0 and u1
    f.push_back({0, u0, u1});
    f.push_back({u0, l0, u1});
    f.push_back({l0, l1, u1});
    f.push_back({11, l1, l0});
  }
  for (auto& t : f) {  // orient ccw seen from outside
    const V3 a = v[t[0]], b = v[t[1]], c = v[t[2]];
    if (dot(cross(sub(b, a), sub(c, a)), add(add(a, b), c)) < 0) std::swap(t[1], t[2]);
  }
}

V3 circumcenter_dir(const V3& r0, const V3& r1, const V3& r2) {
  const V3 r10 = sub(r1, r0), r20 = sub(r2, r0);
  const V3 t = sub(mul(dot(r10, r10), r20), mul(dot(r20, r20), r10));
  const V3 c = cross(r10, r20);
  const V3 p = add(mul(1.0 / (2.0 * dot(c, c)), cross(t, c)), r0);
  return normalized(p);
}

// Polar angles (theta, phi) with phi in (-pi, pi].
inline void polar(const V3& x, double& th, double& ph) {
  th = std::acos(std::max(-1.0, std::min(1.0, x.z)));
  ph = std::atan2(x.y, x.x);
}

// Angle of the unit tangent at x pointing along the geodesic toward y, measured in
// the orthonormal frame (e_theta, e_phi) at x.  The reference obtains it from the
// geodesic solution (test_sorting_simp.cc:238-252): alpha = acos(dtheta/ds) signed by
// sin(theta) dphi/ds; the result lies in [-pi, pi].
double tangent_angle(const V3& x, const V3& y) {
  double th, ph;
  polar(x, th, ph);
  const V3 t = sub(y, mul(dot(x, y), x));
  const V3 et = {std::cos(th) * std::cos(ph), std::cos(th) * std::sin(ph), -std::sin(th)};
  const V3 ep = {-std::sin(ph), std::cos(ph), 0.0};
  return std::atan2(dot(t, ep), dot(t, et));
}

inline double wrap_pi(double a) {  // into [-pi, pi)
  while (a >= M_PI) a -= 2.0 * M_PI;
  while (a < -M_PI) a += 2.0 * M_PI;
  return a;
}
inline double mod_2pi(double a) {  // into [0, 2pi), floor convention of geodesic.h Mod
  return a - 2.0 * M_PI * std::floor(a / (2.0 * M_PI));
}

// omega_xy = integral of cos(theta) dphi along the geodesic x->y (the quantity the
// reference integrates numerically, test_sorting_simp.cc:113-148), in closed form by
// Gauss-Bonnet: the integral equals dphi - (signed area of the spherical triangle
// N,x,y), where dphi in (-pi,pi) is the azimuth change of the minor arc and the signed
// area follows from Van Oosterom & Strackee, IEEE TBME 30 (1983) 125.
double connection_angle(const V3& x, const V3& y) {
  double thx, phx, thy, phy;
  polar(x, thx, phx);
  polar(y, thy, phy);
  const double dphi = wrap_pi(phy - phx);
  const V3 N{0, 0, 1};
  const double num = dot(N, cross(x, y));
  const double den = 1.0 + x.z + y.z + dot(x, y);
  const double area = 2.0 * std::atan2(num, den);
  return dphi - area;
}

// Branch cut on the meridian phi0 = pi/48 (test_sorting_simp.cc:163-195): links whose
// endpoints straddle it get omega -= 2pi in their stored orientation.
bool crosses_cut(const V3& x, const V3& y) {
  const double phi0 = M_PI / 48.0;
  double th, p1, p2;
  polar(x, th, p1);
  polar(y, th, p2);
  const double d1 = wrap_pi(p1 - phi0), d2 = wrap_pi(p2 - phi0), d = wrap_pi(p1 - p2);
  return std::abs(d1) + std::abs(d2) - std::abs(d) < 1.0e-6;
}

void spin_structure(const std::vector<V3>& pts, const std::vector<std::array<int, 2>>& links,
                    std::vector<double>& afwd, std::vector<double>& abwd,
                    std::vector<double>& om) {
  const size_t E = links.size();
  afwd.resize(E);
  abwd.resize(E);
  om.resize(E);
  for (size_t k = 0; k < E; k++) {
    const V3& x = pts[links[k][0]];
    const V3& y = pts[links[k][1]];
    afwd[k] = tangent_angle(x, y);
    abwd[k] = mod_2pi(tangent_angle(y, x));
    double w = connection_angle(x, y);
    if (crosses_cut(x, y)) w -= 2.0 * M_PI;
    om[k] = w;
  }
}

}  // namespace

double geodesic_length(const V3& a, const V3& b) {
  const double c = dot(a, b) / (norm(a) * norm(b));
  return std::acos(std::max(-1.0, std::min(1.0, c)));
}

double spherical_area(const V3& p1, const V3& p2, const V3& p3) {
  const double a = geodesic_length(p2, p3), b = geodesic_length(p3, p1),
               c = geodesic_length(p1, p2);
  const double s = 0.5 * (a + b + c);
  const double t = std::tan(0.5 * s) * std::tan(0.5 * (s - a)) * std::tan(0.5 * (s - b)) *
                   std::tan(0.5 * (s - c));
  return 4.0 * std::atan(std::sqrt(std::max(t, 0.0)));
}

Geometry make_geometry(int n) {
  if (n < 1 || n > 2048) throw std::invalid_argument("n_refine out of range");
  std::vector<V3> iv;
  std::vector<std::array<int, 3>> ifc;
  icosahedron(iv, ifc);

  // Points keyed by their integer barycentric weights over icosahedron vertices, so
  // that a point shared by several base faces is generated bit-identically once.
  using Key = std::array<int, 6>;  // (v0,w0,v1,w1,v2,w2), sorted by vertex, unused = -1
  std::map<Key, int> id_of;
  std::vector<V3> pts;
  auto point = [&](const std::array<int, 3>& f, int i, int j) {
    std::array<std::pair<int, int>, 3> vw = {std::make_pair(f[0], n - i - j),
                                             std::make_pair(f[1], i), std::make_pair(f[2], j)};
    std::sort(vw.begin(), vw.end());
    Key key;
    key.fill(-1);
    int m = 0;
    for (auto& e : vw)
      if (e.second > 0) {
        key[2 * m] = e.first;
        key[2 * m + 1] = e.second;
        m++;
      }
    auto it = id_of.find(key);
    if (it != id_of.end()) return it->second;
    V3 p{0, 0, 0};
    for (int q = 0; q < m; q++) p = add(p, mul(double(key[2 * q + 1]) / n, iv[key[2 * q]]));
    const int id = int(pts.size());
    pts.push_back(normalized(p));
    id_of.emplace(key, id);
    return id;
  };
  std::vector<std::array<int, 3>> tris;
  for (const auto& f : ifc) {
    for (int i = 0; i < n; i++)
      for (int j = 0; i + j < n; j++) {
        const int a = point(f, i, j), b = point(f, i + 1, j), c = point(f, i, j + 1);
        tris.push_back({a, b, c});
        if (i + j <= n - 2) {
          const int d = point(f, i + 1, j + 1);
          tris.push_back({b, d, c});
        }
      }
  }
  const int V = int(pts.size());
  if (V != 10 * n * n + 2) throw std::logic_error("site count mismatch");

  // Hilbert renumbering of sites.
  std::vector<int> order(V);
  std::iota(order.begin(), order.end(), 0);
  std::vector<uint64_t> hk(V);
  for (int i = 0; i < V; i++) hk[i] = hilbert_key(pts[i]);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return hk[a] < hk[b]; });
  std::vector<int> newid(V);
  for (int i = 0; i < V; i++) newid[order[i]] = i;

  Geometry g;
  g.n = n;
  g.sites.resize(V);
  for (int i = 0; i < V; i++) g.sites[newid[i]] = pts[i];

  // Faces, ccw, ordered by the Hilbert key of the centroid.
  for (auto& t : tris) {
    for (auto& v : t) v = newid[v];
    // rotate so the smallest id comes first (orientation preserved)
    while (t[0] > t[1] || t[0] > t[2]) std::rotate(t.begin(), t.begin() + 1, t.end());
  }
  {
    const size_t F = tris.size();
    std::vector<uint64_t> fk(F);
    for (size_t k = 0; k < F; k++)
      fk[k] = hilbert_key(normalized(add(add(g.sites[tris[k][0]], g.sites[tris[k][1]]),
                                         g.sites[tris[k][2]])));
    std::vector<int> fo(F);
    std::iota(fo.begin(), fo.end(), 0);
    std::stable_sort(fo.begin(), fo.end(), [&](int a, int b) { return fk[a] < fk[b]; });
    g.faces.resize(F);
    for (size_t k = 0; k < F; k++) g.faces[k] = tris[fo[k]];
  }
  const int F = int(g.faces.size());
  if (F != 20 * n * n) throw std::logic_error("face count mismatch");

  // Links (i<j) and the two faces adjacent to each.
  std::map<std::pair<int, int>, std::vector<int>> link_faces;
  for (int f = 0; f < F; f++)
    for (int e = 0; e < 3; e++) {
      int a = g.faces[f][e], b = g.faces[f][(e + 1) % 3];
      if (a > b) std::swap(a, b);
      link_faces[{a, b}].push_back(f);
    }
  const int E = int(link_faces.size());
  if (E != 30 * n * n) throw std::logic_error("link count mismatch");
  g.nns.assign(V, {});
  for (auto& kv : link_faces) {
    if (kv.second.size() != 2) throw std::logic_error("non-manifold link");
    g.links.push_back({kv.first.first, kv.first.second});
    const int fa = std::min(kv.second[0], kv.second[1]);
    const int fb = std::max(kv.second[0], kv.second[1]);
    g.dual_links.push_back({fa, fb});
    g.nns[kv.first.first].push_back(kv.first.second);
    g.nns[kv.first.second].push_back(kv.first.first);
  }
  for (auto& v : g.nns) std::sort(v.begin(), v.end());

  // Dual sites (normalised circumcentres, get_geometry.cc:94-104) and dual neighbours.
  g.dual_sites.resize(F);
  for (int f = 0; f < F; f++)
    g.dual_sites[f] =
        circumcenter_dir(g.sites[g.faces[f][0]], g.sites[g.faces[f][1]], g.sites[g.faces[f][2]]);
  {
    std::vector<std::vector<int>> dn(F);
    for (auto& dl : g.dual_links) {
      dn[dl[0]].push_back(dl[1]);
      dn[dl[1]].push_back(dl[0]);
    }
    g.dual_nns.resize(F);
    for (int f = 0; f < F; f++) {
      if (dn[f].size() != 3) throw std::logic_error("dual site not trivalent");
      std::sort(dn[f].begin(), dn[f].end());
      g.dual_nns[f] = {dn[f][0], dn[f][1], dn[f][2]};
    }
  }

  // Dual faces: the ring of faces around each site, ccw seen from outside,
  // starting at the smallest face id (get_geometry.cc:280-337 orders them likewise).
  {
    std::vector<std::vector<int>> around(V);
    for (int f = 0; f < F; f++)
      for (int e = 0; e < 3; e++) around[g.faces[f][e]].push_back(f);
    g.dual_faces.resize(V);
    for (int x = 0; x < V; x++) {
      auto& fs = around[x];
      const V3 p = g.sites[x];
      // sort by angle around p in the tangent plane
      V3 e1 = sub(g.dual_sites[fs[0]], mul(dot(g.dual_sites[fs[0]], p), p));
      e1 = normalized(e1);
      const V3 e2 = cross(p, e1);
      std::vector<std::pair<double, int>> ang;
      for (int f : fs) {
        const V3 d = g.dual_sites[f];
        ang.push_back({std::atan2(dot(d, e2), dot(d, e1)), f});
      }
      std::sort(ang.begin(), ang.end());
      std::vector<int> ring;
      for (auto& a : ang) ring.push_back(a.second);
      const auto mn = std::min_element(ring.begin(), ring.end());
      std::rotate(ring.begin(), mn, ring.end());
      g.dual_faces[x] = ring;
    }
  }

  // Spin structure. Simplicial: frame rotated by R_y(t), t = 1/n.
  {
    const double t = 1.0 / n;
    std::vector<V3> rot(V);
    for (int i = 0; i < V; i++) {
      const V3& s = g.sites[i];
      rot[i] = {std::cos(t) * s.x + std::sin(t) * s.z, s.y, -std::sin(t) * s.x + std::cos(t) * s.z};
    }
    spin_structure(rot, g.links, g.alpha_fwd, g.alpha_bwd, g.omega);
    spin_structure(g.dual_sites, g.dual_links, g.dual_alpha_fwd, g.dual_alpha_bwd, g.dual_omega);
  }
  return g;
}

HopTables make_hop_tables(const Geometry& g, int lattice, int Lt, double r, double M5,
                          double at) {
  if (lattice != 0 && lattice != 1) throw std::invalid_argument("lattice must be 0 or 1");
  if (Lt < 1) throw std::invalid_argument("Lt must be >= 1");
  HopTables h;
  h.lattice = lattice;
  h.r = r;
  h.M5 = M5;
  h.at = at;
  h.Lt = Lt;
  const bool dual = lattice == 1;
  const std::vector<V3>& S = dual ? g.dual_sites : g.sites;
  const std::vector<V3>& D = dual ? g.sites : g.dual_sites;
  h.links = dual ? g.dual_links : g.links;
  const auto& dlinks = dual ? g.links : g.dual_links;  // the crossing link of each link
  const auto& afwd = dual ? g.dual_alpha_fwd : g.alpha_fwd;
  const auto& abwd = dual ? g.dual_alpha_bwd : g.alpha_bwd;
  const auto& om = dual ? g.dual_omega : g.omega;
  h.V = int(S.size());
  h.E = int(h.links.size());
  h.max_nn = dual ? 3 : 6;

  // ell, link volume (s2n_simp.h:301-364, s2n_dual.h:109-160)
  h.ell.resize(h.E);
  h.link_volume.resize(h.E);
  double sum_ell = 0;
  for (int l = 0; l < h.E; l++) {
    const V3& x = S[h.links[l][0]];
    const V3& y = S[h.links[l][1]];
    h.ell[l] = geodesic_length(x, y);
    h.link_volume[l] = spherical_area(x, y, D[dlinks[l][0]]) + spherical_area(x, y, D[dlinks[l][1]]);
    sum_ell += h.ell[l];
  }
  h.mean_ell = sum_ell / h.E;
  h.kappa.resize(h.E);
  for (int l = 0; l < h.E; l++) h.kappa[l] = 2.0 * h.link_volume[l] / h.ell[l] / h.mean_ell;

  // dual areas: area of the dual cell around each site.
  h.dual_area.assign(h.V, 0.0);
  if (!dual) {
    for (int x = 0; x < h.V; x++) {
      const auto& ring = g.dual_faces[x];
      double a = 0;
      for (size_t i = 0; i < ring.size(); i++)
        a += spherical_area(g.dual_sites[ring[i]], g.dual_sites[ring[(i + 1) % ring.size()]],
                            g.sites[x]);
      h.dual_area[x] = a;
    }
    for (const auto& f : g.faces) {
      h.face_rings.push_back({f[0], f[1], f[2]});
      h.face_area.push_back(spherical_area(g.sites[f[0]], g.sites[f[1]], g.sites[f[2]]));
    }
  } else {
    for (int f = 0; f < h.V; f++)
      h.dual_area[f] =
          spherical_area(g.sites[g.faces[f][0]], g.sites[g.faces[f][1]], g.sites[g.faces[f][2]]);
    for (int x = 0; x < int(g.sites.size()); x++) {
      const auto& ring = g.dual_faces[x];
      h.face_rings.push_back(ring);
      double a = 0;
      for (size_t i = 0; i < ring.size(); i++)
        a += spherical_area(g.dual_sites[ring[i]], g.dual_sites[ring[(i + 1) % ring.size()]],
                            g.sites[x]);
      h.face_area.push_back(a);
    }
  }
  h.F = int(h.face_rings.size());

  // kappa_t (dirac_ext.h:307-320)
  h.kappa_t.resize(h.V);
  for (int x = 0; x < h.V; x++) h.kappa_t[x] = (Lt != 1) ? h.dual_area[x] / h.mean_ell / at : 0.0;

  // Slots.
  std::vector<std::vector<std::pair<int, int>>> adj(h.V);  // (neighbour, signed link)
  for (int l = 0; l < h.E; l++) {
    adj[h.links[l][0]].push_back({h.links[l][1], l});
    adj[h.links[l][1]].push_back({h.links[l][0], ~l});
  }
  const int V = h.V, M = h.max_nn;
  h.nn_count.resize(V);
  h.nbr.assign(size_t(M) * V, 0);
  h.lnk.assign(size_t(M) * V, 0);
  h.hopD.assign(size_t(M) * V * 4, 0.0);
  h.hopDdag.assign(size_t(M) * V * 4, 0.0);
  h.slot_alpha.assign(size_t(M) * V, 0.0);
  h.slot_alpha_nb.assign(size_t(M) * V, 0.0);
  h.diag.resize(V);
  for (int x = 0; x < V; x++) {
    auto& a = adj[x];
    std::sort(a.begin(), a.end());
    if (int(a.size()) > M) throw std::logic_error("site valence exceeds slot count");
    h.nn_count[x] = int(a.size());
    double c = 0;
    for (int j = 0; j < M; j++) {
      const size_t sidx = size_t(j) * V + x;
      if (j >= int(a.size())) {  // zero-weight pad slot
        h.nbr[sidx] = x;
        h.lnk[sidx] = 0;
        continue;
      }
      const int y = a[j].first, sl = a[j].second;
      const bool fwd = sl >= 0;
      const int l = fwd ? sl : ~sl;
      h.nbr[sidx] = y;
      h.lnk[sidx] = sl;
      const double kap = h.kappa[l];
      const double w = fwd ? om[l] : -om[l];            // omega_xy
      const double axy = fwd ? afwd[l] : abwd[l];       // alpha_xy
      const double ayx = fwd ? abwd[l] : afwd[l];       // alpha_yx
      h.slot_alpha[sidx] = axy;
      h.slot_alpha_nb[sidx] = ayx;
      const double C = 0.5 * kap * std::cos(0.5 * w), Sg = 0.5 * kap * std::sin(0.5 * w);
      double* hd = &h.hopD[sidx * 4];
      hd[0] = C;
      hd[1] = Sg;
      hd[2] = C * std::cos(axy) + Sg * std::sin(axy);
      hd[3] = C * std::sin(axy) - Sg * std::cos(axy);
      // H_yx has (C, -S, a_yx, b_yx); its adjoint is again of the H form with
      // (C, S, C cos(alpha_yx) - S sin(alpha_yx), C sin(alpha_yx) + S cos(alpha_yx)).
      double* hh = &h.hopDdag[sidx * 4];
      hh[0] = C;
      hh[1] = Sg;
      hh[2] = C * std::cos(ayx) - Sg * std::sin(ayx);
      hh[3] = C * std::sin(ayx) + Sg * std::cos(ayx);
      c += 0.5 * r * kap;
    }
    c += r * h.kappa_t[x];
    c += M5;
    h.diag[x] = c;
  }
  return h;
}

}  // namespace qed3
