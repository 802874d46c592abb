// C ABI of libqed3 (declared in include/qed3.h).  Every entry point converts C++
// exceptions into status codes; the text of the last error is kept per thread.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "../../include/qed3.h"
#include "cg.hpp"
#include "dirac.hpp"
#include "geometry.hpp"
#include "batch.hpp"
#include "fcg.hpp"
#include "hmc.hpp"
#include "horner.hpp"
#include "mpnn.hpp"
#include "msolve.hpp"
#include "rng.hpp"

struct qed3_geom {
  qed3::Geometry g;
};

struct qed3_ctx {
  qed3_params p;
  std::unique_ptr<qed3::Geometry> g;
  std::unique_ptr<qed3::Dirac> D;
};

struct qed3_hmc {
  qed3_ctx* ctx;
  std::unique_ptr<qed3::HMC> hmc;
};

struct qed3_overlap {
  qed3_ctx* ctx;
  std::unique_ptr<qed3::Overlap> ov;
};

struct qed3_rng {
  std::unique_ptr<qed3::ParallelRng> rng;
};

struct qed3_net {
  qed3_ctx* ctx;
  std::unique_ptr<qed3::Net> net;
};

struct qed3_precond {
  qed3_ctx* ctx;
  std::unique_ptr<qed3::Horner> K;
};

namespace {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const qed3::HipError& e) {
    return fail(QED3_ERR_HIP, e.what());
  } catch (const std::invalid_argument& e) {
    return fail(QED3_ERR_ARG, e.what());
  } catch (const std::bad_alloc& e) {
    return fail(QED3_ERR_HIP, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(QED3_ERR_ARG, e.what());
  }
}
}  // namespace

extern "C" {

const char* qed3_last_error(void) { return g_err.c_str(); }

int qed3_malloc(void** d_ptr, size_t bytes) {
  if (!d_ptr) return fail(QED3_ERR_ARG, "null pointer");
  return guarded([&]() -> int {
    QED3_HIP(hipMalloc(d_ptr, bytes));
    return QED3_OK;
  });
}

int qed3_free(void* d_ptr) {
  return guarded([&]() -> int {
    QED3_HIP(hipFree(d_ptr));
    return QED3_OK;
  });
}

int qed3_memcpy(void* dst, const void* src, size_t bytes, int kind) {
  static const hipMemcpyKind kinds[3] = {hipMemcpyHostToDevice, hipMemcpyDeviceToHost,
                                         hipMemcpyDeviceToDevice};
  if (kind < 0 || kind > 2) return fail(QED3_ERR_ARG, "bad memcpy kind");
  return guarded([&]() -> int {
    QED3_HIP(hipMemcpy(dst, src, bytes, kinds[kind]));
    return QED3_OK;
  });
}

int qed3_device_synchronize(void) {
  return guarded([&]() -> int {
    QED3_HIP(hipDeviceSynchronize());
    return QED3_OK;
  });
}

int qed3_geom_create(int n_refine, qed3_geom** out) {
  if (!out) return fail(QED3_ERR_ARG, "null out");
  return guarded([&]() -> int {
    auto* g = new qed3_geom;
    try {
      g->g = qed3::make_geometry(n_refine);
    } catch (...) {
      delete g;
      throw;
    }
    *out = g;
    return QED3_OK;
  });
}

void qed3_geom_destroy(qed3_geom* g) { delete g; }

int qed3_geom_counts(const qed3_geom* g, int* V, int* E, int* F) {
  if (!g) return fail(QED3_ERR_ARG, "null geometry");
  if (V) *V = int(g->g.sites.size());
  if (E) *E = int(g->g.links.size());
  if (F) *F = int(g->g.faces.size());
  return QED3_OK;
}

int qed3_geom_export(const qed3_geom* gg, double* sites, int* faces, int* links, int* nn_offsets,
                     int* nns, double* dual_sites, int* dual_links, int* dual_nns,
                     int* dual_face_offsets, int* dual_faces, double* alpha, double* omega,
                     double* dual_alpha, double* dual_omega) {
  if (!gg) return fail(QED3_ERR_ARG, "null geometry");
  const qed3::Geometry& g = gg->g;
  const size_t V = g.sites.size(), E = g.links.size(), F = g.faces.size();
  if (sites)
    for (size_t i = 0; i < V; i++) {
      sites[3 * i] = g.sites[i].x;
      sites[3 * i + 1] = g.sites[i].y;
      sites[3 * i + 2] = g.sites[i].z;
    }
  if (faces)
    for (size_t f = 0; f < F; f++)
      for (int k = 0; k < 3; k++) faces[3 * f + k] = g.faces[f][k];
  if (links)
    for (size_t l = 0; l < E; l++) {
      links[2 * l] = g.links[l][0];
      links[2 * l + 1] = g.links[l][1];
    }
  if (nn_offsets || nns) {
    int o = 0;
    for (size_t x = 0; x < V; x++) {
      if (nn_offsets) nn_offsets[x] = o;
      for (int y : g.nns[x]) {
        if (nns) nns[o] = y;
        o++;
      }
    }
    if (nn_offsets) nn_offsets[V] = o;
  }
  if (dual_sites)
    for (size_t f = 0; f < F; f++) {
      dual_sites[3 * f] = g.dual_sites[f].x;
      dual_sites[3 * f + 1] = g.dual_sites[f].y;
      dual_sites[3 * f + 2] = g.dual_sites[f].z;
    }
  if (dual_links)
    for (size_t l = 0; l < E; l++) {
      dual_links[2 * l] = g.dual_links[l][0];
      dual_links[2 * l + 1] = g.dual_links[l][1];
    }
  if (dual_nns)
    for (size_t f = 0; f < F; f++)
      for (int k = 0; k < 3; k++) dual_nns[3 * f + k] = g.dual_nns[f][k];
  if (dual_face_offsets || dual_faces) {
    int o = 0;
    for (size_t x = 0; x < V; x++) {
      if (dual_face_offsets) dual_face_offsets[x] = o;
      for (int f : g.dual_faces[x]) {
        if (dual_faces) dual_faces[o] = f;
        o++;
      }
    }
    if (dual_face_offsets) dual_face_offsets[V] = o;
  }
  for (size_t l = 0; l < E; l++) {
    if (alpha) {
      alpha[2 * l] = g.alpha_fwd[l];
      alpha[2 * l + 1] = g.alpha_bwd[l];
    }
    if (omega) omega[l] = g.omega[l];
    if (dual_alpha) {
      dual_alpha[2 * l] = g.dual_alpha_fwd[l];
      dual_alpha[2 * l + 1] = g.dual_alpha_bwd[l];
    }
    if (dual_omega) dual_omega[l] = g.dual_omega[l];
  }
  return QED3_OK;
}

int qed3_geom_import(int n_refine, int V, int E, int F, const double* sites, const int* faces,
                     const int* links, const int* nn_offsets, const int* nns,
                     const double* dual_sites, const int* dual_links, const int* dual_nns,
                     const int* dual_face_offsets, const int* dual_faces, const double* alpha,
                     const double* omega, const double* dual_alpha, const double* dual_omega,
                     qed3_geom** out) {
  if (!out || !sites || !faces || !links || !nn_offsets || !nns || !dual_sites || !dual_links ||
      !dual_nns || !dual_face_offsets || !dual_faces || !alpha || !omega || !dual_alpha || !dual_omega)
    return fail(QED3_ERR_ARG, "null table");
  if (V < 4 || E < 6 || F < 4 || V - E + F != 2) return fail(QED3_ERR_ARG, "V - E + F must be 2");
  return guarded([&]() -> int {
    auto g = std::make_unique<qed3_geom>();
    qed3::Geometry& G = g->g;
    G.n = n_refine;
    for (int i = 0; i < V; i++) G.sites.push_back({sites[3 * i], sites[3 * i + 1], sites[3 * i + 2]});
    for (int f = 0; f < F; f++) {
      G.faces.push_back({faces[3 * f], faces[3 * f + 1], faces[3 * f + 2]});
      G.dual_sites.push_back({dual_sites[3 * f], dual_sites[3 * f + 1], dual_sites[3 * f + 2]});
      G.dual_nns.push_back({dual_nns[3 * f], dual_nns[3 * f + 1], dual_nns[3 * f + 2]});
    }
    auto in = [](int v, int hi) {
      if (v < 0 || v >= hi) throw std::invalid_argument("table index out of range");
      return v;
    };
    for (int l = 0; l < E; l++) {
      G.links.push_back({in(links[2 * l], V), in(links[2 * l + 1], V)});
      G.dual_links.push_back({in(dual_links[2 * l], F), in(dual_links[2 * l + 1], F)});
      G.alpha_fwd.push_back(alpha[2 * l]);
      G.alpha_bwd.push_back(alpha[2 * l + 1]);
      G.omega.push_back(omega[l]);
      G.dual_alpha_fwd.push_back(dual_alpha[2 * l]);
      G.dual_alpha_bwd.push_back(dual_alpha[2 * l + 1]);
      G.dual_omega.push_back(dual_omega[l]);
    }
    for (int x = 0; x < V; x++)
      if (nn_offsets[x + 1] < nn_offsets[x] || dual_face_offsets[x + 1] < dual_face_offsets[x])
        throw std::invalid_argument("offsets must not decrease");
    if (nn_offsets[0] != 0 || nn_offsets[V] != 2 * E)
      throw std::invalid_argument("nn_offsets must run from 0 to 2E");
    if (dual_face_offsets[0] != 0 || dual_face_offsets[V] != 3 * F)
      throw std::invalid_argument("dual_face_offsets must run from 0 to 3F");
    G.nns.resize(V);
    G.dual_faces.resize(V);
    for (int x = 0; x < V; x++) {
      for (int k = nn_offsets[x]; k < nn_offsets[x + 1]; k++) G.nns[x].push_back(in(nns[k], V));
      for (int k = dual_face_offsets[x]; k < dual_face_offsets[x + 1]; k++)
        G.dual_faces[x].push_back(in(dual_faces[k], F));
    }
    for (auto& f : G.faces)
      for (int v : f) in(v, V);
    for (auto& f : G.dual_nns)
      for (int v : f) in(v, F);
    *out = g.release();
    return QED3_OK;
  });
}

int qed3_random_gauge(int Lt, int E, int V, int seed, double kappa, double* spatial,
                      double* temporal) {
  if (Lt < 1 || E < 1 || V < 1 || !spatial || !temporal) return fail(QED3_ERR_ARG, "bad argument");
  return guarded([&]() -> int {
    qed3::draw_gauge(Lt, E, V, seed, kappa, spatial, temporal);
    return QED3_OK;
  });
}

int qed3_random_source(int Lt, int E, int V, int seed, double* xi) {
  if (Lt < 1 || E < 1 || V < 1 || !xi) return fail(QED3_ERR_ARG, "bad argument");
  return guarded([&]() -> int {
    qed3::draw_source(Lt, E, V, seed, xi);
    return QED3_OK;
  });
}

int qed3_ctx_create(const qed3_params* p, qed3_ctx** out) {
  if (!p || !out) return fail(QED3_ERR_ARG, "null argument");
  if (p->lattice != 0 && p->lattice != 1) return fail(QED3_ERR_ARG, "lattice must be 0 or 1");
  if (p->Lt < 1 || p->n_refine < 1) return fail(QED3_ERR_ARG, "Lt and n_refine must be >= 1");
  if (p->bc_time != 0 && p->bc_time != 1) return fail(QED3_ERR_ARG, "bc_time must be 0 or 1");
  if (!(p->at > 0)) return fail(QED3_ERR_ARG, "at must be positive");
  return guarded([&]() -> int {
    auto c = std::make_unique<qed3_ctx>();
    c->p = *p;
    c->g = std::make_unique<qed3::Geometry>(qed3::make_geometry(p->n_refine));
    c->D = std::make_unique<qed3::Dirac>(*c->g, p->lattice, p->Lt, p->bc_time, p->r, p->M5, p->at);
    *out = c.release();
    return QED3_OK;
  });
}

int qed3_ctx_create_geom(const qed3_params* p, const qed3_geom* g, qed3_ctx** out) {
  if (!p || !g || !out) return fail(QED3_ERR_ARG, "null argument");
  if (p->lattice != 0 && p->lattice != 1) return fail(QED3_ERR_ARG, "lattice must be 0 or 1");
  if (p->Lt < 1) return fail(QED3_ERR_ARG, "Lt must be >= 1");
  if (p->bc_time != 0 && p->bc_time != 1) return fail(QED3_ERR_ARG, "bc_time must be 0 or 1");
  if (!(p->at > 0)) return fail(QED3_ERR_ARG, "at must be positive");
  return guarded([&]() -> int {
    auto c = std::make_unique<qed3_ctx>();
    c->p = *p;
    c->g = std::make_unique<qed3::Geometry>(g->g);
    c->D = std::make_unique<qed3::Dirac>(*c->g, p->lattice, p->Lt, p->bc_time, p->r, p->M5, p->at);
    *out = c.release();
    return QED3_OK;
  });
}

void qed3_ctx_destroy(qed3_ctx* ctx) { delete ctx; }

int qed3_ctx_sizes(const qed3_ctx* ctx, int* V, int* E, int* F, double* mean_ell) {
  if (!ctx) return fail(QED3_ERR_ARG, "null context");
  const auto& h = ctx->D->tables();
  if (V) *V = h.V;
  if (E) *E = h.E;
  if (F) *F = h.F;
  if (mean_ell) *mean_ell = h.mean_ell;
  return QED3_OK;
}

int qed3_ctx_geometry(const qed3_ctx* ctx, double* kappa, double* kappa_t, double* link_volume,
                      double* ell, double* dual_area) {
  if (!ctx) return fail(QED3_ERR_ARG, "null context");
  const auto& h = ctx->D->tables();
  if (kappa) std::memcpy(kappa, h.kappa.data(), h.E * sizeof(double));
  if (kappa_t) std::memcpy(kappa_t, h.kappa_t.data(), h.V * sizeof(double));
  if (link_volume) std::memcpy(link_volume, h.link_volume.data(), h.E * sizeof(double));
  if (ell) std::memcpy(ell, h.ell.data(), h.E * sizeof(double));
  if (dual_area) std::memcpy(dual_area, h.dual_area.data(), h.V * sizeof(double));
  return QED3_OK;
}

int qed3_set_gauge(qed3_ctx* ctx, const double* spatial, const double* temporal, int where) {
  if (!ctx || !spatial || !temporal) return fail(QED3_ERR_ARG, "null argument");
  if (where != QED3_HOST && where != QED3_DEVICE) return fail(QED3_ERR_ARG, "bad where");
  return guarded([&]() -> int {
    ctx->D->set_gauge(spatial, temporal, where == QED3_DEVICE);
    return QED3_OK;
  });
}

namespace {
// qed3_apply and qed3_apply_async; the two-pass products keep their intermediate field in
// the context's scratch slot kApplySlot, so the asynchronous form needs no allocation that
// would outlive the call.
constexpr int kApplySlot = 11;
int apply_entry(qed3_ctx* ctx, int op, int prec, void* d_out, const void* d_in, bool sync) {
  if (!ctx || !d_out || !d_in) return fail(QED3_ERR_ARG, "null argument");
  if (op < 0 || op > 3) return fail(QED3_ERR_ARG, "bad op");
  if (prec != QED3_PREC_F64 && prec != QED3_PREC_F32 && prec != QED3_PREC_F16)
    return fail(QED3_ERR_ARG, "bad precision");
  if (d_out == d_in) return fail(QED3_ERR_ARG, "in-place apply is not supported");
  if (!ctx->D->has_gauge()) return fail(QED3_ERR_STATE, "no gauge field set");
  return guarded([&]() -> int {
    auto& D = *ctx->D;
    auto go = [&](auto tag) {
      using T = decltype(tag);
      auto* o = static_cast<qed3::Spin<T>*>(d_out);
      auto* i = static_cast<const qed3::Spin<T>*>(d_in);
      if (op == QED3_OP_D || op == QED3_OP_DDAG) {
        D.apply<T>(op == QED3_OP_DDAG, o, i, sync);
        return;
      }
      auto* tmp = static_cast<qed3::Spin<T>*>(D.scratch(kApplySlot, D.points() * sizeof(qed3::Spin<T>)));
      if (op == QED3_OP_DDAGD)
        D.apply_dagd<T>(o, i, tmp, sync);
      else
        D.apply_ddag<T>(o, i, tmp, sync);
    };
    if (prec == QED3_PREC_F16) {
      auto* o = static_cast<qed3::SpinH*>(d_out);
      auto* i = static_cast<const qed3::SpinH*>(d_in);
      if (op == QED3_OP_D || op == QED3_OP_DDAG) {
        D.apply_half(op == QED3_OP_DDAG, o, i, sync);
        return QED3_OK;
      }
      auto* tmp = static_cast<qed3::SpinH*>(D.scratch(kApplySlot, D.points() * sizeof(qed3::SpinH)));
      D.apply_half(op == QED3_OP_DDAGD ? false : true, tmp, i, false);
      D.apply_half(op == QED3_OP_DDAGD ? true : false, o, tmp, sync);
      return QED3_OK;
    }
    if (prec == QED3_PREC_F64)
      go(double{});
    else
      go(float{});
    return QED3_OK;
  });
}
}  // namespace

int qed3_apply(qed3_ctx* ctx, int op, int prec, void* d_out, const void* d_in) {
  return apply_entry(ctx, op, prec, d_out, d_in, true);
}

int qed3_apply_async(qed3_ctx* ctx, int op, int prec, void* d_out, const void* d_in) {
  return apply_entry(ctx, op, prec, d_out, d_in, false);
}

int qed3_ctx_set_stream(qed3_ctx* ctx, struct ihipStream_t* stream) {
  if (!ctx) return fail(QED3_ERR_ARG, "null context");
  return guarded([&]() -> int {
    ctx->D->set_stream(stream);
    return QED3_OK;
  });
}

namespace {
int cg_entry(qed3_ctx* ctx, void* d_x, const void* d_b, const qed3_cg_opts* opts,
             qed3_cg_report* rep, qed3::KernelTimes* times) {
  if (!ctx || !d_x || !d_b || !opts) return fail(QED3_ERR_ARG, "null argument");
  if (d_x == d_b) return fail(QED3_ERR_ARG, "x and b must differ");
  if (!(opts->tol > 0) || opts->maxiter < 1) return fail(QED3_ERR_ARG, "bad tolerance/maxiter");
  if (opts->solver != QED3_SOLVER_CG_F64 && opts->solver != QED3_SOLVER_CG_MIXED &&
      opts->solver != QED3_SOLVER_CG_HALF)
    return fail(QED3_ERR_ARG, "bad solver");
  if (!ctx->D->has_gauge()) return fail(QED3_ERR_STATE, "no gauge field set");
  return guarded([&]() -> int {
    qed3::CGOptions o;
    o.tol = opts->tol;
    o.maxiter = opts->maxiter;
    o.mixed = opts->solver != QED3_SOLVER_CG_F64;
    o.half = opts->solver == QED3_SOLVER_CG_HALF;
    o.delta = opts->delta > 0 ? opts->delta : 0.1;
    o.chunk = opts->check_every;
    o.times = times;
    const qed3::CGResult r = qed3::cg_solve(*ctx->D, static_cast<qed3::Spin<double>*>(d_x),
                                            static_cast<const qed3::Spin<double>*>(d_b), o);
    if (rep) {
      rep->status = r.status == 0 ? QED3_OK : (r.status == 4 ? QED3_ERR_NAN : QED3_ERR_NOCONV);
      rep->iterations = r.iterations;
      rep->reliable_updates = r.reliable_updates;
      rep->true_resid = r.true_resid;
      rep->seconds = r.seconds;
    }
    if (r.status == 4) return fail(QED3_ERR_NAN, "CG breakdown (non-finite or zero curvature)");
    if (r.status != 0) return fail(QED3_ERR_NOCONV, "CG did not reach the tolerance");
    return QED3_OK;
  });
}
}  // namespace

int qed3_cg(qed3_ctx* ctx, void* d_x, const void* d_b, const qed3_cg_opts* opts,
            qed3_cg_report* rep) {
  return cg_entry(ctx, d_x, d_b, opts, rep, nullptr);
}

int qed3_cg_noalloc(qed3_ctx* ctx, void* d_x, const void* d_b, const qed3_cg_opts* opts, qed3_cg_report* rep) {
  if (!ctx || !d_x || !d_b || !opts) return fail(QED3_ERR_ARG, "null argument");
  if (d_x == d_b) return fail(QED3_ERR_ARG, "x and b must differ");
  if (!(opts->tol > 0) || opts->maxiter < 1) return fail(QED3_ERR_ARG, "bad tolerance/maxiter");
  if (opts->solver != QED3_SOLVER_CG_F64) return fail(QED3_ERR_ARG, "qed3_cg_noalloc runs the FP64 CG only");
  if (!ctx->D->has_gauge()) return fail(QED3_ERR_STATE, "no gauge field set");
  return guarded([&]() -> int {
    qed3::CGResult r;
    std::string err;
    const int rc = qed3::cg_noalloc(*ctx->D, d_x, d_b, opts->tol, opts->maxiter, &r, &err);
    if (rc != QED3_OK) return fail(rc, err);
    if (rep) {
      rep->status = r.status == 0 ? QED3_OK : (r.status == 4 ? QED3_ERR_NAN : QED3_ERR_NOCONV);
      rep->iterations = r.iterations;
      rep->reliable_updates = r.reliable_updates;
      rep->true_resid = r.true_resid;
      rep->seconds = r.seconds;
    }
    if (r.status == 4) return fail(QED3_ERR_NAN, "CG breakdown (non-finite or zero curvature)");
    if (r.status != 0) return fail(QED3_ERR_NOCONV, "CG did not reach the tolerance");
    return QED3_OK;
  });
}

int qed3_cg_kernel_times(qed3_ctx* ctx, void* d_x, const void* d_b, const qed3_cg_opts* opts,
                         qed3_cg_report* rep, double* ms) {
  if (!ms) return fail(QED3_ERR_ARG, "null argument");
  qed3::KernelTimes t;
  const int rc = cg_entry(ctx, d_x, d_b, opts, rep, &t);
  for (int k = 0; k < 3; k++) ms[k] = t.ms[k];
  return rc;
}

namespace {
int pcg_entry(qed3_ctx* ctx, qed3_precond* pc, void* d_x, const void* d_b, const qed3_cg_opts* opts,
              qed3_cg_report* rep, qed3::KernelTimes* times, qed3::KernelTimes* ltimes) {
  if (!ctx || !pc || !d_x || !d_b || !opts) return fail(QED3_ERR_ARG, "null argument");
  if (pc->ctx != ctx) return fail(QED3_ERR_ARG, "preconditioner belongs to another context");
  if (d_x == d_b) return fail(QED3_ERR_ARG, "x and b must differ");
  if (!(opts->tol > 0) || opts->maxiter < 1) return fail(QED3_ERR_ARG, "bad tolerance/maxiter");
  if (!ctx->D->has_gauge()) return fail(QED3_ERR_STATE, "no gauge field set");
  return guarded([&]() -> int {
    qed3::CGOptions o;
    o.tol = opts->tol;
    o.maxiter = opts->maxiter;
    o.mixed = o.half = true;
    o.delta = opts->delta > 0 ? opts->delta : 0.01;
    o.chunk = opts->check_every;
    o.times = times;
    const qed3::CGResult r = qed3::pcg_solve(*pc->K, static_cast<qed3::Spin<double>*>(d_x),
                                             static_cast<const qed3::Spin<double>*>(d_b), o, ltimes);
    if (rep) {
      rep->status = r.status == 0 ? QED3_OK : (r.status == 4 ? QED3_ERR_NAN : QED3_ERR_NOCONV);
      rep->iterations = r.iterations;
      rep->reliable_updates = r.reliable_updates;
      rep->true_resid = r.true_resid;
      rep->seconds = r.seconds;
    }
    if (r.status == 4) return fail(QED3_ERR_NAN, "preconditioned CG breakdown (non-finite or <r, M r> <= 0)");
    if (r.status != 0) return fail(QED3_ERR_NOCONV, "preconditioned CG did not reach the tolerance");
    return QED3_OK;
  });
}
}  // namespace

int qed3_precond_create(qed3_ctx* ctx, int layers, qed3_precond** out) {
  if (!ctx || !out) return fail(QED3_ERR_ARG, "null argument");
  if (layers < 1 || layers > 16) return fail(QED3_ERR_ARG, "layers must be in [1,16]");
  if (ctx->p.r != 1.0) return fail(QED3_ERR_ARG, "the time-exact preconditioner needs r = 1");
  return guarded([&]() -> int {
    auto pc = std::make_unique<qed3_precond>();
    pc->ctx = ctx;
    pc->K = std::make_unique<qed3::Horner>(*ctx->D, layers);
    *out = pc.release();
    return QED3_OK;
  });
}

void qed3_precond_destroy(qed3_precond* pc) { delete pc; }

int qed3_precond_nparams(const qed3_precond* pc, int* n) {
  if (!pc || !n) return fail(QED3_ERR_ARG, "null argument");
  *n = pc->K->params().count();
  return QED3_OK;
}

int qed3_precond_get_params(const qed3_precond* pc, double* p) {
  if (!pc || !p) return fail(QED3_ERR_ARG, "null argument");
  pc->K->params().pack(reinterpret_cast<std::complex<double>*>(p));
  return QED3_OK;
}

int qed3_precond_set_params(qed3_precond* pc, const double* p) {
  if (!pc || !p) return fail(QED3_ERR_ARG, "null argument");
  return guarded([&]() -> int {
    qed3::HornerParams q = pc->K->params();
    q.unpack(reinterpret_cast<const std::complex<double>*>(p));
    for (int j = 0; j < q.count(); j++)
      if (!std::isfinite(p[2 * j]) || !std::isfinite(p[2 * j + 1])) return fail(QED3_ERR_ARG, "non-finite parameter");
    pc->K->set_params(q);
    return QED3_OK;
  });
}

int qed3_precond_apply(qed3_precond* pc, int op, void* d_out, const void* d_in) {
  if (!pc || !d_out || !d_in) return fail(QED3_ERR_ARG, "null argument");
  if (op != QED3_OP_D && op != QED3_OP_DDAG) return fail(QED3_ERR_ARG, "op must be QED3_OP_D or QED3_OP_DDAG");
  if (!pc->ctx->D->has_gauge()) return fail(QED3_ERR_STATE, "no gauge field set");
  return guarded([&]() -> int {
    pc->K->apply(op == QED3_OP_DDAG, static_cast<qed3::Spin<double>*>(d_out),
                 static_cast<const qed3::Spin<double>*>(d_in));
    return QED3_OK;
  });
}


int qed3_precond_apply_solver(qed3_precond* pc, int op, void* d_out, const void* d_in) {
  if (!pc || !d_out || !d_in) return fail(QED3_ERR_ARG, "null argument");
  if (op != QED3_OP_D && op != QED3_OP_DDAG && op != QED3_OP_DDDAG)
    return fail(QED3_ERR_ARG, "op must be QED3_OP_D, QED3_OP_DDAG or QED3_OP_DDDAG");
  if (d_out == d_in) return fail(QED3_ERR_ARG, "in-place apply is not supported");
  if (!pc->ctx->D->has_gauge()) return fail(QED3_ERR_STATE, "no gauge field set");
  return guarded([&]() -> int {
    qed3::horner_apply_solver(*pc->K, op == QED3_OP_DDAG ? 1 : (op == QED3_OP_DDDAG ? 2 : 0),
                              static_cast<qed3::Spin<double>*>(d_out), static_cast<const qed3::Spin<double>*>(d_in));
    return QED3_OK;
  });
}

int qed3_precond_loss_grad(qed3_precond* pc, const void* d_v, const void* d_e, int batch, double* loss,
                           double* grad) {
  if (!pc || !d_v || !d_e || !loss) return fail(QED3_ERR_ARG, "null argument");
  if (batch < 1) return fail(QED3_ERR_ARG, "batch must be positive");
  if (!pc->ctx->D->has_gauge()) return fail(QED3_ERR_STATE, "no gauge field set");
  return guarded([&]() -> int {
    std::vector<std::complex<double>> g;
    *loss = pc->K->loss_grad(static_cast<const qed3::Spin<double>*>(d_v),
                             static_cast<const qed3::Spin<double>*>(d_e), batch, grad ? &g : nullptr);
    if (grad)
      for (size_t j = 0; j < g.size(); j++) grad[2 * j] = g[j].real(), grad[2 * j + 1] = g[j].imag();
    return QED3_OK;
  });
}

int qed3_precond_train(qed3_precond* pc, const void* d_v, const void* d_e, int npairs, int batch, int epochs,
                       double lr, double* history) {
  if (!pc || !d_v || !d_e) return fail(QED3_ERR_ARG, "null argument");
  if (npairs < 1 || batch < 1 || epochs < 1 || !(lr > 0)) return fail(QED3_ERR_ARG, "bad training sizes");
  if (!pc->ctx->D->has_gauge()) return fail(QED3_ERR_STATE, "no gauge field set");
  return guarded([&]() -> int {
    const std::vector<double> h = pc->K->train(static_cast<const qed3::Spin<double>*>(d_v),
                                               static_cast<const qed3::Spin<double>*>(d_e), npairs, batch,
                                               epochs, lr);
    if (history)
      for (int k = 0; k < epochs; k++) history[k] = h[k];
    return QED3_OK;
  });
}

int qed3_pcg(qed3_ctx* ctx, qed3_precond* pc, void* d_x, const void* d_b, const qed3_cg_opts* opts,
             qed3_cg_report* rep) {
  return pcg_entry(ctx, pc, d_x, d_b, opts, rep, nullptr, nullptr);
}

int qed3_pcg_kernel_times(qed3_ctx* ctx, qed3_precond* pc, void* d_x, const void* d_b,
                          const qed3_cg_opts* opts, qed3_cg_report* rep, double* ms) {
  if (!ms) return fail(QED3_ERR_ARG, "null argument");
  qed3::KernelTimes t, lt;
  const int rc = pcg_entry(ctx, pc, d_x, d_b, opts, rep, &t, &lt);
  for (int k = 0; k < 3; k++) ms[k] = t.ms[k];
  for (int k = 0; k < qed3::KernelTimes::kClasses; k++) ms[3 + k] = lt.ms[k];
  return rc;
}

int qed3_net_create(qed3_ctx* ctx, int layers, int seed, qed3_net** out) {
  if (!ctx || !out) return fail(QED3_ERR_ARG, "null argument");
  if (layers < 2 || layers > 16) return fail(QED3_ERR_ARG, "layers must be in [2,16]");
  return guarded([&]() -> int {
    auto n = std::make_unique<qed3_net>();
    n->ctx = ctx;
    n->net = std::make_unique<qed3::Net>(*ctx->D, layers, seed);
    *out = n.release();
    return QED3_OK;
  });
}

void qed3_net_destroy(qed3_net* net) { delete net; }

int qed3_net_nparams(const qed3_net* net, int* n) {
  if (!net || !n) return fail(QED3_ERR_ARG, "null argument");
  *n = int(net->net->nparams());
  return QED3_OK;
}

int qed3_net_get_params(const qed3_net* net, double* p) {
  if (!net || !p) return fail(QED3_ERR_ARG, "null argument");
  net->net->get_params(p);
  return QED3_OK;
}

int qed3_net_set_params(qed3_net* net, const double* p) {
  if (!net || !p) return fail(QED3_ERR_ARG, "null argument");
  return guarded([&]() -> int {
    net->net->set_params(p);
    return QED3_OK;
  });
}

int qed3_net_get_skip(const qed3_net* net, double* c0) {
  if (!net || !c0) return fail(QED3_ERR_ARG, "null argument");
  *c0 = net->net->skip();
  return QED3_OK;
}

int qed3_net_set_skip(qed3_net* net, double c0) {
  if (!net || !std::isfinite(c0) || c0 < 0) return fail(QED3_ERR_ARG, "skip must be finite and >= 0");
  net->net->set_skip(c0);
  return QED3_OK;
}

int qed3_net_apply(qed3_net* net, void* d_out, const void* d_in, int batch) {
  if (!net || !d_out || !d_in || batch < 1) return fail(QED3_ERR_ARG, "bad argument");
  if (!net->ctx->D->has_gauge()) return fail(QED3_ERR_STATE, "no gauge field set");
  return guarded([&]() -> int {
    net->net->apply(static_cast<qed3::Spin<float>*>(d_out), static_cast<const qed3::Spin<float>*>(d_in), batch);
    QED3_HIP(hipStreamSynchronize(net->ctx->D->stream()));
    return QED3_OK;
  });
}

int qed3_net_apply_adjoint(qed3_net* net, void* d_out, const void* d_in, int batch) {
  if (!net || !d_out || !d_in || batch < 1) return fail(QED3_ERR_ARG, "bad argument");
  if (d_out == d_in) return fail(QED3_ERR_ARG, "d_out must not alias d_in");
  if (!net->ctx->D->has_gauge()) return fail(QED3_ERR_STATE, "no gauge field set");
  return guarded([&]() -> int {
    net->net->apply_adj(static_cast<qed3::Spin<float>*>(d_out), static_cast<const qed3::Spin<float>*>(d_in),
                        batch);
    QED3_HIP(hipStreamSynchronize(net->ctx->D->stream()));
    return QED3_OK;
  });
}

int qed3_net_loss_grad(qed3_net* net, const void* d_r, const void* d_e, int batch, double* loss,
                       double* grad) {
  if (!net || !d_r || !d_e || batch < 1) return fail(QED3_ERR_ARG, "bad argument");
  if (!net->ctx->D->has_gauge()) return fail(QED3_ERR_STATE, "no gauge field set");
  return guarded([&]() -> int {
    std::vector<double> g;
    const double l = net->net->loss_grad(static_cast<const qed3::Spin<float>*>(d_r),
                                         static_cast<const qed3::Spin<float>*>(d_e), batch, g);
    if (loss) *loss = l;
    if (grad) std::copy(g.begin(), g.end(), grad);
    return QED3_OK;
  });
}

int qed3_net_train(qed3_net* net, const void* d_r, const void* d_e, int npairs, int batch,
                   int epochs, double lr, double* history) {
  if (!net || !d_r || !d_e || npairs < 1 || batch < 1 || batch > npairs || epochs < 1)
    return fail(QED3_ERR_ARG, "bad argument");
  if (!net->ctx->D->has_gauge()) return fail(QED3_ERR_STATE, "no gauge field set");
  return guarded([&]() -> int {
    std::vector<double> h;
    net->net->train(static_cast<const qed3::Spin<float>*>(d_r), static_cast<const qed3::Spin<float>*>(d_e),
                    npairs, batch, epochs, lr, &h);
    if (history) std::copy(h.begin(), h.end(), history);
    return QED3_OK;
  });
}

int qed3_fcg(qed3_ctx* ctx, qed3_net* net, void* d_x, const void* d_b, const qed3_cg_opts* opts,
             qed3_cg_report* rep, int* precond_applies, int* used_fallback) {
  if (!ctx || !net || !d_x || !d_b || !opts) return fail(QED3_ERR_ARG, "null argument");
  if (net->ctx != ctx) return fail(QED3_ERR_ARG, "net belongs to another context");
  if (!(opts->tol > 0) || opts->maxiter < 1) return fail(QED3_ERR_ARG, "bad tolerance/maxiter");
  if (!ctx->D->has_gauge()) return fail(QED3_ERR_STATE, "no gauge field set");
  return guarded([&]() -> int {
    qed3::CGOptions o;
    o.tol = opts->tol;
    o.maxiter = opts->maxiter;
    o.delta = opts->delta > 0 ? opts->delta : 0.1;
    qed3::FCGStats s;
    const qed3::CGResult r = qed3::fcg_solve(*ctx->D, *net->net, static_cast<qed3::Spin<double>*>(d_x),
                                             static_cast<const qed3::Spin<double>*>(d_b), o, &s);
    if (rep) {
      rep->status = r.status == 0 ? QED3_OK : QED3_ERR_NOCONV;
      rep->iterations = r.iterations;
      rep->reliable_updates = r.reliable_updates;
      rep->true_resid = r.true_resid;
      rep->seconds = r.seconds;
    }
    if (precond_applies) *precond_applies = s.precond_applies;
    if (used_fallback) *used_fallback = s.fallback;
    if (r.status != 0) return fail(QED3_ERR_NOCONV, "FCG did not reach the tolerance");
    return QED3_OK;
  });
}

int qed3_batch_cg(qed3_ctx* const* ctxs, int nsys, void* const* d_x, const void* const* d_b,
                  const qed3_cg_opts* opts, qed3_cg_report* reps) {
  if (!ctxs || nsys < 1 || !d_x || !d_b || !opts) return fail(QED3_ERR_ARG, "bad argument");
  if (!(opts->tol > 0) || opts->maxiter < 1) return fail(QED3_ERR_ARG, "bad tolerance/maxiter");
  std::vector<qed3::Dirac*> Ds;
  std::vector<qed3::Spin<double>*> xs;
  std::vector<const qed3::Spin<double>*> bs;
  for (int k = 0; k < nsys; k++) {
    if (!ctxs[k] || !d_x[k] || !d_b[k]) return fail(QED3_ERR_ARG, "null system");
    if (!ctxs[k]->D->has_gauge()) return fail(QED3_ERR_STATE, "a system has no gauge field");
    Ds.push_back(ctxs[k]->D.get());
    xs.push_back(static_cast<qed3::Spin<double>*>(d_x[k]));
    bs.push_back(static_cast<const qed3::Spin<double>*>(d_b[k]));
  }
  return guarded([&]() -> int {
    qed3::CGOptions o;
    o.tol = opts->tol;
    o.maxiter = opts->maxiter;
    o.mixed = true;
    o.delta = opts->delta > 0 ? opts->delta : 0.1;
    o.chunk = opts->check_every;
    const std::vector<qed3::CGResult> r = qed3::batch_cg(Ds, xs, bs, o);
    int worst = QED3_OK;
    for (int k = 0; k < nsys; k++) {
      const int st = r[k].status == 0 ? QED3_OK : (r[k].status == 4 ? QED3_ERR_NAN : QED3_ERR_NOCONV);
      if (st != QED3_OK) worst = st;
      if (reps) {
        reps[k].status = st;
        reps[k].iterations = r[k].iterations;
        reps[k].reliable_updates = r[k].reliable_updates;
        reps[k].true_resid = r[k].true_resid;
        reps[k].seconds = r[k].seconds;
      }
    }
    if (worst != QED3_OK) return fail(worst, "a batched system did not converge");
    return QED3_OK;
  });
}

int qed3_dot(qed3_ctx* ctx, const void* d_x, const void* d_y, double* re, double* im) {
  if (!ctx || !d_x || !d_y) return fail(QED3_ERR_ARG, "null argument");
  return guarded([&]() -> int {
    ctx->D->dot(static_cast<const qed3::Spin<double>*>(d_x),
                static_cast<const qed3::Spin<double>*>(d_y), re, im);
    return QED3_OK;
  });
}

int qed3_pf_action(qed3_ctx* ctx, const void* d_phi, const void* d_eta, double* S) {
  if (!ctx || !d_phi || !d_eta || !S) return fail(QED3_ERR_ARG, "null argument");
  return guarded([&]() -> int {
    double im = 0.0;
    ctx->D->dot(static_cast<const qed3::Spin<double>*>(d_phi), static_cast<const qed3::Spin<double>*>(d_eta), S,
                &im);
    return QED3_OK;
  });
}

int qed3_pf_force(qed3_ctx* ctx, const void* d_eta, double* d_fs, double* d_ft) {
  if (!ctx || !d_eta || !d_fs || !d_ft) return fail(QED3_ERR_ARG, "null argument");
  if (!ctx->D->has_gauge()) return fail(QED3_ERR_STATE, "no gauge field set");
  return guarded([&]() -> int {
    ctx->D->force(static_cast<const qed3::Spin<double>*>(d_eta), d_fs, d_ft);
    return QED3_OK;
  });
}

int qed3_gauge_action(qed3_ctx* ctx, double gsq, int compact, const double* spatial,
                      const double* temporal, double* S, double* fs, double* ft) {
  if (!ctx || !spatial || !temporal) return fail(QED3_ERR_ARG, "null argument");
  return guarded([&]() -> int {
    const qed3::GaugeAction A(ctx->D->tables(), gsq, ctx->p.at, compact != 0);
    if (S) *S = A.S(spatial, temporal);
    if (fs && ft) A.force(spatial, temporal, fs, ft);
    return QED3_OK;
  });
}

int qed3_plaquette_correlator(const qed3_geom* g, const qed3_params* p, const double* spatial,
                              double* corr) {
  if (!g || !p || !spatial || !corr) return fail(QED3_ERR_ARG, "null argument");
  if ((p->lattice != 0 && p->lattice != 1) || p->Lt < 1) return fail(QED3_ERR_ARG, "bad params");
  return guarded([&]() -> int {
    const qed3::HopTables h = qed3::make_hop_tables(g->g, p->lattice, p->Lt, p->r, p->M5, p->at);
    const qed3::GaugeAction A(h, 1.0, p->at, false);
    A.plaquette_correlator(spatial, corr);
    return QED3_OK;
  });
}

int qed3_hmc_create(qed3_ctx* ctx, const qed3_hmc_params* p, qed3_hmc** out) {
  if (!ctx || !p || !out) return fail(QED3_ERR_ARG, "null argument");
  if (p->solver != QED3_SOLVER_CG_F64 && p->solver != QED3_SOLVER_CG_MIXED &&
      p->solver != QED3_SOLVER_CG_HALF && p->solver != QED3_SOLVER_PCG)
    return fail(QED3_ERR_ARG, "bad solver");
  return guarded([&]() -> int {
    qed3::HMCParams hp;
    hp.gsq = p->gsq;
    hp.compact = p->compact != 0;
    hp.npf = p->npf;
    hp.tau = p->tau;
    hp.nsteps = p->nsteps;
    hp.nsteps_inner = p->nsteps_inner;
    hp.tol = p->tol;
    hp.mixed = p->solver != QED3_SOLVER_CG_F64;
    hp.half = p->solver == QED3_SOLVER_CG_HALF;
    hp.pcg = p->solver == QED3_SOLVER_PCG;
    hp.seed = p->seed;
    auto h = std::make_unique<qed3_hmc>();
    h->ctx = ctx;
    h->hmc = std::make_unique<qed3::HMC>(*ctx->D, hp);
    *out = h.release();
    return QED3_OK;
  });
}

void qed3_hmc_destroy(qed3_hmc* h) { delete h; }

int qed3_hmc_set_steps(qed3_hmc* h, double tau, int nsteps, int nsteps_inner) {
  if (!h) return fail(QED3_ERR_ARG, "null handle");
  if (!(tau > 0) || nsteps < 1 || nsteps_inner < 1) return fail(QED3_ERR_ARG, "bad trajectory length or steps");
  return guarded([&]() -> int {
    h->hmc->set_steps(tau, nsteps, nsteps_inner);
    return QED3_OK;
  });
}

int qed3_hmc_set_overlap(qed3_hmc* h, qed3_overlap* ov, int refresh) {
  if (!h) return fail(QED3_ERR_ARG, "null argument");
  if (ov && ov->ctx != h->ctx) return fail(QED3_ERR_ARG, "overlap handle of another context");
  return guarded([&]() -> int {
    h->hmc->set_overlap(ov ? ov->ov.get() : nullptr, refresh != 0);
    return QED3_OK;
  });
}

int qed3_hmc_trajectory(qed3_hmc* h, double* spatial, double* temporal, int no_reject,
                        qed3_hmc_report* rep) {
  if (!h || !spatial || !temporal) return fail(QED3_ERR_ARG, "null argument");
  return guarded([&]() -> int {
    const qed3::HMCReport r = h->hmc->trajectory(spatial, temporal, no_reject != 0);
    if (rep) {
      rep->dH = r.dH;
      rep->h0 = r.h0;
      rep->h1 = r.h1;
      rep->accept = r.accept;
      rep->cg_solves = r.cg_solves;
      rep->cg_iterations = r.cg_iterations;
      rep->max_resid = r.max_resid;
      rep->seconds = r.seconds;
    }
    return QED3_OK;
  });
}

int qed3_zolotarev(int n, double k, int* size, double* A, double* c, double* cp, double* consts) {
  if (!size || !A || !c || !cp || !consts) return fail(QED3_ERR_ARG, "null argument");
  return guarded([&]() -> int {
    const qed3::Zolotarev z(n, k);
    *size = z.size;
    for (int m = 0; m < z.size; m++) {
      A[m] = z.A[m];
      c[m] = z.c[m];
      cp[m] = z.cp[m];
    }
    consts[0] = z.M;
    consts[1] = z.lambda_inv;
    consts[2] = z.E;
    consts[3] = z.C;
    consts[4] = z.delta();
    return QED3_OK;
  });
}

int qed3_multishift_cg(qed3_ctx* ctx, int nshift, const double* shifts, void* const* d_x,
                       const void* d_b, const qed3_cg_opts* opts, qed3_cg_report* reps) {
  if (!ctx || !shifts || !d_x || !d_b || !opts || nshift < 1) return fail(QED3_ERR_ARG, "bad argument");
  if (!ctx->D->has_gauge()) return fail(QED3_ERR_STATE, "no gauge field set");
  return guarded([&]() -> int {
    qed3::CGOptions o;
    o.tol = opts->tol;
    o.maxiter = opts->maxiter;
    std::vector<double> sh(shifts, shifts + nshift);
    std::vector<qed3::Spin<double>*> xs(nshift);
    for (int i = 0; i < nshift; i++) {
      if (!d_x[i] || d_x[i] == d_b) throw std::invalid_argument("solution vectors must be distinct from b");
      xs[i] = static_cast<qed3::Spin<double>*>(d_x[i]);
    }
    const auto r = qed3::multishift_cg(*ctx->D, sh, xs, static_cast<const qed3::Spin<double>*>(d_b), o);
    int worst = QED3_OK;
    for (int i = 0; i < nshift; i++) {
      if (reps) {
        reps[i].status = r[i].status;
        reps[i].iterations = r[i].iterations;
        reps[i].reliable_updates = 0;
        reps[i].true_resid = r[i].true_resid;
        reps[i].seconds = r[i].seconds;
      }
      if (r[i].status != 0) worst = QED3_ERR_NOCONV;
    }
    if (worst != QED3_OK) return fail(worst, "a shifted system did not reach the tolerance");
    return QED3_OK;
  });
}

int qed3_spectral_bounds(qed3_ctx* ctx, double tol, int maxiter, double* mu_min, double* mu_max) {
  if (!ctx || !mu_min || !mu_max || !(tol > 0) || maxiter < 1) return fail(QED3_ERR_ARG, "bad argument");
  if (!ctx->D->has_gauge()) return fail(QED3_ERR_STATE, "no gauge field set");
  return guarded([&]() -> int {
    qed3::spectral_bounds(*ctx->D, tol, maxiter, mu_min, mu_max);
    return QED3_OK;
  });
}

int qed3_overlap_create(qed3_ctx* ctx, int n, double lambda_min, double lambda_max, double tol,
                        qed3_overlap** out) {
  if (!ctx || !out || n < 1 || !(tol > 0)) return fail(QED3_ERR_ARG, "bad argument");
  return guarded([&]() -> int {
    auto o = std::make_unique<qed3_overlap>();
    o->ctx = ctx;
    o->ov = std::make_unique<qed3::Overlap>(*ctx->D, n, lambda_min, lambda_max, tol);
    *out = o.release();
    return QED3_OK;
  });
}

void qed3_overlap_destroy(qed3_overlap* ov) { delete ov; }

int qed3_overlap_apply(qed3_overlap* ov, int op, void* d_out, const void* d_in) {
  if (!ov || !d_out || !d_in || d_out == d_in) return fail(QED3_ERR_ARG, "bad argument");
  if (op != QED3_OP_D && op != QED3_OP_DDAG) return fail(QED3_ERR_ARG, "op must be QED3_OP_D or QED3_OP_DDAG");
  if (!ov->ctx->D->has_gauge()) return fail(QED3_ERR_STATE, "no gauge field set");
  return guarded([&]() -> int {
    ov->ov->apply(op == QED3_OP_DDAG, static_cast<qed3::Spin<double>*>(d_out),
                  static_cast<const qed3::Spin<double>*>(d_in));
    return QED3_OK;
  });
}

int qed3_overlap_force(qed3_overlap* ov, const void* d_eta, double* d_fs, double* d_ft) {
  if (!ov || !d_eta || !d_fs || !d_ft || d_fs == d_ft) return fail(QED3_ERR_ARG, "bad argument");
  if (!ov->ctx->D->has_gauge()) return fail(QED3_ERR_STATE, "no gauge field set");
  return guarded([&]() -> int {
    ov->ov->force(static_cast<const qed3::Spin<double>*>(d_eta), d_fs, d_ft);
    return QED3_OK;
  });
}

int qed3_overlap_solve(qed3_overlap* ov, void* d_x, const void* d_b, double tol, int maxiter,
                       qed3_cg_report* rep) {
  if (!ov || !d_x || !d_b || d_x == d_b || !(tol > 0) || maxiter < 0) return fail(QED3_ERR_ARG, "bad argument");
  if (!ov->ctx->D->has_gauge()) return fail(QED3_ERR_STATE, "no gauge field set");
  return guarded([&]() -> int {
    const qed3::CGResult r = ov->ov->solve_dhd(static_cast<qed3::Spin<double>*>(d_x),
                                               static_cast<const qed3::Spin<double>*>(d_b), tol, maxiter);
    if (rep) {
      rep->status = r.status == 0 ? QED3_OK : QED3_ERR_NOCONV;
      rep->iterations = r.iterations;
      rep->reliable_updates = 0;
      rep->true_resid = r.true_resid;
      rep->seconds = r.seconds;
    }
    if (r.status != 0) return fail(QED3_ERR_NOCONV, "overlap CG did not reach the tolerance");
    return QED3_OK;
  });
}

int qed3_gauge_write(const char* path, int Lt, int E, int V, const double* spatial,
                     const double* temporal) {
  if (!path || !spatial || !temporal || Lt < 1 || E < 1 || V < 1) return fail(QED3_ERR_ARG, "bad argument");
  std::FILE* f = std::fopen(path, "wb");
  if (!f) return fail(QED3_ERR_ARG, std::string("cannot write ") + path);
  const size_t nS = size_t(Lt) * E, nT = size_t(Lt) * V;
  const bool ok = std::fwrite(spatial, sizeof(double), nS, f) == nS &&
                  std::fwrite(temporal, sizeof(double), nT, f) == nT;
  const bool closed = std::fclose(f) == 0;
  if (!ok || !closed) return fail(QED3_ERR_ARG, std::string("write failed: ") + path);
  return QED3_OK;
}

int qed3_gauge_read(const char* path, int Lt, int E, int V, double* spatial, double* temporal) {
  if (!path || !spatial || !temporal || Lt < 1 || E < 1 || V < 1) return fail(QED3_ERR_ARG, "bad argument");
  std::FILE* f = std::fopen(path, "rb");
  if (!f) return fail(QED3_ERR_ARG, std::string("cannot read ") + path);
  const size_t nS = size_t(Lt) * E, nT = size_t(Lt) * V;
  bool ok = std::fread(spatial, sizeof(double), nS, f) == nS &&
            std::fread(temporal, sizeof(double), nT, f) == nT;
  char extra;
  ok = ok && std::fread(&extra, 1, 1, f) == 0;  // exactly Lt*(E+V) doubles
  std::fclose(f);
  if (!ok) return fail(QED3_ERR_ARG, std::string("gauge file size does not match Lt*(E+V): ") + path);
  return QED3_OK;
}

int qed3_rng_create(int Lt, int E, int V, int seed, qed3_rng** out) {
  if (!out || Lt < 1 || E < 1 || V < 1) return fail(QED3_ERR_ARG, "bad argument");
  return guarded([&]() -> int {
    auto r = std::make_unique<qed3_rng>();
    r->rng = std::make_unique<qed3::ParallelRng>(Lt, E, V, seed);
    *out = r.release();
    return QED3_OK;
  });
}

void qed3_rng_destroy(qed3_rng* r) { delete r; }

int qed3_rng_save(const qed3_rng* r, const char* path) {
  if (!r || !path) return fail(QED3_ERR_ARG, "null argument");
  return guarded([&]() -> int {
    r->rng->save(path);
    return QED3_OK;
  });
}

int qed3_rng_load(qed3_rng* r, const char* path) {
  if (!r || !path) return fail(QED3_ERR_ARG, "null argument");
  return guarded([&]() -> int {
    r->rng->load(path);
    return QED3_OK;
  });
}

int qed3_rng_gaussian(qed3_rng* r, int which, int s, int i, double* out) {
  if (!r || !out) return fail(QED3_ERR_ARG, "null argument");
  const qed3::ParallelRng& p = *r->rng;
  if (which == 0) {
    *out = r->rng->gaussian_master();
  } else if (which == 1 && s >= 0 && s < p.Lt() && i >= 0 && i < p.E()) {
    *out = r->rng->gaussian_link(s, i);
  } else if (which == 2 && s >= 0 && s < p.Lt() && i >= 0 && i < p.V()) {
    *out = r->rng->gaussian_site(s, i);
  } else {
    return fail(QED3_ERR_ARG, "stream index out of range");
  }
  return QED3_OK;
}

int qed3_hmc_rng_save(const qed3_hmc* h, const char* path) {
  if (!h || !path) return fail(QED3_ERR_ARG, "null argument");
  return guarded([&]() -> int {
    h->hmc->rng().save(path);
    return QED3_OK;
  });
}

int qed3_hmc_rng_load(qed3_hmc* h, const char* path) {
  if (!h || !path) return fail(QED3_ERR_ARG, "null argument");
  return guarded([&]() -> int {
    h->hmc->rng().load(path);
    return QED3_OK;
  });
}

int qed3_time_copy(size_t bytes, int reps, double* ms) {
  if (!ms || bytes < 16) return fail(QED3_ERR_ARG, "bad argument");
  return guarded([&]() -> int {
    *ms = qed3::time_copy(bytes, reps);
    return QED3_OK;
  });
}

int qed3_tile_info(const qed3_ctx* ctx, int op, int prec, int* out) {
  if (!ctx || !out) return fail(QED3_ERR_ARG, "null argument");
  const bool dag = op == QED3_OP_DDAG;
  const qed3::StencilTile t = prec != QED3_PREC_F64 ? ctx->D->tile<float>(dag) : ctx->D->tile<double>(dag);
  out[0] = t.block;
  out[1] = t.ts;
  out[2] = t.remap;
  out[3] = t.lds;
  return QED3_OK;
}

int qed3_set_tile(qed3_ctx* ctx, int op, int prec, const int* in) {
  if (!ctx || !in) return fail(QED3_ERR_ARG, "null argument");
  if (op != QED3_OP_D && op != QED3_OP_DDAG) return fail(QED3_ERR_ARG, "op must be D or DDAG");
  if (in[0] != 64 && in[0] != 128 && in[0] != 256) return fail(QED3_ERR_ARG, "block must be 64/128/256");
  if (in[1] < 1 || in[1] > 64) return fail(QED3_ERR_ARG, "time tile must be in [1,64]");
  return guarded([&]() -> int {
    ctx->D->set_tile(prec != QED3_PREC_F64, op == QED3_OP_DDAG,
                     qed3::StencilTile{in[0], in[1], in[2] ? 1 : 0, in[3] ? 1 : 0});
    return QED3_OK;
  });
}

int qed3_autotune(qed3_ctx* ctx) {
  if (!ctx) return fail(QED3_ERR_ARG, "null context");
  return guarded([&]() -> int {
    ctx->D->autotune();
    qed3::tune_cg_tiles(*ctx->D);
    return QED3_OK;
  });
}

int qed3_time_apply(qed3_ctx* ctx, int op, int prec, int warmup, int reps, double* ms) {
  if (!ctx || !ms) return fail(QED3_ERR_ARG, "null argument");
  if (op != QED3_OP_D && op != QED3_OP_DDAG) return fail(QED3_ERR_ARG, "op must be D or DDAG");
  if (prec != QED3_PREC_F64 && prec != QED3_PREC_F32 && prec != QED3_PREC_F16)
    return fail(QED3_ERR_ARG, "bad precision");
  if (!ctx->D->has_gauge()) return fail(QED3_ERR_STATE, "no gauge field set");
  return guarded([&]() -> int {
    *ms = ctx->D->time_apply(op == QED3_OP_DDAG, prec, warmup, reps);
    return QED3_OK;
  });
}

}  // extern "C"
