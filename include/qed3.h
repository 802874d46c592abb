/* qed3_amd: MI355X (gfx950) Wilson-Dirac operator on the refined icosahedron S^2 x time,
 * CG on D^dagger D and the pseudofermion force -- C ABI.
 *
 * Conventions shared by every entry point:
 *  - Return value: 0 = QED3_OK, otherwise a QED3_ERR_* code; qed3_last_error() gives text.
 *  - Spinor fields use the reference's layout (src/both_3d/dirac_ext.h:63): complex index
 *    i = Nx*s + 2*x + a with Nx = 2V, s in [0,Lt), x in [0,V), a in {0,1}; FP64 fields
 *    are interleaved (re,im) doubles, FP32 fields interleaved floats.
 *  - Gauge angles use the reference's GaugeExt layout (src/both_3d/gauge_ext.h:13-14):
 *    spatial[Lt][E] along the stored link orientation (i<j), temporal[Lt][V].
 *  - Pointers named d_* are device pointers; everything runs on the current HIP device
 *    on a stream owned by the context.  A context is not thread safe.
 */
#ifndef QED3_H
#define QED3_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  QED3_OK = 0,
  QED3_ERR_ARG = 1,    /* invalid argument or size */
  QED3_ERR_HIP = 2,    /* HIP runtime error */
  QED3_ERR_NOCONV = 3, /* solver hit maxiter before reaching the tolerance */
  QED3_ERR_NAN = 4,    /* non-finite value / breakdown in a solver */
  QED3_ERR_STATE = 5   /* call order violated (e.g. no gauge field set) */
};

enum { QED3_LATTICE_SIMPLICIAL = 0, QED3_LATTICE_DUAL = 1 };
enum { QED3_BC_ANTIPERIODIC = 0, QED3_BC_PERIODIC = 1 };
enum { QED3_OP_D = 0, QED3_OP_DDAG = 1, QED3_OP_DDAGD = 2, QED3_OP_DDDAG = 3 };
/* field precision: complex double, complex float, or binary16 pairs (4 x 16 bit per
 * point, FP32 arithmetic, 16-bit fixed-point link angles: the half-storage CG's operator) */
enum { QED3_PREC_F64 = 0, QED3_PREC_F32 = 1, QED3_PREC_F16 = 2 };
/* FP64 CG; FP32 inner CG with FP64 reliable updates; the same with the inner r, p and w
 * stored in IEEE binary16 (scaled to unit rms at every reliable update) */
enum { QED3_SOLVER_CG_F64 = 0, QED3_SOLVER_CG_MIXED = 1, QED3_SOLVER_CG_HALF = 2 };
/* The learned method: flexible CG with the time-exact preconditioner (qed3_pcg, four layers,
 * Neumann coefficients); a solver of the HMC's pseudofermion solves (qed3_hmc_params). */
enum { QED3_SOLVER_PCG = 3 };
enum { QED3_HOST = 0, QED3_DEVICE = 1 };

const char* qed3_last_error(void);

/* ---- device memory plumbing (hipMalloc/hipMemcpy on the current device) ---- */
int qed3_malloc(void** d_ptr, size_t bytes);
int qed3_free(void* d_ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device */
int qed3_memcpy(void* dst, const void* src, size_t bytes, int kind);
int qed3_device_synchronize(void);

/* ---- geometry: refined icosahedron, its dual, spin structure ----------------------
 * Replaces the offline table generator and its .dat files
 * (integrator/get_geometry.cc:41-435, integrator/test_sorting_simp.cc:66-470,
 *  integrator/test_sorting.cc:57-450) that src/both_3d/s2n_simp.h:43-269 reads. */
typedef struct qed3_geom qed3_geom;
int qed3_geom_create(int n_refine, qed3_geom** out);
void qed3_geom_destroy(qed3_geom* g);
/* V = 10n^2+2 sites, E = 30n^2 links, F = 20n^2 triangles */
int qed3_geom_counts(const qed3_geom* g, int* V, int* E, int* F);
/* Export every table (host arrays sized from qed3_geom_counts):
 *  sites[3V], faces[3F] (ccw), links[2E] (i<j), nn_offsets[V+1] + nns[2E],
 *  dual_sites[3F], dual_links[2E] (faces sharing links[k]), dual_nns[3F],
 *  dual_face_offsets[V+1] + dual_faces[3F] (ccw ring of faces around each site),
 *  alpha[2E] (alpha_ij, alpha_ji per simplicial link), omega[E] (omega_ij),
 *  dual_alpha[2E], dual_omega[E] (same for dual links). */
int qed3_geom_export(const qed3_geom* g, double* sites, int* faces, int* links,
                     int* nn_offsets, int* nns, double* dual_sites, int* dual_links,
                     int* dual_nns, int* dual_face_offsets, int* dual_faces, double* alpha,
                     double* omega, double* dual_alpha, double* dual_omega);

/* Build a geometry from external tables with the same meaning and sizes as
 * qed3_geom_export (e.g. the reference's QFE-generated .dat files, read as in
 * s2n_simp.h:43-200 and dirac_simp.h:12-59).  alpha/omega follow each link's stored
 * orientation (links[k] = (i,j): alpha[2k] = alpha_ij, alpha[2k+1] = alpha_ji). */
int qed3_geom_import(int n_refine, int V, int E, int F, const double* sites, const int* faces,
                     const int* links, const int* nn_offsets, const int* nns,
                     const double* dual_sites, const int* dual_links, const int* dual_nns,
                     const int* dual_face_offsets, const int* dual_faces, const double* alpha,
                     const double* omega, const double* dual_alpha, const double* dual_omega,
                     qed3_geom** out);

/* ---- seeded synthetic inputs (host) ------------------------------------------------
 * The reference's RNG scheme (src/both_3d/rng.h:91-128): a master std::mt19937_64
 * seeded with `seed` reseeds one stream per (t,link), then one per (t,site).
 * qed3_random_gauge: spatial[Lt][E] from the link streams, temporal[Lt][V] from the
 *   site streams (the order of GaugeExt::gaussian, gauge_ext.h:146-149), each angle
 *   von Mises(0, kappa) (Best & Fisher 1979); kappa <= 0 gives uniform [-pi, pi).
 * qed3_random_source: xi[2*V*Lt complex] = (g1 + i g2)/sqrt2 per component from the
 *   site streams (ParallelRngExt::fill_gaussian, rng.h:122-128), real part first. */
int qed3_random_gauge(int Lt, int E, int V, int seed, double kappa, double* spatial,
                      double* temporal);
int qed3_random_source(int Lt, int E, int V, int seed, double* xi);

/* ---- operator context --------------------------------------------------------------
 * Replaces DiracExt<Base,BaseDirac>(lattice, m, r, M5, at) (src/both_3d/dirac_ext.h:27-46)
 * plus DWDevice/DiracPf construction (src/both_3d/sparse_dirac.h:52-129,
 * src/both_3d/dirac_pf.h:24-36).  The bare mass of the reference is unused there; the
 * diagonal mass term is M5 (dirac_ext.h:201). */
typedef struct qed3_params {
  int lattice;  /* QED3_LATTICE_* */
  int n_refine; /* refinement level n */
  int Lt;       /* time slices (1 = the 2D operator) */
  int bc_time;  /* QED3_BC_*; the reference is antiperiodic (dirac_ext.h:162-165) */
  double r;     /* Wilson parameter */
  double M5;    /* diagonal mass */
  double at;    /* temporal lattice spacing */
} qed3_params;

typedef struct qed3_ctx qed3_ctx;
int qed3_ctx_create(const qed3_params* p, qed3_ctx** out);
/* same, on a given geometry (p->n_refine is ignored; the geometry is copied) */
int qed3_ctx_create_geom(const qed3_params* p, const qed3_geom* g, qed3_ctx** out);
void qed3_ctx_destroy(qed3_ctx* ctx);
/* sizes of the chosen lattice (dual: V=20n^2, F=10n^2+2) and <ell> */
int qed3_ctx_sizes(const qed3_ctx* ctx, int* V, int* E, int* F, double* mean_ell);
/* Host copies of the derived per-link / per-site quantities: kappa[E], kappa_t[V],
 * link_volume[E], ell[E], dual_area[V] (s2n_simp.h:301-402, dirac_simp.h:270-277,
 * dirac_ext.h:307-320).  Any pointer may be NULL. */
int qed3_ctx_geometry(const qed3_ctx* ctx, double* kappa, double* kappa_t,
                      double* link_volume, double* ell, double* dual_area);

/* Set the U(1) gauge field (replaces DiracPf::update -> DWDevice::update,
 * src/both_3d/dirac_pf.h:61-64, sparse_dirac.h:145-151).  where = QED3_HOST/DEVICE. */
int qed3_set_gauge(qed3_ctx* ctx, const double* spatial, const double* temporal, int where);

/* out = op(in) (replaces DiracPf::{mult,adj,DHD,DDH}_deviceAsyncLaunch,
 * src/both_3d/dirac_pf.h:66-92).  prec selects complex double, complex float or binary16
 * fields (QED3_PREC_*).  Returns when out is written. */
int qed3_apply(qed3_ctx* ctx, int op, int prec, void* d_out, const void* d_in);
/* The same, queued on the context's stream without waiting (the role of the stream argument
 * of LinOp::Async / *_deviceAsyncLaunch, src/both_3d/sparse_matrix.h:4-9,
 * dirac_pf.h:66-92): out is written when the work queued before it on that stream and the
 * apply itself have run.  The first call on a lattice may still wait (it builds tables). */
int qed3_apply_async(qed3_ctx* ctx, int op, int prec, void* d_out, const void* d_in);
/* Every later call on the context queues its device work on `stream` (a hipStream_t; NULL
 * selects the context's own stream again), ordered after what the caller queued there
 * before; the call first waits for the work already queued on the previous stream.  The
 * blocking entry points still return only when their results are complete.  The stream
 * must not be the null stream (the solvers capture hipGraphs on it) and must outlive its
 * use by the context. */
int qed3_ctx_set_stream(qed3_ctx* ctx, struct ihipStream_t* stream);

/* ---- CG on D^dagger D (replaces MatPoly::solve, src/both_3d/matpoly.h:204-267) --- */
typedef struct qed3_cg_opts {
  double tol;      /* stop when ||b - D^dag D x|| <= tol ||b|| (true FP64 residual) */
  int maxiter;     /* inner iteration cap */
  int solver;      /* QED3_SOLVER_* */
  double delta;    /* reliable-update threshold on ||r|| / max ||r|| (mixed solver) */
  int check_every; /* iterations per device-side chunk between host polls (0 = auto) */
} qed3_cg_opts;

typedef struct qed3_cg_report {
  int status;           /* QED3_OK / QED3_ERR_NOCONV / QED3_ERR_NAN */
  int iterations;       /* inner CG iterations */
  int reliable_updates; /* FP64 residual recomputations */
  double true_resid;    /* ||b - D^dag D x|| / ||b|| in FP64 */
  double seconds;       /* wall time of the solve */
} qed3_cg_report;

/* Solve D^dag D x = b; d_x, d_b complex double device fields of length 2*V*Lt. x starts
 * from zero, as in the reference. */
int qed3_cg(qed3_ctx* ctx, void* d_x, const void* d_b, const qed3_cg_opts* opts,
            qed3_cg_report* rep);
/* The same solve with direct launches (no graph) and HIP events around the two stencil
 * kernels of every iteration, for the roofline of the CG's own kernels (no counterpart in
 * the reference).  ms[0] = mean K1 (D with the search-direction update fused), ms[1] =
 * mean K2 (D^dag with the residual update fused; for QED3_SOLVER_CG_HALF the iterations
 * whose x term stays pending), ms[2] = mean K2 that also updates x (half-storage solver,
 * once per four iterations; = ms[1] for the other solvers); means over the iterations run,
 * each kernel's reduction finaliser excluded. */
int qed3_cg_kernel_times(qed3_ctx* ctx, void* d_x, const void* d_b, const qed3_cg_opts* opts,
                         qed3_cg_report* rep, double* ms);
/* The FP64 solve of qed3_cg (opts->solver must be QED3_SOLVER_CG_F64) that allocates
 * nothing: no counterpart in the reference; it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its work fields live in one array of 64 MiB that the library holds on every
 * GPU it runs on until the process ends, which limits the lattice to V*Lt <= 349,520 points
 * (six FP64 fields, 192 bytes of work per point); larger lattices return QED3_ERR_ARG with the limit in
 * qed3_last_error().  Calls share that GPU's array: calls that overlap on one GPU (other
 * host threads, other contexts) race and must be serialised by the caller.  Same arguments,
 * results and report as qed3_cg; the search runs with direct kernel launches. */
int qed3_cg_noalloc(qed3_ctx* ctx, void* d_x, const void* d_b, const qed3_cg_opts* opts,
                    qed3_cg_report* rep);

/* ---- learned preconditioner --------------------------------------------------------
 * No counterpart in the reference (it has no ML part).  A gauge-covariant linear
 * message-passing network on the context's lattice (16 complex channels, `layers` >= 2):
 * input head 1->16, (layers-2) BF16-MFMA channel-mixing layers, output head 16->1; its
 * messages are the link parallel transports exp(i theta) Omega, the gamma-weighted
 * transports and the two temporal transports (with kappa_t/2), its weights may differ per
 * spin component, so the net commutes with U(1) gauge and local spin-frame
 * transformations and can represent D and D^dagger.  It reads the context's current gauge field.
 * Fields are complex float device arrays of `batch` consecutive spinor fields. */
typedef struct qed3_net qed3_net;
int qed3_net_create(qed3_ctx* ctx, int layers, int seed, qed3_net** out);
void qed3_net_destroy(qed3_net* net);
int qed3_net_nparams(const qed3_net* net, int* n); /* complex parameters */
/* parameters as interleaved complex doubles, per output spin a:
 * v[2][5][16], W[layers-2][2][5][16][16], u[2][5][16] */
int qed3_net_get_params(const qed3_net* net, double* p);
int qed3_net_set_params(qed3_net* net, const double* p);
/* the fixed skip coefficient c0 of out = N(r) + c0 r; the default 1/B, B a Gershgorin
 * bound on |D|_2, is set at creation (c0 = 0 leaves the bare network) */
int qed3_net_get_skip(const qed3_net* net, double* c0);
int qed3_net_set_skip(qed3_net* net, double c0);
int qed3_net_apply(qed3_net* net, void* d_out, const void* d_in, int batch);
/* d_out = N^dagger(d_in), the adjoint network (reverse layers, adjoint feature maps) */
int qed3_net_apply_adjoint(qed3_net* net, void* d_out, const void* d_in, int batch);
/* loss = mean_b |N(r_b) + c0 r_b - e_b|^2/|e_b|^2 and d loss/d Re p + i d loss/d Im p (grad: 2n doubles) */
int qed3_net_loss_grad(qed3_net* net, const void* d_r, const void* d_e, int batch, double* loss,
                       double* grad);
/* Adam over npairs (r, e) device pairs in mini-batches; history[epochs] = mean loss/epoch */
int qed3_net_train(qed3_net* net, const void* d_r, const void* d_e, int npairs, int batch,
                   int epochs, double lr, double* history);
/* Flexible CG (FP32 vectors, FP64 reliable updates) on D^dag D preconditioned by
 * M = N N^dagger (net trained as an approximate inverse of D, so M is Hermitian positive
 * semi-definite); plain mixed CG takes over on breakdown or stall and the FP64 true
 * residual decides convergence.
 * precond_applies / used_fallback may be NULL. */
int qed3_fcg(qed3_ctx* ctx, qed3_net* net, void* d_x, const void* d_b, const qed3_cg_opts* opts,
             qed3_cg_report* rep, int* precond_applies, int* used_fallback);

/* ---- time-exact learned preconditioner and flexible CG (the learned method) ---------
 * No counterpart in the reference (it has no preconditioner, MatPoly::solve is plain CG,
 * matpoly.h:204-267).  For r = 1 the operator splits as D = T - S with T = c_x + the
 * temporal hops with the time wrap left out (bidiagonal in time per site and spin row:
 * the hops of dirac_ext.h:158-187 are spin projectors) and S = -(spatial hops + the wrap
 * hop from slice Lt-1 to 0).  The
 * network, a gauge-covariant linear message-passing net on the lattice graph, is
 *   K v = sum_{j<layers} c_j (T^-1 S_th)^j T^-1 v
 * evaluated in Horner form (one launch per layer: an exact first-order recurrence along
 * each site's time column plus one spatial transport step), S_th = S with its same-spin
 * spatial entries of spin row a scaled by A_a and its spin-flip ones by B_a.  Defaults
 * c_j = A_a = B_a = 1 give the Neumann series of D^-1; training adjusts them.
 * Parameters (complex, interleaved doubles): c[0..layers), A0, A1, B0, B1. */
typedef struct qed3_precond qed3_precond;
/* needs the context's r = 1 (QED3_ERR_ARG otherwise); layers in [1,16] */
int qed3_precond_create(qed3_ctx* ctx, int layers, qed3_precond** out);
void qed3_precond_destroy(qed3_precond* pc);
int qed3_precond_nparams(const qed3_precond* pc, int* n);
int qed3_precond_get_params(const qed3_precond* pc, double* p);
int qed3_precond_set_params(qed3_precond* pc, const double* p);
/* d_out = K d_in (op = QED3_OP_D) or K^dagger d_in (op = QED3_OP_DDAG); complex double
 * device fields, FP32 arithmetic on the context's link phases */
int qed3_precond_apply(qed3_precond* pc, int op, void* d_out, const void* d_in);
/* The same with the solver's own kernels (no counterpart in the reference; for checking
 * them): the input rounded to binary16 in the solver's temporal gauge, the
 * parallel-in-time layers (binary16 fields and transports, FP32 recurrences), the output
 * rotated back to the caller's gauge as complex double.  op = QED3_OP_DDDAG applies the
 * solver's M = K K^dagger (K's first layer fused into K^dagger's last). */
int qed3_precond_apply_solver(qed3_precond* pc, int op, void* d_out, const void* d_in);
/* loss = mean_b |K v_b - e_b|^2 / |e_b|^2 over `batch` consecutive device field pairs and
 * its gradient d loss / d Re p + i d loss / d Im p (2 * nparams doubles), forward and
 * backward on the GPU */
int qed3_precond_loss_grad(qed3_precond* pc, const void* d_v, const void* d_e, int batch,
                           double* loss, double* grad);
/* Adam over npairs (v, e) device pairs in mini-batches; history[epochs] = mean loss */
int qed3_precond_train(qed3_precond* pc, const void* d_v, const void* d_e, int npairs, int batch,
                       int epochs, double lr, double* history);
/* Flexible CG (FCG(1)) on D^dag D preconditioned by M = K K^dagger: binary16-stored
 * fields with FP32 arithmetic, FP64 reliable updates every time |r| drops by
 * opts->delta, convergence on the FP64 true residual (opts->solver is ignored).
 * Replaces MatPoly::solve (matpoly.h:204-267) like qed3_cg. */
int qed3_pcg(qed3_ctx* ctx, qed3_precond* pc, void* d_x, const void* d_b, const qed3_cg_opts* opts,
             qed3_cg_report* rep);
/* The same solve with direct launches and HIP events around its kernels, ms[8]:
 * ms[0] K1, ms[1] K2 (its x term left pending), ms[2] 0 (kept for the layout: no K2 updates
 * x; directions still pending when the ring is reused are folded by a pass of their own), ms[3]
 * the first layer of K^dagger (recurrence only), ms[4] a transport layer of K^dagger, ms[5]
 * a transport layer of K, ms[6] the last layer of K^dagger with K's first layer fused in,
 * ms[7] the last layer of K with the FCG reductions (means per launch over the iterations
 * run; 0 where no such launch ran) */
int qed3_pcg_kernel_times(qed3_ctx* ctx, qed3_precond* pc, void* d_x, const void* d_b,
                          const qed3_cg_opts* opts, qed3_cg_report* rep, double* ms);

/* ---- batched CG ----------------------------------------------------------------------
 * Solves D_k^dag D_k x_k = b_k for nsys contexts (any mix of lattices, refinements, Lt and
 * masses, all on the current device) together, binary16-storage inner iteration (as
 * QED3_SOLVER_CG_HALF) with FP64 reliable updates and the true-residual stop (opts->solver
 * is ignored), time tiles from each context's autotuned shapes.  Systems large enough to
 * cover the GPU with one launch (>= 256 blocks of 256) run in rounds: one hipGraph chunk of
 * iterations of every unfinished system queued on the system's own stream, then one host
 * poll per system; systems too small for that share launches in which every CG phase runs
 * over all of them (per-system scalars and convergence masks, iterations captured in a
 * hipGraph).  The reference runs one CG per right-hand side (PseudoFermion::update_eta,
 * pseudofermion.h:51, looped in integrator.h:156-166).  reps[nsys] may be NULL. */
int qed3_batch_cg(qed3_ctx* const* ctxs, int nsys, void* const* d_x, const void* const* d_b,
                  const qed3_cg_opts* opts, qed3_cg_report* reps);

/* ---- BLAS-1 on complex double device fields ---------------------------------------
 * <x,y> = sum conj(x_i) y_i (the role of cublasZdotc in MatPoly::dot, matpoly.h:130-144) */
int qed3_dot(qed3_ctx* ctx, const void* d_x, const void* d_y, double* re, double* im);

/* ---- pseudofermion action ---------------------------------------------------------
 * S_f = Re <phi, eta> with eta = (D^dag D)^{-1} phi from qed3_cg (replaces
 * PseudoFermion::S, pseudofermion.h:53-58); complex double device fields. */
int qed3_pf_action(qed3_ctx* ctx, const void* d_phi, const void* d_eta, double* S);

/* ---- pseudofermion force -----------------------------------------------------------
 * F_l = -2 Re <D eta, dD/dtheta_l eta> for every spatial link (s,l) and temporal link
 * (s,x) at once (replaces GaugeExt::compute's per-link loop, gauge_ext.h:204-222, over
 * DiracPf::precalc_grad/grad_deviceAsyncLaunch, dirac_pf.h:95-121).  Outputs are device
 * arrays in the gauge layout: d_fs[Lt][E], d_ft[Lt][V]. */
int qed3_pf_force(qed3_ctx* ctx, const void* d_eta, double* d_fs, double* d_ft);

/* ---- gauge action and HMC (host FP64 over the GPU fermion path) --------------------
 * U(1) plaquette action on the context's lattice with beta_s = at/(vol_f g^2) per face
 * and beta_t = 2 vol_l/(at ell_l^2 g^2) per temporal plaquette (action_ext.h:150-159):
 * compact = 0 is the reference's 1/2 beta P^2 (action_ext.h:33-80), compact = 1 the
 * Wilson form beta (1 - cos P).  Host arrays; fs/ft (gauge layout) may be NULL. */
int qed3_gauge_action(qed3_ctx* ctx, double gsq, int compact, const double* spatial,
                      const double* temporal, double* S, double* fs, double* ft);

/* Plaquette-angle correlator of gauge.cu:313-343 on the host (no device needed):
 * corr[t] = sum_{s,f} [P_f(s) P_f(s+t) + P_f(s) P_f(s-t)] / vol_f^2 / (2 Lt F), t = 0..Lt-1,
 * times mod Lt, P_f(s) the oriented sum of the spatial angles around face f at slice s
 * (faces of the params' lattice: triangles, or the dual cells), vol_f its area. */
int qed3_plaquette_correlator(const qed3_geom* g, const qed3_params* p, const double* spatial,
                              double* corr);

typedef struct qed3_hmc_params {
  double gsq;       /* gauge coupling g^2 = 1/beta */
  int compact;      /* gauge action form (see qed3_gauge_action) */
  int npf;          /* pseudofermions (Nf = 2 npf); 0: pure-gauge HMC (hmc.h:212-296) */
  double tau;       /* trajectory length */
  int nsteps;       /* outer (fermion) leapfrog steps */
  int nsteps_inner; /* gauge steps per outer step (integrator.h:58-73) */
  double tol;       /* true-residual tolerance of every pseudofermion solve */
  int solver;       /* QED3_SOLVER_* (QED3_SOLVER_PCG: needs r = 1) */
  int seed;         /* ParallelRngExt master seed (rng.h:91-120) */
} qed3_hmc_params;

typedef struct qed3_hmc_report {
  double dH, h0, h1;
  int accept;
  int cg_solves;
  long long cg_iterations;
  double max_resid;
  double seconds;
} qed3_hmc_report;

/* HMC2 (hmc.h:103-201) with ExplicitLeapfrogML (integrator.h:58-170); the RNG streams
 * persist across trajectories of one handle. */
typedef struct qed3_hmc qed3_hmc;
int qed3_hmc_create(qed3_ctx* ctx, const qed3_hmc_params* p, qed3_hmc** out);
void qed3_hmc_destroy(qed3_hmc* h);
/* One trajectory from the host gauge field (spatial[Lt][E], temporal[Lt][V]), replaced
 * in place by the new field when accepted (or always when no_reject != 0). */
int qed3_hmc_trajectory(qed3_hmc* h, double* spatial, double* temporal, int no_reject,
                        qed3_hmc_report* rep);
/* Trajectory length and leapfrog steps of the following trajectories (no counterpart in the
 * reference, whose HMC2 keeps them fixed): thermalisation from an ordered start converges
 * geometrically only when the length varies from trajectory to trajectory (a fixed length
 * leaves the modes whose period divides it nearly untouched). */
int qed3_hmc_set_steps(qed3_hmc* h, double tau, int nsteps, int nsteps_inner);
/* Overlap fermions in the HMC (hmc.cu:155-186 with IS_OVERLAP, Fermion =
 * Overlap<WilsonDirac>): phi = D_ov^dag xi, eta = (D_ov + D_ov^dag)^{-1} phi, the overlap
 * force, all through ov (created on the HMC's context; it must outlive its use here).
 * refresh != 0 recomputes the spectral bounds and Zolotarev coefficients after every gauge
 * update as Overlap::update does (overlap.h:200-209); 0 keeps ov's, so H is a smooth
 * function of the links.  ov = NULL returns to Wilson fermions.  (qed3_overlap: the
 * overlap operator handle of the section below.) */
typedef struct qed3_overlap qed3_overlap;
int qed3_hmc_set_overlap(qed3_hmc* h, qed3_overlap* ov, int refresh);

/* ---- multi-shift CG and the overlap operator (overlap.h:7-465) --------------------
 * Zolotarev coefficients of the overlap's sign approximation (Zolotarev, overlap.h:7-135):
 * size = n/2 + 1 entries in A, c, cp (index 0 unused), consts = {M, lambda_inv, E, C,
 * Delta}; eps(x) = C x (1 + sum_m A_m / (x^2 - k^2/cp_m)), max |eps - 1| = Delta on
 * [k, 1].  Host only. */
int qed3_zolotarev(int n, double k, int* size, double* A, double* c, double* cp, double* consts);
/* (D^dag D + shifts[i]) x_i = b for nshift <= 32 shifts >= 0 at once (one Krylov space,
 * FP64); reps[i] holds each shift's iterations and FP64 true residual.  Replaces the
 * size-1 independent MatPoly::solve calls of Overlap::mult/adj (overlap.h:309-361). */
int qed3_multishift_cg(qed3_ctx* ctx, int nshift, const double* shifts, void* const* d_x,
                       const void* d_b, const qed3_cg_opts* opts, qed3_cg_report* reps);
/* Extreme eigenvalues of D^dag D (power and inverse iteration to relative change tol);
 * Overlap::compute_lambda_max (overlap.h:217-306) derives lambda_min/max from them. */
int qed3_spectral_bounds(qed3_ctx* ctx, double tol, int maxiter, double* mu_min, double* mu_max);
/* Overlap operator D_ov = 1 + X eps(X^dag X), X = D / lambda_max, k = lambda_min/lambda_max,
 * Zolotarev degree n (<= 65), shifted solves to tol (Overlap, overlap.h:137-361);
 * op = QED3_OP_D applies D_ov, QED3_OP_DDAG its adjoint.  FP64 device fields. */
int qed3_overlap_create(qed3_ctx* ctx, int n, double lambda_min, double lambda_max, double tol,
                        qed3_overlap** out);
void qed3_overlap_destroy(qed3_overlap* ov);
int qed3_overlap_apply(qed3_overlap* ov, int op, void* d_out, const void* d_in);
/* Overlap pseudofermion force for every link, f_l = -2 Re <eta, dD_ov/dtheta_l eta> with
 * lambda_max and the Zolotarev coefficients held fixed; d_fs [Lt][E], d_ft [Lt][V] FP64
 * device arrays.  Replaces Overlap::precalc_grad_deviceAsyncLaunch + the per-link
 * grad_deviceAsyncLaunch loop of GaugeExt::compute (overlap.h:393-465, gauge_ext.h:204-222). */
int qed3_overlap_force(qed3_overlap* ov, const void* d_eta, double* d_fs, double* d_ft);
/* x = (D_ov + D_ov^dag)^{-1} b by FP64 CG from x = 0 (stop: recursive |r| <= tol |b|);
 * rep gets the iterations and the true residual.  Replaces PseudoFermion<Overlap>::update_eta
 * (MatPoly::solve on Overlap::DHD_deviceAsyncLaunch, pseudofermion.h:51, overlap.h:364-377). */
int qed3_overlap_solve(qed3_overlap* ov, void* d_x, const void* d_b, double tol, int maxiter,
                       qed3_cg_report* rep);

/* ---- checkpoint and resume (hmc.cu:423-447 scans ckpoint_lat.k / ckpoint_rng.k) ------
 * Gauge file: raw native-endian FP64, spatial[Lt][E] then temporal[Lt][V], nothing else
 * (GaugeExt::ckpoint/read, gauge_ext.h:225-274).  qed3_gauge_read fails unless the file
 * holds exactly Lt*(E+V) doubles. */
int qed3_gauge_write(const char* path, int Lt, int E, int V, const double* spatial,
                     const double* temporal);
int qed3_gauge_read(const char* path, int Lt, int E, int V, double* spatial, double* temporal);

/* RNG streams of ParallelRngExt (rng.h:91-168) on the host, no device needed: the master
 * mt19937_64 seeded with `seed` reseeds one stream per (t, link) and then per (t, site).
 * The checkpoint text holds four lines per stream (engine, normal, uniform, {1,2}
 * distributions), master first, then links [t][l], then sites [t][x].  Loading restores
 * the engines and resets the distributions, which is what the reference reader ends up
 * doing (rng.h:36-52). */
typedef struct qed3_rng qed3_rng;
int qed3_rng_create(int Lt, int E, int V, int seed, qed3_rng** out);
void qed3_rng_destroy(qed3_rng* r);
int qed3_rng_save(const qed3_rng* r, const char* path);
int qed3_rng_load(qed3_rng* r, const char* path);
/* next N(0,1) of stream `which` (0 master, 1 link (s, i), 2 site (s, i)) */
int qed3_rng_gaussian(qed3_rng* r, int which, int s, int i, double* out);
/* the HMC handle's own streams */
int qed3_hmc_rng_save(const qed3_hmc* h, const char* path);
int qed3_hmc_rng_load(qed3_hmc* h, const char* path);

/* Pick the stencil launch shape (threads per block, time slices per thread) for this
 * lattice by timing every candidate on the device (project-written autotuner); with a
 * gauge field set, the FP32 / binary16 shape is then refined on a few half-storage CG
 * iterations, whose fused kernels prefer other shapes than a bare D apply. */
int qed3_autotune(qed3_ctx* ctx);

/* ---- timing of the operator kernels (HIP events on the context stream) ------------- */
/* Average device time in ms per launch of op (QED3_OP_D / QED3_OP_DDAG) at prec over
 * `reps` launches, after `warmup` untimed ones. */
int qed3_time_apply(qed3_ctx* ctx, int op, int prec, int warmup, int reps, double* ms);
/* Average ms per launch of a plain device copy of `bytes` (read + write): the reference
 * point for HBM bandwidth and for calibrating the FETCH_SIZE/WRITE_SIZE counters. */
int qed3_time_copy(size_t bytes, int reps, double* ms);
/* Launch shape the autotuner chose for op (QED3_OP_D / QED3_OP_DDAG) at prec (binary16
 * fields use the FP32 shapes):
 * out[0] threads per block, out[1] time slices per thread, out[2] XCD-aware order,
 * out[3] LDS-staged tile. */
int qed3_tile_info(const qed3_ctx* ctx, int op, int prec, int* out);
/* Fix the launch shape of op at prec: in[0] threads per block (64/128/256), in[1] time
 * slices per thread, in[2] XCD-aware order (0/1), in[3] LDS-staged tile (0/1). */
int qed3_set_tile(qed3_ctx* ctx, int op, int prec, const int* in);

#ifdef __cplusplus
}
#endif
#endif /* QED3_H */
