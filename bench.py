"""Headline benchmark: CG solves/s on D^dag D to a true FP64 relative residual of 1e-10.

Learned workload (the value): dual (trivalent) refined icosahedron n=128 (327,680 sites)
x Lt=64, antiperiodic in time, r=1, M5=m=0.02, von Mises(kappa=1.4) link angles, Gaussian
RHS.  Four solvers are timed on the same seeded RHS (one complete solve per step): the
flexible CG with the learned preconditioner M = N N^dagger (FP32 vectors, FP64 reliable
updates; the net trained first, outside the timed region, on a small lattice of the same
family, dual n=16 x Lt=16, its weights applied unchanged at full size), the
unpreconditioned mixed CG (FP32 + FP64 reliable updates), the same CG with its inner
fields stored in binary16 (FP32 arithmetic, FP64 reliable updates) and the plain FP64 CG.
`value` is the solves/s of the fastest of them, named in config.solver; every method's
figures and its ratio to plain FP64 CG are in config.methods.  `roofline` is that of the
stencil kernel of the fastest method (one D apply at its storage precision).

Second workload: simplicial n=128 (163,842 sites) x Lt=64, m=0.05: the three CGs.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
# Bytes moved per D launch (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors read off a
# 1 GiB copy kernel) from two rocprofv3 --pmc runs of `bench.py --kernels-only --lattice 1
# --m 0.02 --tile <block,ts,remap,lds>` per launch shape the autotuner picks; summary in
# profiles/stencil_D_dual_n128_Lt64_hbm_traffic.txt.  Keyed by (precision or CG kernel,
# lattice, n, Lt, tile); the CG kernels ("K1-half" ...) from the half-storage solve the same
# command runs at the pinned shape.
MEASURED_TRAFFIC = {("f16", 1, 128, 64, 256, 64, 1, 1): 4.9646e8,
                    ("f32", 1, 128, 64, 256, 64, 1, 1): 1.2992e9,
                    ("K1-half", 1, 128, 64, 256, 64, 1, 1): 8.5079e8,
                    ("K2-half", 1, 128, 64, 256, 64, 1, 1): 6.3056e8,
                    ("K2x-half", 1, 128, 64, 256, 64, 1, 1): 1.9851e9}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--n", type=int, default=128)
    p.add_argument("--Lt", type=int, default=64)
    p.add_argument("--lattice", type=int, default=0)
    p.add_argument("--m", type=float, default=0.05)
    p.add_argument("--beta", type=float, default=1.4)
    p.add_argument("--tol", type=float, default=1e-10)
    p.add_argument("--cpu-lt", type=int, default=8)
    p.add_argument("--cpu-iters", type=int, default=600)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-fp64", action="store_true")
    p.add_argument("--no-learned", action="store_true")
    p.add_argument("--no-hmc", action="store_true")
    p.add_argument("--no-overlap", action="store_true")
    p.add_argument("--hmc-traj", type=int, default=3)
    p.add_argument("--hmc-therm", type=int, default=6)
    p.add_argument("--hmc-inner", type=int, default=50)
    p.add_argument("--hmc-nsteps", type=int, default=20)
    p.add_argument("--batched", action="store_true",
                   help="also time the 64-system batched CG (uneven sizes) against one-by-one solves")
    p.add_argument("--batch-nsys", type=int, default=64)
    p.add_argument("--learned-steps", type=int, default=2)
    p.add_argument("--train-n", type=int, default=16)
    p.add_argument("--train-Lt", type=int, default=16)
    p.add_argument("--train-epochs", type=int, default=40)
    p.add_argument("--tile", type=str, default=None,
                   help="block,ts,remap,lds for the FP32 D stencil instead of the autotuned one")
    p.add_argument("--kernels-only", action="store_true",
                   help="only launch the D stencils and the calibration copy (for profilers)")
    return p.parse_args()


def copy_bandwidth(q, nbytes=1 << 30, reps=20):
    """GB/s of a plain device copy (read + write of nbytes) and its ms per launch."""
    import ctypes as C
    ms = C.c_double()
    q.check(q.load().qed3_time_copy(nbytes, reps, C.byref(ms)))
    return 2 * nbytes / (ms.value * 1e-3) / 1e9, ms.value


def tile_info(q, D, op, prec):
    import ctypes as C
    out = (C.c_int * 4)()
    q.check(q.load().qed3_tile_info(D._h, op, prec, out))
    return {"block": out[0], "ts": out[1], "xcd_remap": out[2], "lds": out[3]}


def mean_ell(g, lattice):
    if lattice == 0:
        x, y = g.sites[g.links[:, 0]], g.sites[g.links[:, 1]]
    else:
        x, y = g.dual_sites[g.dual_links[:, 0]], g.dual_sites[g.dual_links[:, 1]]
    return float(np.mean(np.arccos(np.clip(np.einsum("ij,ij->i", x, y), -1, 1))))


def algorithmic_bytes_D(V, E, Lt, real_bytes):
    """B_D = 2 B_vec + B_link: one spinor field read and written once, one angle per link."""
    b_vec = 2 * V * Lt * 2 * real_bytes
    b_link = (E + V) * Lt * real_bytes
    return 2 * b_vec + b_link


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "qed3_amd runs on one GPU; --gpus must be 1"}))
        sys.exit(1)
    import qed3_amd as q

    t_setup = time.time()
    g = q.Geometry(a.n)
    at = 0.5 * mean_ell(g, a.lattice)
    D = q.WilsonDirac(a.lattice, a.n, a.Lt, r=1.0, M5=a.m, at=at)
    sp, tp = q.random_gauge(a.lattice, a.n, a.Lt, seed=1234, kappa=a.beta)
    D.update(sp, tp)
    D.autotune()
    print(f"# setup {time.time() - t_setup:.1f}s V={D.V} E={D.E} Lt={a.Lt} at={at:.5f}",
          file=sys.stderr, flush=True)

    if a.kernels_only:
        if a.tile:
            b_, t_, r_, l_ = (int(v) for v in a.tile.split(","))
            D.set_tile(b_, t_, r_, l_, ops=(q.OP_D, q.OP_DDAG), precs=(q.F32,))
        for prec, nm in ((q.F16, "f16"), (q.F32, "f32"), (q.F64, "f64")):
            for op in (q.OP_D, q.OP_DDAG):
                ms = D.time_apply(op, prec, warmup=2, reps=10)
                rb = {q.F16: 2, q.F32: 4, q.F64: 8}[prec]
                gbs = algorithmic_bytes_D(D.V, D.E, a.Lt, rb) / (ms * 1e-3) / 1e9
                print(f"# D op={op} {nm}: {ms:.4f} ms  {gbs:.0f} GB/s algorithmic  "
                      f"tile={tile_info(q, D, op, prec)}", file=sys.stderr, flush=True)
        # one half-storage CG solve at the same shape: its K1 / K2 kernels for the counters
        d_b = q.DeviceArray.from_numpy(q.random_source(a.lattice, a.n, a.Lt, seed=1000))
        d_x = q.DeviceArray(D.N * 16)
        rep, ms = D.solve_kernel_times(d_x, d_b, tol=a.tol, solver=q.CG_HALF)
        print(f"# half CG: {rep.iterations} its, K1 {ms[0]:.4f} ms, K2 {ms[1]:.4f} ms, "
              f"K2 with x update {ms[2]:.4f} ms", file=sys.stderr, flush=True)
        gbs, ms = copy_bandwidth(q, 1 << 30, reps=10)
        print(f"# copy 1 GiB: {ms:.4f} ms  {gbs:.0f} GB/s", file=sys.stderr, flush=True)
        return

    nbytes = D.N * 16
    d_x = q.DeviceArray(nbytes)
    rhs = [q.DeviceArray.from_numpy(q.random_source(a.lattice, a.n, a.Lt, seed=1000 + k))
           for k in range(a.warmup + a.steps)]

    def run(solver, k):
        return D.solve_device(d_x, rhs[k], tol=a.tol, solver=solver)

    results = {}
    for name, solver in (("half", q.CG_HALF), ("mixed", q.CG_MIXED), ("fp64", q.CG_F64)):
        if name == "fp64" and a.no_fp64:
            continue
        for k in range(a.warmup):
            run(solver, k)
        q.load().qed3_device_synchronize()
        t0 = time.time()
        reps = []
        for k in range(a.warmup, a.warmup + a.steps):
            r = run(solver, k)
            reps.append(r)
            print(f"# {name} solve {k}: iters={r.iterations} ru={r.reliable_updates} "
                  f"resid={r.true_resid:.2e} {r.seconds:.3f}s", file=sys.stderr, flush=True)
        q.load().qed3_device_synchronize()
        dt = time.time() - t0
        results[name] = dict(solves_per_s=a.steps / dt, ms_per_solve=1e3 * dt / a.steps,
                             iters=float(np.mean([r.iterations for r in reps])),
                             reliable_updates=float(np.mean([r.reliable_updates for r in reps])),
                             max_true_resid=float(max(r.true_resid for r in reps)))

    # roofline of the dominant kernel: the FP32 stencil (D) inside the mixed solver
    ms32 = D.time_apply(q.OP_D, q.F32, warmup=5, reps=50)
    tile32 = tile_info(q, D, q.OP_D, q.F32)
    D_V, D_E = D.V, D.E
    ms64 = D.time_apply(q.OP_D, q.F64, warmup=5, reps=50)
    b32 = algorithmic_bytes_D(D.V, D.E, a.Lt, 4)
    b64 = algorithmic_bytes_D(D.V, D.E, a.Lt, 8)
    gbs32 = b32 / (ms32 * 1e-3) / 1e9
    gbs64 = b64 / (ms64 * 1e-3) / 1e9
    copy_gbs, _ = copy_bandwidth(q)

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(a, g, at, results.get("fp64", results["mixed"])["iters"], sp, tp)
    del D, d_x, rhs
    learned = learned_leg(a, q, g, with_net=not a.no_learned)
    hmc = None if a.no_hmc else hmc_leg(a, q)
    overlap = None if a.no_overlap else overlap_leg(q)
    batched = batched_leg(a, q) if a.batched else None

    simp = {"workload": f"{'simplicial' if a.lattice == 0 else 'dual'} n={a.n} Lt={a.Lt} "
                        f"m={a.m} tol={a.tol} D^dag D CG", "methods": results,
            "speedup_mixed_vs_fp64": (round(results["mixed"]["solves_per_s"]
                                            / results["fp64"]["solves_per_s"], 3)
                                      if "fp64" in results else None),
            "speedup_half_vs_fp64": (round(results["half"]["solves_per_s"]
                                           / results["fp64"]["solves_per_s"], 3)
                                     if "fp64" in results else None)}
    solver_names = {"learned": "flexible CG, M = N N^dag (4-layer net), FP32 + FP64 reliable updates",
                    "half": "CG, binary16-stored r/p/w (FP32 arithmetic, FP32 x) + FP64 reliable updates",
                    "mixed": "CG, FP32 + FP64 reliable updates", "fp64": "CG, FP64"}
    if True:
        meth = learned["methods"]
        winner = max(meth, key=lambda k: meth[k]["solves_per_s"])
        best = meth[winner]
        for k in meth:
            if "fp64" in meth:
                meth[k]["vs_fp64"] = round(meth[k]["solves_per_s"] / meth["fp64"]["solves_per_s"], 3)
        steps, ms = a.learned_steps, best["ms_per_solve"]
        config = {"workload": learned["workload"], "n_refine": a.n, "Lt": a.Lt,
                  "sites": learned["sites"], "links": learned["links"],
                  "solver": solver_names[winner],
                  "methods": learned["methods"], "training": learned["training"],
                  "speedup_learned_vs_fp64": learned["speedup_vs_fp64"],
                  "speedup_mixed_vs_fp64": learned["speedup_mixed_vs_fp64"],
                  "speedup_half_vs_fp64": learned["speedup_half_vs_fp64"],
                  "net_apply": learned["net_apply"], "solution_agreement": learned["agreement"],
                  "second_workload": simp, "hmc": hmc, "overlap": overlap, "batched": batched}
        # the dominant kernel of the fastest CG: its K1 (the most time per solve)
        ck = learned["cg_kernels"]
        src = winner if winner in ck else "mixed"
        k1 = ck[src]["K1"]
        roof = {"bound": "hbm", "achieved": k1["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": k1["frac"], "traffic": k1["traffic"],
                "kernel": k1["kernel"] + f" (CG K1 of the {src} solver: p = r + beta p_old, w = D p, |w|^2)",
                "bytes_per_launch": k1["bytes_per_launch"], "ms_per_launch": k1["ms_per_launch"],
                "tile": ck[src]["tile"], "cg_kernels": ck, "bare_D": learned["roofline"]}
        dtype = ("f32 inner (bf16 net activations) / f64 reliable updates" if winner == "learned"
                 else "f16 storage, f32 arithmetic / f64 reliable updates" if winner == "half"
                 else "f32 inner / f64 reliable updates" if winner == "mixed" else "f64")
    line = {
        "metric": "CG solves/s to 1e-10 true resid, n=128 Lt=64: learned+FP32 vs plain FP64 CG",
        "value": round(best["solves_per_s"], 4),
        "unit": "solves/s",
        "n_gpus": 1, "steps": steps, "warmup": a.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic: von Mises(kappa=%.2f) links, Gaussian RHS, seeded" % a.beta,
        "config": config,
        "roofline": dict(roof, copy_kernel_GBs=round(copy_gbs, 1),
                         simplicial={"kernel": ("k_stencil_lds" if tile32["lds"] else "k_stencil")
                                     + "<float,6,false,EpiStore> (D, FP32, simplicial n=128 Lt=64)",
                                     "achieved": round(gbs32, 1),
                                     "frac": round(gbs32 / HBM_PEAK_GBS, 4),
                                     "bytes_per_launch": b32, "ms_per_launch": round(ms32, 5),
                                     "tile": tile32,
                                     "fp64_D": {"achieved": round(gbs64, 1),
                                                "frac": round(gbs64 / HBM_PEAK_GBS, 4),
                                                "bytes_per_launch": b64,
                                                "ms_per_launch": round(ms64, 5)}}),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def timed_solves(q, a, name, run, rhs, warmup, steps):
    for k in range(warmup):
        run(rhs[k])
    q.load().qed3_device_synchronize()
    t0 = time.time()
    reps = []
    for k in range(warmup, warmup + steps):
        r = run(rhs[k])
        reps.append(r)
        print(f"# {name} solve {k}: iters={r.iterations} ru={r.reliable_updates} "
              f"resid={r.true_resid:.2e} {r.seconds:.3f}s", file=sys.stderr, flush=True)
    q.load().qed3_device_synchronize()
    dt = time.time() - t0
    return dict(solves_per_s=steps / dt, ms_per_solve=1e3 * dt / steps,
                iters=float(np.mean([r.iterations for r in reps])),
                reliable_updates=float(np.mean([r.reliable_updates for r in reps])),
                max_true_resid=float(max(r.true_resid for r in reps)))


def learned_leg(a, q, g, with_net=True):
    """Dual n x Lt at m=0.02: train on dual train_n x train_Lt, then time FCG with the net,
    mixed CG and FP64 CG on the same seeded RHS (with_net=False: the two CGs only), and the
    roofline of this workload's FP32 stencil."""
    from qed3_amd import train
    m, lattice = 0.02, 1
    if not with_net:
        return plain_leg(a, q, g, m, lattice)
    t0 = time.time()
    gt = q.Geometry(a.train_n)
    Dt = q.WilsonDirac(lattice, a.train_n, a.train_Lt, r=1.0, M5=m, at=0.5 * mean_ell(gt, lattice))
    data = train.make_pairs(Dt, lattice, a.train_n, a.train_Lt, cfg_seeds=[11, 12, 13],
                            betas=[1.0, 1.4, 2.0])
    t_pairs = time.time() - t0
    net_t = q.Net(Dt, layers=4, seed=5)
    hist = train.train(net_t, Dt, data, epochs=a.train_epochs, batch=4, lr=3e-3)
    t_train = time.time() - t0
    print(f"# trained on dual n={a.train_n} Lt={a.train_Lt}: loss {hist[0]:.3f} -> {hist[-1]:.4f} "
          f"in {t_train:.1f}s", file=sys.stderr, flush=True)
    # on the training lattice itself: an unseen configuration and RHS
    sp_s, tp_s = q.random_gauge(lattice, a.train_n, a.train_Lt, seed=99, kappa=a.beta)
    Dt.update(sp_s, tp_s)
    b_s = q.DeviceArray.from_numpy(q.random_source(lattice, a.train_n, a.train_Lt, seed=77))
    x_s = q.DeviceArray(Dt.N * 16)
    small_cg = Dt.solve_device(x_s, b_s, tol=a.tol, solver=q.CG_F64)
    small_fcg, _, small_fb = net_t.fcg(x_s, b_s, tol=a.tol)
    small = {"lattice": f"dual n={a.train_n} Lt={a.train_Lt} (the training lattice, unseen links and RHS)",
             "plain_cg_iterations": small_cg.iterations, "learned_fcg_iterations": small_fcg.iterations,
             "fallback": small_fb}
    print(f"# on the training lattice: plain CG {small_cg.iterations} its, learned FCG "
          f"{small_fcg.iterations} its", file=sys.stderr, flush=True)
    at = 0.5 * mean_ell(g, lattice)
    D = dual_context(a, q, g, m, lattice)
    net = q.Net(D, layers=4, seed=5)
    net.params = net_t.params
    del net_t, Dt, data
    nbytes = D.N * 16
    d_x = q.DeviceArray(nbytes)
    rhs = [q.DeviceArray.from_numpy(q.random_source(lattice, a.n, a.Lt, seed=2000 + k))
           for k in range(a.warmup + a.learned_steps)]
    methods = {}
    stats = []

    def run_fcg(b):
        rep, applies, fb = net.fcg(d_x, b, tol=a.tol)
        stats.append((applies, fb))
        return rep

    methods["learned"] = timed_solves(q, a, "learned", run_fcg, rhs, a.warmup, a.learned_steps)
    x_learned = d_x.to_numpy(np.complex128)
    methods["learned"]["fallbacks"] = int(sum(fb for _, fb in stats[a.warmup:]))
    methods["half"] = timed_solves(q, a, "half", lambda b: D.solve_device(d_x, b, tol=a.tol,
                                   solver=q.CG_HALF), rhs, a.warmup, a.learned_steps)
    x_half = d_x.to_numpy(np.complex128)
    methods["mixed"] = timed_solves(q, a, "mixed", lambda b: D.solve_device(d_x, b, tol=a.tol,
                                    solver=q.CG_MIXED), rhs, a.warmup, a.learned_steps)
    x_mixed = d_x.to_numpy(np.complex128)
    agreement = None
    if not a.no_fp64:
        methods["fp64"] = timed_solves(q, a, "fp64", lambda b: D.solve_device(d_x, b, tol=a.tol,
                                       solver=q.CG_F64), rhs, a.warmup, a.learned_steps)
        x64 = d_x.to_numpy(np.complex128)
        nrm = np.linalg.norm(x64)
        agreement = {"rhs_seed": 2000 + a.warmup + a.learned_steps - 1,
                     "mixed_vs_fp64": float(np.linalg.norm(x_mixed - x64) / nrm),
                     "half_vs_fp64": float(np.linalg.norm(x_half - x64) / nrm),
                     "learned_vs_fp64": float(np.linalg.norm(x_learned - x64) / nrm)}
    # one net application N (and N^dag) at full size, device-resident fields
    d_y = q.DeviceArray(D.N * 8)
    d_z = q.DeviceArray(D.N * 8)
    for fn in (net.apply_device, net.apply_adj_device):
        fn(d_z, d_y, 1)
    q.load().qed3_device_synchronize()
    reps = 10
    t1 = time.time()
    for _ in range(reps):
        net.apply_device(d_z, d_y, 1)
    q.load().qed3_device_synchronize()
    ms_net = 1e3 * (time.time() - t1) / reps
    t1 = time.time()
    for _ in range(reps):
        net.apply_adj_device(d_z, d_y, 1)
    q.load().qed3_device_synchronize()
    ms_adj = 1e3 * (time.time() - t1) / reps
    pts = D.V * a.Lt
    # algorithmic bytes of one N: input spinor read (8 B complex x 2) + 3 layers' BF16
    # activations written and read (16 channels x 2 spins x 4 B) + output spinor written,
    # plus the link phases once per layer
    b_net = pts * (16 + 3 * 2 * 128 + 16) + 4 * (D.E + D.V) * a.Lt * 8
    fp64 = methods.get("fp64")
    return {"workload": f"dual n={a.n} Lt={a.Lt} m={m} tol={a.tol} D^dag D CG",
            "agreement": agreement,
            "roofline": stencil_roofline(a, q, D),
            "cg_kernels": cg_kernel_rooflines(a, q, D, d_x, rhs[-1]),
            "sites": D.V, "links": D.E, "methods": methods,
            "training": {"lattice": f"dual n={a.train_n} Lt={a.train_Lt}", "pairs": 3 * 40,
                         "epochs": a.train_epochs, "loss_first": round(hist[0], 4),
                         "on_training_lattice": small,
                         "loss_last": round(hist[-1], 5), "seconds": round(t_train, 1),
                         "pair_generation_s": round(t_pairs, 1)},
            "speedup_vs_fp64": (round(methods["learned"]["solves_per_s"] / fp64["solves_per_s"], 3)
                                if fp64 else None),
            "speedup_mixed_vs_fp64": (round(methods["mixed"]["solves_per_s"] / fp64["solves_per_s"], 3)
                                      if fp64 else None),
            "speedup_half_vs_fp64": (round(methods["half"]["solves_per_s"] / fp64["solves_per_s"], 3)
                                     if fp64 else None),
            "net_apply": {"ms": round(ms_net, 4), "ms_adjoint": round(ms_adj, 4),
                          "algorithmic_GB": round(b_net / 1e9, 3),
                          "achieved_GBs": round(b_net / (ms_net * 1e-3) / 1e9, 1)}}


def dual_context(a, q, g, m, lattice):
    at = 0.5 * mean_ell(g, lattice)
    D = q.WilsonDirac(lattice, a.n, a.Lt, r=1.0, M5=m, at=at)
    sp, tp = q.random_gauge(lattice, a.n, a.Lt, seed=4321, kappa=a.beta)
    D.update(sp, tp)
    D.autotune()
    return D


def cg_kernel_bytes(V, E, Lt, solver):
    """Algorithmic bytes per launch of the CG's two stencil kernels (fields read and written
    once, one real per link at the inner precision): K1 reads r, p_old and writes p, w;
    K2 reads w, r (and x, p) and writes r (and x).  Half storage: 4 B per complex,
    2 B per link, FP32 x read and written by K2 once per four iterations with the four
    directions of its group; its fields live in the temporal gauge, where the temporal
    links are 1 except one wrap phase per site, so it reads the E Lt spatial links and
    V wrap phases only."""
    pts = V * Lt
    if solver == "half":
        bv, bl, bx = pts * 2 * 4, (E * Lt + V) * 2, pts * 2 * 8
        return {"K1": 4 * bv + bl, "K2": 3 * bv + bl, "K2x": 7 * bv + 2 * bx + bl}
    s_r = 4 if solver == "mixed" else 8
    bv, bl = pts * 2 * 2 * s_r, (E + V) * Lt * s_r
    return {"K1": 4 * bv + bl, "K2": 6 * bv + bl, "K2x": 6 * bv + bl}


CG_KERNEL_NAMES = {
    "half": {"K1": "k_stencil_lds<float,NS,false,EpiK1PH,SrcPUpdateH,phase16_t>",
             "K2": "k_stencil_lds<float,NS,true,EpiK2H<0..2>,SrcHalf,phase16_t>",
             "K2x": "k_stencil_lds<float,NS,true,EpiK2H<3>,SrcHalf,phase16_t>"},
    "mixed": {"K1": "k_stencil_lds<float,NS,false,EpiK1P<float>,SrcPUpdate<float>>",
              "K2": "k_stencil_lds<float,NS,true,EpiK2<float>,SrcPlain<float>>"},
    "fp64": {"K1": "k_stencil_lds<double,NS,false,EpiK1P<double>,SrcPUpdate<double>>",
             "K2": "k_stencil_lds<double,NS,true,EpiK2<double>,SrcPlain<double>>"},
}


def cg_kernel_rooflines(a, q, D, d_x, d_b):
    """The CG's own stencil kernels (fused BLAS-1 epilogues), timed with HIP events around
    every launch of one solve per solver (direct launches, outside the timed solves),
    against their algorithmic bytes."""
    out = {}
    ns = "6" if D.lattice == 0 else "3"
    for name, solver in (("half", q.CG_HALF), ("mixed", q.CG_MIXED), ("fp64", q.CG_F64)):
        if name == "fp64" and a.no_fp64:
            continue
        rep, ms = D.solve_kernel_times(d_x, d_b, tol=a.tol, solver=solver)
        nbytes = cg_kernel_bytes(D.V, D.E, a.Lt, name)
        prec = q.F64 if name == "fp64" else q.F32
        tile = tile_info(q, D, q.OP_D, prec)
        ent = {}
        for k, t in zip(("K1", "K2", "K2x"), ms):
            if k not in CG_KERNEL_NAMES[name]:
                continue
            gbs = nbytes[k] / (t * 1e-3) / 1e9
            ent[k] = {"kernel": CG_KERNEL_NAMES[name][k].replace("NS", ns), "ms_per_launch": round(t, 5),
                      "bytes_per_launch": nbytes[k], "achieved": round(gbs, 1),
                      "frac": round(gbs / HBM_PEAK_GBS, 4),
                      "traffic": MEASURED_TRAFFIC.get((f"{k}-{name}", D.lattice, a.n, a.Lt, tile["block"],
                                                       tile["ts"], tile["xcd_remap"], tile["lds"]))}
        ent["iterations"] = rep.iterations
        ent["tile"] = tile
        out[name] = ent
        print(f"# {name} CG kernels: " + "  ".join(f"{k} {v['ms_per_launch']:.4f} ms {v['achieved']:.0f} GB/s"
                                                  for k, v in ent.items() if k.startswith("K")),
              file=sys.stderr, flush=True)
    return out


def stencil_roofline(a, q, D):
    """The dominant kernel of the CG: the stencil (one D apply) at the inner solver's
    storage precision, timed with HIP events on the context's stream, against
    B_D = 2 B_vec + B_link (one real per link at that precision).  Returns both the
    binary16-storage and the FP32 kernel, keyed "f16" / "f32"."""
    out = {}
    for nm, prec, rb, kern in (("f16", q.F16, 2, "EpiStoreH,SrcHalf,phase16_t> (D, binary16 storage)"),
                               ("f32", q.F32, 4, "EpiStore,SrcPlain<float>> (D, FP32)")):
        ms = D.time_apply(q.OP_D, prec, warmup=5, reps=50)
        b = algorithmic_bytes_D(D.V, D.E, a.Lt, rb)
        gbs = b / (ms * 1e-3) / 1e9
        tile = tile_info(q, D, q.OP_D, prec)
        out[nm] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": round(gbs / HBM_PEAK_GBS, 4),
                   "traffic": MEASURED_TRAFFIC.get((nm, D.lattice, a.n, a.Lt, tile["block"], tile["ts"],
                                                    tile["xcd_remap"], tile["lds"])),
                   "kernel": ("k_stencil_lds" if tile["lds"] else "k_stencil")
                   + f"<float,{6 if D.lattice == 0 else 3},false," + kern,
                   "bytes_per_launch": b, "ms_per_launch": round(ms, 5), "tile": tile}
    return out


def plain_leg(a, q, g, m, lattice):
    D = dual_context(a, q, g, m, lattice)
    d_x = q.DeviceArray(D.N * 16)
    rhs = [q.DeviceArray.from_numpy(q.random_source(lattice, a.n, a.Lt, seed=2000 + k))
           for k in range(a.warmup + a.learned_steps)]
    methods = {"half": timed_solves(q, a, "half", lambda b: D.solve_device(
        d_x, b, tol=a.tol, solver=q.CG_HALF), rhs, a.warmup, a.learned_steps)}
    x_half = d_x.to_numpy(np.complex128)
    methods["mixed"] = timed_solves(q, a, "mixed", lambda b: D.solve_device(
        d_x, b, tol=a.tol, solver=q.CG_MIXED), rhs, a.warmup, a.learned_steps)
    x_mixed = d_x.to_numpy(np.complex128)
    agreement = None
    if not a.no_fp64:
        methods["fp64"] = timed_solves(q, a, "fp64", lambda b: D.solve_device(
            d_x, b, tol=a.tol, solver=q.CG_F64), rhs, a.warmup, a.learned_steps)
        x64 = d_x.to_numpy(np.complex128)
        agreement = {"rhs_seed": 2000 + a.warmup + a.learned_steps - 1,
                     "mixed_vs_fp64": float(np.linalg.norm(x_mixed - x64) / np.linalg.norm(x64)),
                     "half_vs_fp64": float(np.linalg.norm(x_half - x64) / np.linalg.norm(x64))}
    fp64 = methods.get("fp64")
    return {"workload": f"dual n={a.n} Lt={a.Lt} m={m} tol={a.tol} D^dag D CG", "agreement": agreement,
            "roofline": stencil_roofline(a, q, D), "cg_kernels": cg_kernel_rooflines(a, q, D, d_x, rhs[-1]),
            "sites": D.V, "links": D.E,
            "methods": methods, "training": None, "speedup_vs_fp64": None,
            "speedup_mixed_vs_fp64": (round(methods["mixed"]["solves_per_s"] / fp64["solves_per_s"], 3)
                                      if fp64 else None),
            "speedup_half_vs_fp64": (round(methods["half"]["solves_per_s"] / fp64["solves_per_s"], 3)
                                     if fp64 else None), "net_apply": None}


def overlap_leg(q, n=16, Lt=16, M5=-0.5, degree=21, tol=1e-10):
    """Overlap fermion pieces (overlap.h:137-465) on simplicial n=16 x Lt=16 with the
    Wilson kernel at M5 = -0.5: spectral bounds of D^dag D, one D_ov apply, the
    pseudofermion force of every link, the CG on D_ov + D_ov^dag.  The force is checked
    at this size against a central difference of Re <eta, D_ov eta> along a random link
    direction (lambda_max and the Zolotarev coefficients fixed, eps = 1e-5)."""
    g = q.Geometry(n)
    at = 0.5 * mean_ell(g, 0)
    D = q.WilsonDirac(0, n, Lt, r=1.0, M5=M5, at=at)
    sp, tp = q.random_gauge(0, n, Lt, seed=77, kappa=1.4)
    D.update(sp, tp)
    t0 = time.time()
    mu_min, mu_max = q.spectral_bounds(D, tol=1e-8, maxiter=2000)
    t_bounds = time.time() - t0
    ov = q.Overlap(D, n=degree, lambda_min=np.sqrt((1 - 1e-6) * mu_min),
                   lambda_max=np.sqrt((1 + 1e-6) * mu_max), tol=tol)
    z = q.zolotarev(degree, ov.lambda_min / ov.lambda_max)
    eta = q.random_source(0, n, Lt, seed=3)
    ov.apply(eta)  # warm-up
    t0 = time.time()
    w = ov.apply(eta)
    t_apply = time.time() - t0
    t0 = time.time()
    fs, ft = ov.force(eta)
    t_force = time.time() - t0
    rng = np.random.default_rng(5)
    vs, vt = rng.standard_normal(sp.shape), rng.standard_normal(tp.shape)
    eps = 1e-5

    def g_of(sgn):
        D.update(sp + sgn * eps * vs, tp + sgn * eps * vt)
        return float(np.real(np.vdot(eta, ov.apply(eta))))

    fd = -2.0 * (g_of(1) - g_of(-1)) / (2 * eps)
    D.update(sp, tp)
    dirf = float(np.sum(fs * vs) + np.sum(ft * vt))
    phi = q.random_source(0, n, Lt, seed=4)
    t0 = time.time()
    _, rep = ov.solve(phi, tol=tol)
    t_solve = time.time() - t0
    print(f"# overlap n={n} Lt={Lt}: k={ov.lambda_min / ov.lambda_max:.3e} apply {1e3 * t_apply:.1f} ms "
          f"force {1e3 * t_force:.1f} ms solve {t_solve:.2f}s ({rep.iterations} its)",
          file=sys.stderr, flush=True)
    return {"workload": f"simplicial n={n} Lt={Lt} M5={M5} kappa=1.4, Zolotarev degree {degree}, "
                        f"shifted solves to {tol}",
            "points": D.V * Lt, "mu_min": mu_min, "mu_max": mu_max,
            "k": ov.lambda_min / ov.lambda_max, "zolotarev_delta": z["Delta"],
            "bounds_s": round(t_bounds, 3), "apply_ms": round(1e3 * t_apply, 2),
            "unitarity_dev": float(abs(np.linalg.norm(w - eta) / np.linalg.norm(eta) - 1)),
            "force_ms": round(1e3 * t_force, 2),
            "force_vs_central_difference": abs(dirf - fd) / abs(fd),
            "solve_s": round(t_solve, 3), "solve_iterations": rep.iterations,
            "solve_true_resid": rep.true_resid}


def hmc_leg(a, q):
    """Full HMC (simplicial n=64, Lt=32, m=0.1, compact U(1) at beta=1.4, tau=1, 20 outer
    (fermion) steps of hmc_inner gauge steps each, Nf=2 via one pseudofermion, half-storage
    CG):
    hmc_therm forced-accept trajectories from the von Mises start, then hmc_traj timed
    trajectories with the accept step (seconds per trajectory, dH, acceptance, CG work);
    the pseudofermion force at the same size against the FP64 CPU path (the oracle's local
    route) on the same eta, and the force launch against its algorithmic bytes.  The gauge
    coupling per face, beta_s = a_t/(vol_f g^2), is ~220 at n=64 with a_t = <l>/2, which
    sets the small gauge step."""
    import oracle as O
    n, Lt, m, lattice = 64, 32, 0.1, 0
    g = q.Geometry(n)
    at = 0.5 * mean_ell(g, lattice)
    D = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=m, at=at)
    sp, tp = q.random_gauge(lattice, n, Lt, seed=55, kappa=1.4)
    sp, tp = np.ascontiguousarray(sp), np.ascontiguousarray(tp)
    D.update(sp, tp)
    D.autotune()
    H = q.HMC(D, 1 / 1.4, compact=True, npf=1, tau=1.0, nsteps=a.hmc_nsteps, nsteps_inner=a.hmc_inner,
              tol=1e-10, solver=q.CG_HALF, seed=7)
    for k in range(a.hmc_therm):
        r = H.trajectory(sp, tp, no_reject=True)
        print(f"# hmc thermalisation {k}: dH={r['dH']:.3e}", file=sys.stderr, flush=True)
    trajs = []
    for k in range(a.hmc_traj):
        t0 = time.time()
        r = H.trajectory(sp, tp)
        r["wall_s"] = time.time() - t0
        trajs.append(r)
        print(f"# hmc traj {k}: dH={r['dH']:.3e} accept={r['accept']} cg={r['cg_solves']}/"
              f"{r['cg_iterations']} {r['wall_s']:.2f}s", file=sys.stderr, flush=True)
    # force at this size: GPU vs the FP64 CPU path on the same eta
    D.update(sp, tp)
    phi = q.random_source(lattice, n, Lt, seed=3)
    eta, _ = D.solve(phi, tol=1e-12, solver=q.CG_F64)
    d_eta = q.DeviceArray.from_numpy(eta)
    d_fs, d_ft = q.DeviceArray(8 * Lt * D.E), q.DeviceArray(8 * Lt * D.V)
    for _ in range(2):
        q.check(q.load().qed3_pf_force(D._h, d_eta.ptr, d_fs.ptr, d_ft.ptr))
    q.load().qed3_device_synchronize()
    t0 = time.time()
    reps = 20
    for _ in range(reps):
        q.check(q.load().qed3_pf_force(D._h, d_eta.ptr, d_fs.ptr, d_ft.ptr))
    q.load().qed3_device_synchronize()
    ms_force = 1e3 * (time.time() - t0) / reps
    fs = d_fs.to_numpy(np.float64).reshape(Lt, D.E)
    ft = d_ft.to_numpy(np.float64).reshape(Lt, D.V)
    t0 = time.time()
    L = O.Lattice(g, lattice)
    Do = O.Dirac(L, Lt, r=1.0, M5=m, at=at)
    Do.update(sp, tp)
    t_setup = time.time() - t0
    t0 = time.time()
    fs_o, ft_o = Do.force(eta)
    t_cpu = time.time() - t0
    scale = max(np.abs(fs_o).max(), np.abs(ft_o).max())
    err = max(np.abs(fs - fs_o).max(), np.abs(ft - ft_o).max()) / scale
    vec = 2 * D.V * Lt * 16
    link = (D.E + D.V) * Lt * 16
    # one FP64 D apply (D eta), then the link kernel: eta and D eta in, phases, one double out
    b_force = (2 * vec + link) + (2 * vec + link + (D.E + D.V) * Lt * 8)
    return {"workload": f"simplicial n={n} Lt={Lt} m={m} compact beta=1.4 tau=1 "
                        f"nsteps={a.hmc_nsteps}x{a.hmc_inner} npf=1, after {a.hmc_therm} forced-accept "
                        "trajectories",
            "acceptance": float(np.mean([t["accept"] for t in trajs])),
            "seconds_per_trajectory": round(float(np.mean([t["wall_s"] for t in trajs])), 3),
            "trajectories": [{k: t[k] for k in ("dH", "accept", "cg_solves", "cg_iterations",
                                                "max_resid")} for t in trajs],
            "force": {"ms_gpu": round(ms_force, 4), "algorithmic_GB": round(b_force / 1e9, 4),
                      "achieved_GBs": round(b_force / (ms_force * 1e-3) / 1e9, 1),
                      "max_rel_diff_vs_cpu_fp64": float(err), "cpu_fp64_s": round(t_cpu, 3),
                      "cpu_setup_s": round(t_setup, 1)}}


def batched_leg(a, q):
    """Batched half-storage CG over nsys systems of uneven shapes: refinement n in
    {32,64,96,128}, Lt in {16,32,64}, both lattices, m ~ U[0.02, 0.2] (seeded), von
    Mises(1.4) links, one Gaussian RHS each; one batched call against the same systems
    solved one by one with the single-system half-storage CG.  The launch shapes are
    autotuned once per (lattice, n, Lt) and shared by the systems of that shape."""
    rng = np.random.default_rng(31)
    geos, tuned = {}, {}
    ops, xs, bs, desc = [], [], [], []
    t0 = time.time()
    for k in range(a.batch_nsys):
        n = int(rng.choice([32, 64, 96, 128]))
        Lt = int(rng.choice([16, 32, 64]))
        lattice = int(rng.integers(0, 2))
        m = float(rng.uniform(0.02, 0.2))
        if n not in geos:
            geos[n] = q.Geometry(n)
        D = q.WilsonDirac(lattice, n, Lt, r=1.0, M5=m, at=0.5 * mean_ell(geos[n], lattice))
        sp, tp = q.random_gauge(lattice, n, Lt, seed=500 + k, kappa=a.beta)
        D.update(sp, tp)
        key = (lattice, n, Lt)
        if key not in tuned:
            D.autotune()
            tuned[key] = {(op, prec): D.tile(op, prec) for op in (q.OP_D, q.OP_DDAG) for prec in (q.F32, q.F64)}
        else:
            for (op, prec), t in tuned[key].items():
                D.set_tile(t["block"], t["ts"], t["remap"], t["lds"], ops=(op,), precs=(prec,))
        ops.append(D)
        b = q.random_source(lattice, n, Lt, seed=900 + k)
        bs.append(q.DeviceArray.from_numpy(b))
        xs.append(q.DeviceArray(b.nbytes))
        desc.append((lattice, n, Lt, round(m, 4)))
    t_setup = time.time() - t0
    pts = sum(D.V * D.Lt for D in ops)
    print(f"# batched: {a.batch_nsys} systems, {pts} points, setup {t_setup:.1f}s",
          file=sys.stderr, flush=True)
    # warm-up: one full call allocates every context's work fields and binary16 phase
    # and halo tables, which the timed calls then reuse
    q.batch_solve(ops, xs, bs, tol=a.tol)
    q.load().qed3_device_synchronize()
    t0 = time.time()
    reps = q.batch_solve(ops, xs, bs, tol=a.tol)
    q.load().qed3_device_synchronize()
    t_batch = time.time() - t0
    t0 = time.time()
    seq = [D.solve_device(x, b, tol=a.tol, solver=q.CG_HALF) for D, x, b in zip(ops, xs, bs)]
    q.load().qed3_device_synchronize()
    t_seq = time.time() - t0
    return {"systems": a.batch_nsys, "points": int(pts), "setup_s": round(t_setup, 1),
            "batched_s": round(t_batch, 3), "one_by_one_s": round(t_seq, 3),
            "solves_per_s_batched": round(a.batch_nsys / t_batch, 2),
            "solves_per_s_one_by_one": round(a.batch_nsys / t_seq, 2),
            "iterations_max": int(max(r.iterations for r in reps)),
            "iterations_mean": float(np.mean([r.iterations for r in reps])),
            "max_true_resid": float(max(r.true_resid for r in reps)),
            "one_by_one_iterations_mean": float(np.mean([r.iterations for r in seq])),
            "shapes_tuned": len(tuned),
            "shapes": desc[:8]}


def cpu_baseline(a, g, at, gpu_iters, sp_full, tp_full):
    """Oracle CSR CG (the reference's CPU path restated) on a bounded sample: the same
    lattice at Lt=cpu_lt, `cpu_iters` CG iterations, scaled to Lt and to the GPU FP64
    iteration count (CG iteration cost is linear in the number of points)."""
    try:
        import oracle as O
        import qed3_amd as q
    except Exception as e:  # noqa: BLE001
        return {"error": str(e)}
    t0 = time.time()
    L = O.Lattice(g, a.lattice)
    Do = O.Dirac(L, a.cpu_lt, r=1.0, M5=a.m, at=at)
    sp, tp = sp_full[: a.cpu_lt].copy(), tp_full[: a.cpu_lt].copy()
    Do.update(sp, tp)
    b = q.random_source(a.lattice, a.n, a.cpu_lt, seed=77)
    setup = time.time() - t0
    t1 = time.time()
    _, it, _ = Do.solve(b, tol=1e-30, maxiter=a.cpu_iters)
    dt = time.time() - t1
    per_iter = dt / max(a.cpu_iters, 1) * (a.Lt / a.cpu_lt)
    sec_per_solve = per_iter * gpu_iters
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(1.0 / sec_per_solve, 6), "unit": "solves/s", "cores": cores,
            "kind": "port",
            "sample": f"oracle CSR CG (matpoly.h:204-267 restated), n={a.n} Lt={a.cpu_lt}, "
                      f"{a.cpu_iters} iterations in {dt:.1f}s (setup {setup:.1f}s), scaled x"
                      f"{a.Lt // a.cpu_lt} in volume and x{gpu_iters:.0f} GPU FP64 iterations",
            "s_per_iteration_full_volume": round(per_iter, 4)}


if __name__ == "__main__":
    main()
