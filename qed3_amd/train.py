"""Training data and training loop for the learned preconditioner (host side).

The net N is trained as an approximate inverse of D on pairs (D e, e), each scaled to
|D e| = 1: Gaussian e (every mode weighs the same) and the errors e_k = x* - x_k of the
plain CG iterates after k steps (low modes dominate).  Flexible CG on D^dag D then uses
M = N N^dagger.  Weights are per hop type, so a net trained on a small lattice of the same
family (same lattice kind, r, M5, a_t / <l>, gauge coupling) applies unchanged to a large one.
"""
import numpy as np

from . import CG_F64, DeviceArray, Net, WilsonDirac, random_gauge, random_source


def make_pairs(D: WilsonDirac, lattice, n, Lt, cfg_seeds, betas, rhs_seeds=(1, 2, 3, 4),
               ks=(0, 1, 2, 4, 8, 16, 32, 64), nwhite=8):
    """[(spatial, temporal, r[npairs, N], e[npairs, N])] per gauge configuration."""
    data = []
    for cs, beta in zip(cfg_seeds, betas):
        sp, tp = random_gauge(lattice, n, Lt, seed=cs, kappa=beta)
        D.update(sp, tp)
        rs, es = [], []
        for j in range(nwhite):
            e = random_source(lattice, n, Lt, seed=1000 * cs + j)
            r = D.mult(e)
            s = 1.0 / np.linalg.norm(r)
            rs.append(r * s)
            es.append(e * s)
        for rsd in rhs_seeds:
            b = random_source(lattice, n, Lt, seed=rsd)
            xs, _ = D.solve(b, tol=1e-11, solver=CG_F64)
            for k in ks:
                if k == 0:
                    xk = np.zeros_like(b)
                else:
                    d_b = DeviceArray.from_numpy(b)
                    d_x = DeviceArray(b.nbytes)
                    D.solve_device(d_x, d_b, tol=1e-14, maxiter=k, solver=CG_F64,
                                   raise_on_fail=False)
                    xk = d_x.to_numpy(np.complex128)
                ek = xs - xk
                rk = D.mult(ek)
                s = 1.0 / np.linalg.norm(rk)
                rs.append(rk * s)
                es.append(ek * s)
        data.append((sp, tp, np.array(rs), np.array(es)))
    return data


def train(net: Net, D: WilsonDirac, data, epochs=40, batch=4, lr=3e-3):
    """Adam over the pairs, one configuration at a time per epoch; returns the loss history."""
    hist = []
    dev = [(sp, tp, DeviceArray.from_numpy(rs.astype(np.complex64)),
            DeviceArray.from_numpy(es.astype(np.complex64)), len(rs)) for sp, tp, rs, es in data]
    for _ in range(epochs):
        for sp, tp, d_r, d_e, npairs in dev:
            D.update(sp, tp)
            hist.extend(net.train_device(d_r, d_e, npairs=npairs, batch=batch, epochs=1, lr=lr))
    return hist
