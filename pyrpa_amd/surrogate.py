"""Learned warm start of the FLEX loop (no reference counterpart).

A fully connected network maps the parameters of the one-orbital square
lattice -- filling n, Hubbard U, temperature T and next-nearest hopping t'
(in units of t) -- to the self-energy Sigma(k, i w_n) on the first `nw_out`
Matsubara frequencies of a 32x32 k-mesh (irreducible points, real and
imaginary parts) and the chemical potential.  It is trained on FLEX solves
that the GPU plain method itself generates, by this library's kernels
(gmod_nn_*: BF16 matrix-core GEMMs with FP32 accumulation, Adam).  At a new
parameter point its prediction is re-gridded on the device
(gmod_regrid_sigma) onto the target mesh and handed to
mkself(sw_in=True) as the starting Sigma; the plain loop still runs to the
same tolerance, so the network only changes where the iteration starts.

The network weights and the feature / target scalings are saved with numpy
(.npz, no pickles).
"""
from __future__ import annotations

import time
from ctypes import c_double, c_int64, c_uint64

import numpy as np

from . import flibs as F
from . import lattice as L
from ._loader import check, lib

RANGES = {"n": (0.75, 0.95), "U": (2.0, 4.0), "T": (0.01, 0.05), "tp": (-0.35, 0.0)}
KEYS = ("n", "U", "T", "tp")

_bound = False


def _L():
    global _bound
    lb = lib()
    if not _bound:
        nd = np.ctypeslib.ndpointer
        f32, i64 = nd(dtype=np.float32), nd(dtype=np.int64)
        lb.gmod_nn_create.argtypes = [c_int64, i64, c_uint64]
        lb.gmod_nn_create.restype = c_int64
        lb.gmod_nn_destroy.argtypes = [c_int64]
        lb.gmod_nn_num_params.argtypes = [c_int64]
        lb.gmod_nn_num_params.restype = c_int64
        lb.gmod_nn_get_params.argtypes = [c_int64, f32]
        lb.gmod_nn_set_params.argtypes = [c_int64, f32]
        lb.gmod_nn_forward.argtypes = [c_int64, f32, c_int64, f32]
        lb.gmod_nn_grad.argtypes = [c_int64, f32, f32, c_int64, f32]
        lb.gmod_nn_grad.restype = c_double
        lb.gmod_nn_set_data.argtypes = [c_int64, f32, f32, c_int64]
        lb.gmod_nn_train_epoch.argtypes = [c_int64, i64, c_int64, c_int64, c_double]
        lb.gmod_nn_train_epoch.restype = c_double
        _bound = True
    return lb


class Net:
    """Handle of a network that lives on the GPU (libgmod's nn kernels)."""

    def __init__(self, dims, seed=0):
        self.dims = [int(d) for d in dims]
        self.h = _L().gmod_nn_create(len(self.dims) - 1, np.array(self.dims, dtype=np.int64), seed)
        check()

    def __del__(self):
        try:
            _L().gmod_nn_destroy(self.h)
        except Exception:
            pass

    @property
    def num_params(self):
        return int(_L().gmod_nn_num_params(self.h))

    def get_params(self):
        out = np.zeros(self.num_params, dtype=np.float32)
        _L().gmod_nn_get_params(self.h, out)
        check()
        return out

    def set_params(self, flat):
        _L().gmod_nn_set_params(self.h, np.ascontiguousarray(flat, dtype=np.float32))
        check()

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], self.dims[-1]), dtype=np.float32)
        _L().gmod_nn_forward(self.h, X, X.shape[0], Y)
        check()
        return Y

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(self.num_params, dtype=np.float32)
        loss = _L().gmod_nn_grad(self.h, X, T, X.shape[0], g)
        check()
        return loss, g

    def set_data(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        _L().gmod_nn_set_data(self.h, X, T, X.shape[0])
        check()

    def train_epoch(self, order, batch, lr):
        loss = _L().gmod_nn_train_epoch(self.h, np.ascontiguousarray(order, dtype=np.int64), len(order), batch, lr)
        check()
        return loss


# --------------------------------------------------------------------------- data
def sample_params(count, seed, ranges=None):
    """Seeded uniform samples of (n, U, T, t') in `ranges` (default: the
    trained ranges)."""
    rng = np.random.default_rng(seed)
    r = dict(RANGES, **(ranges or {}))
    return np.stack([rng.uniform(*r[k], size=count) for k in KEYS], axis=1)


def lattice_state(p, Nx=32, Ny=32, Nw=256):
    n, U, T, tp = p
    return L.config(5, Nx=Nx, Ny=Ny, Nw=Nw, n=n, U=U, T=T, tp=tp)


def features(P):
    """Parameters scaled to [-1, 1]."""
    P = np.atleast_2d(P)
    lo = np.array([RANGES[k][0] for k in KEYS])
    hi = np.array([RANGES[k][1] for k in KEYS])
    return (2.0 * (P - lo) / (hi - lo) - 1.0).astype(np.float32)


def target_vector(sig, mu, nw_out):
    """[Re Sigma, Im Sigma] on the first nw_out frequencies (all irreducible
    k) and mu, as one vector."""
    s = sig[0, 0, :nw_out, :]
    return np.concatenate([s.real.ravel(), s.imag.ravel(), [mu]])


def split_target(v, nw_out, nk):
    m = nw_out * nk
    s = (v[:m] + 1j * v[m:2 * m]).reshape(nw_out, nk)
    return s, float(v[2 * m])


def generate(count, seed, nw_out=64, eps=1e-8, precision=0, log=None, Nx=32, Ny=32, Nw=256, ranges=None):
    """Run `count` plain FLEX solves at seeded parameter points.  Returns
    (P, Y, iters, seconds) with Y the target vectors, for the solves that
    reached the tolerance (the others are dropped: a non-converged Sigma is
    no target)."""
    P = sample_params(count, seed, ranges)
    Y, iters, keep = [], [], []
    old = F.get_precision()
    t0 = time.perf_counter()
    try:
        F.set_precision(precision)
        for i, p in enumerate(P):
            st = lattice_state(p, Nx, Ny, Nw)
            sig, mu = L.flex(st, eps=eps, scf_loop=300, sw_rescale=True)
            stats = F.flex_stats()
            Y.append(target_vector(sig, mu, nw_out))
            iters.append(int(stats["iterations"]))
            keep.append(stats["esterr"] < eps)
            if log and (i + 1) % max(1, count // 10) == 0:
                log(f"[surrogate] generated {i + 1}/{count} solves ({time.perf_counter() - t0:.1f} s)")
    finally:
        F.set_precision(old)
    keep = np.array(keep, dtype=bool)
    return P[keep], np.array(Y)[keep], np.array(iters)[keep], time.perf_counter() - t0


# --------------------------------------------------------------------------- model
class Surrogate:
    def __init__(self, nw_out=64, nk=None, width=512, layers=4, seed=0):
        self.nw_out = nw_out
        self.nk = nk
        self.width = width
        self.layers = layers
        self.seed = seed
        self.net = None
        self.ymean = self.ystd = None

    def _dims(self, dout):
        return [len(KEYS)] + [self.width] * (self.layers - 1) + [dout]

    def fit(self, P, Y, epochs=300, batch=64, lr=1e-3, log=None):
        """Standardise targets, train with Adam (cosine learning-rate decay)."""
        X = features(P)
        self.ymean = Y.mean(axis=0)
        self.ystd = Y.std(axis=0) + 1e-8 * (np.abs(Y).max() + 1e-30)
        Z = ((Y - self.ymean) / self.ystd).astype(np.float32)
        self.nk = (Y.shape[1] - 1) // (2 * self.nw_out)
        self.net = Net(self._dims(Y.shape[1]), seed=self.seed)
        self.net.set_data(X, Z)
        rng = np.random.default_rng(self.seed + 1)
        hist = []
        for ep in range(epochs):
            lr_e = lr * 0.5 * (1.0 + np.cos(np.pi * ep / epochs))
            hist.append(self.net.train_epoch(rng.permutation(len(X)), batch, lr_e))
            if log and (ep + 1) % max(1, epochs // 10) == 0:
                log(f"[surrogate] epoch {ep + 1}/{epochs}: standardised MSE {hist[-1]:.3e}")
        return np.array(hist)

    def predict(self, P):
        Z = self.net.forward(features(P))
        return Z * self.ystd + self.ymean

    def save(self, path):
        np.savez(path, params=self.net.get_params(), dims=np.array(self.net.dims), ymean=self.ymean, ystd=self.ystd,
                 nw_out=self.nw_out, nk=self.nk, width=self.width, layers=self.layers, seed=self.seed)

    @classmethod
    def load(cls, path):
        d = np.load(path, allow_pickle=False)
        s = cls(int(d["nw_out"]), int(d["nk"]), int(d["width"]), int(d["layers"]), int(d["seed"]))
        s.ymean, s.ystd = d["ymean"], d["ystd"]
        s.net = Net(d["dims"], seed=s.seed)
        s.net.set_params(d["params"])
        return s

    def seed_for(self, p, st, train_mesh=(32, 32)):
        """Warm-start (sigma, mu, mu_old) for mkself on workload `st`
        (one-orbital lattice at parameters p)."""
        v = self.predict(np.atleast_2d(p))[0].astype(np.float64)
        s, mu = split_target(v, self.nw_out, self.nk)
        Nx, Ny = train_mesh
        _, kmap_c, invk_c = F.gen_irr_k_TRS(Nx, Ny, 1)
        sig_c = np.ascontiguousarray(s[None, None, :, :])
        sig = F.regrid_sigma(sig_c, kmap_c, invk_c, (Nx, Ny, 1, self.nw_out), st["kmap"], st["invk"],
                             (st["Nx"], st["Ny"], st["Nz"], st["Nw"]), st["temp"])
        return sig, mu, mu


# --------------------------------------------------------------------------- experiment
def experiment(n_train=2000, n_eval=200, train_mesh=(32, 32, 256), eval_mesh=(64, 64, 1024), nw_out=64,
               epochs=300, batch=64, lr=1e-3, seed=20261207, holdout=0.1, log=print, model_path=None):
    """Generate training solves with the plain GPU method, train the network
    on them, and evaluate the warm start at held-out parameter points on the
    evaluation mesh: FLEX iterations and time to 1e-8 from the cold start and
    from the network's seed, and the converged-Sigma deviation."""
    Nx, Ny, Nw = train_mesh
    t0 = time.perf_counter()
    P, Y, iters, t_gen = generate(n_train, seed, nw_out=nw_out, eps=1e-8, precision=0, log=log, Nx=Nx, Ny=Ny, Nw=Nw)
    n_ok = len(P)
    log(f"[surrogate] {n_train} training solves in {t_gen:.1f} s: {n_ok} converged to 1e-8 "
        f"(mean {iters.mean():.1f} iterations)")
    rng = np.random.default_rng(seed + 1)
    perm = rng.permutation(n_ok)
    n_val = int(round(holdout * n_ok))
    tr, va = perm[n_val:], perm[:n_val]
    sur = Surrogate(nw_out=nw_out, width=512, layers=4, seed=seed % 1000)
    t1 = time.perf_counter()
    hist = sur.fit(P[tr], Y[tr], epochs=epochs, batch=batch, lr=lr, log=log)
    t_fit = time.perf_counter() - t1
    pred = sur.predict(P[va])
    val_rel = float(np.abs(pred[:, :-1] - Y[va, :-1]).max(axis=1).mean() / np.abs(Y[va, :-1]).max(axis=1).mean())
    log(f"[surrogate] trained in {t_fit:.1f} s; validation mean max|dSigma|/max|Sigma| {val_rel:.3e}")
    if model_path:
        sur.save(model_path)
    Ex, Ey, Ew = eval_mesh
    rows = []
    old = F.get_precision()
    try:
        for i, p in enumerate(sample_params(n_eval, seed + 2)):
            st = lattice_state(p, Ex, Ey, Ew)
            F.set_precision(0)
            s_cold, mu_cold = L.flex(st, eps=1e-8, sw_rescale=True)
            c = F.flex_stats()
            tw = time.perf_counter()
            sd = sur.seed_for(p, st, train_mesh=(Nx, Ny))
            t_seed = time.perf_counter() - tw
            F.set_sigma_seed(*sd)
            s_warm, mu_warm = L.flex(st, eps=1e-8, sw_rescale=True, sw_in=True)
            w = F.flex_stats()
            rows.append(dict(p=[float(x) for x in p], it_cold=int(c["iterations"]), it_warm=int(w["iterations"]),
                             t_cold=c["solve_seconds"], t_warm=w["solve_seconds"] + t_seed,
                             conv_cold=bool(c["esterr"] < 1e-8), conv_warm=bool(w["esterr"] < 1e-8),
                             stoner=c["stoner"],
                             seed_rel=float(np.abs(sd[0] - s_cold).max() / np.abs(s_cold).max()),
                             sigma_dev=float(np.abs(s_warm - s_cold).max() / np.abs(s_cold).max()),
                             mu_dev=abs(mu_warm - mu_cold)))
            if (i + 1) % max(1, n_eval // 10) == 0:
                r = rows[-1]
                log(f"[surrogate] eval {i + 1}/{n_eval}: cold {r['it_cold']} it / warm {r['it_warm']} it, "
                    f"seed err {r['seed_rel']:.2e}, converged dev {r['sigma_dev']:.2e}")
    finally:
        F.set_precision(old)
    ok = [r for r in rows if r["conv_cold"] and r["conv_warm"]]
    same = [r for r in ok if r["sigma_dev"] < 1e-6]
    rep = {
        "train": {"solves": n_train, "converged": n_ok, "mesh": list(train_mesh), "seconds": t_gen,
                  "mean_iterations": float(iters.mean()),
                  "fit_seconds": t_fit, "epochs": epochs, "final_loss": float(hist[-1]), "validation_rel_err": val_rel},
        "eval": {"points": n_eval, "mesh": list(eval_mesh), "both_converged": len(ok),
                 "same_fixed_point": len(same),
                 "iterations_cold_mean": float(np.mean([r["it_cold"] for r in same])) if same else None,
                 "iterations_warm_mean": float(np.mean([r["it_warm"] for r in same])) if same else None,
                 "seconds_cold_total": float(sum(r["t_cold"] for r in same)),
                 "seconds_warm_total": float(sum(r["t_warm"] for r in same)),
                 "seed_rel_err_median": float(np.median([r["seed_rel"] for r in rows])) if rows else None,
                 "converged_sigma_dev_max": float(max(r["sigma_dev"] for r in same)) if same else None},
        "seconds": time.perf_counter() - t0,
        "rows": rows,
    }
    return rep


if __name__ == "__main__":
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="train and evaluate the FLEX warm-start surrogate on the GPU")
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--eval", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--out", default=None, help="JSON report path")
    ap.add_argument("--model", default=None, help="save the trained network (.npz)")
    a = ap.parse_args()

    def _log(msg):
        print(msg, file=sys.stderr, flush=True)
    rep = experiment(n_train=a.train, n_eval=a.eval, epochs=a.epochs, log=_log, model_path=a.model)
    txt = json.dumps(rep)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt)
    print(json.dumps({k: v for k, v in rep.items() if k != "rows"}), flush=True)
