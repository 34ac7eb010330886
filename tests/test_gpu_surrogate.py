"""The surrogate network's HIP kernels (BF16 matrix-core GEMMs, SiLU, MSE,
Adam) against the CPU oracle oracle/mlp_oracle.py, which applies the same
BF16 roundings: forward values to 5e-3 and gradients to 2e-2 relative to
each block's largest element (FP32 vs FP64 accumulation can move a value
across a BF16 rounding boundary).  Then training and the warm start itself:
a network trained on small FLEX solves must start the loop closer than the
cold start does."""
import numpy as np
import pytest

import oracle.mlp_oracle as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import pyrpa_amd.surrogate as S
    return S


def _net(S, dims, seed):
    net = S.Net(dims, seed=seed)
    rng = np.random.default_rng(seed)
    flat = (rng.standard_normal(net.num_params) * 0.3).astype(np.float32)
    net.set_params(flat)
    return net, flat


@pytest.mark.parametrize("dims,B", [([4, 64, 48, 80], 37), ([4, 16, 16], 16), ([7, 130, 33], 65)])
def test_forward_matches_oracle(S, dims, B):
    net, flat = _net(S, dims, 1)
    assert np.array_equal(net.get_params(), flat)
    X = np.random.default_rng(2).standard_normal((B, dims[0])).astype(np.float32)
    Y = net.forward(X)
    Yo, _ = M.forward(flat.astype(np.float64), dims, X.astype(np.float64))
    assert np.abs(Y - Yo).max() < 5e-3 * np.abs(Yo).max()


@pytest.mark.parametrize("dims,B", [([4, 64, 48, 80], 37), ([5, 96, 64, 64, 40], 50)])
def test_gradients_match_oracle(S, dims, B):
    net, flat = _net(S, dims, 3)
    rng = np.random.default_rng(4)
    X = rng.standard_normal((B, dims[0])).astype(np.float32)
    T = rng.standard_normal((B, dims[-1])).astype(np.float32)
    loss, g = net.loss_grad(X, T)
    lo, go = M.loss_grad(flat.astype(np.float64), dims, X.astype(np.float64), T.astype(np.float64))
    assert abs(loss - lo) < 1e-3 * abs(lo)
    off = 0
    for i in range(len(dims) - 1):
        for n in (dims[i] * dims[i + 1], dims[i + 1]):
            a, b = g[off:off + n], go[off:off + n]
            assert np.abs(a - b).max() < 2e-2 * np.abs(b).max(), (i, n)
            off += n


def test_training_reduces_the_loss(S):
    rng = np.random.default_rng(5)
    P = rng.uniform(-1, 1, size=(256, 4))
    Y = np.stack([np.sin(2 * P[:, 0]) * P[:, 1], np.cos(P[:, 2] + P[:, 3]), P[:, 0] * P[:, 3]], axis=1)
    net = S.Net([4, 64, 64, 3], seed=6)
    net.set_data(P.astype(np.float32), Y.astype(np.float32))
    first = net.train_epoch(np.arange(256), 32, 3e-3)
    for _ in range(60):
        last = net.train_epoch(rng.permutation(256), 32, 3e-3)
    assert last < 0.05 * first


# parameters away from the magnetic instability (Stoner factor < 1), where
# the FLEX fixed point is unique
SAFE = {"n": (0.75, 0.85), "U": (2.0, 2.5), "T": (0.03, 0.05)}


def test_trained_network_starts_flex_closer_than_cold(S):
    """A small surrogate (16x16 mesh, 64 frequencies, 16 predicted) trained
    on 40 solves: at 3 held-out points on a 32x32 x 128 mesh, the seed it
    gives needs fewer FLEX iterations to 1e-8 than the cold start, and the
    converged Sigma agrees with the cold solve."""
    import pyrpa_amd.flibs as F
    import pyrpa_amd.lattice as L
    P, Y, _, _ = S.generate(40, seed=11, nw_out=16, eps=1e-8, Nx=16, Ny=16, Nw=64, ranges=SAFE)
    sur = S.Surrogate(nw_out=16, width=128, layers=4, seed=0)
    sur.fit(P, Y, epochs=400, batch=16, lr=2e-3)
    gains = []
    for p in S.sample_params(3, seed=99, ranges=SAFE):
        st = S.lattice_state(p, 32, 32, 128)
        F.set_precision(0)
        s_cold, mu_cold = L.flex(st, eps=1e-8, sw_rescale=True)
        c = F.flex_stats()
        seed = sur.seed_for(p, st, train_mesh=(16, 16))
        F.set_sigma_seed(*seed)
        s_warm, mu_warm = L.flex(st, eps=1e-8, sw_rescale=True, sw_in=True)
        w = F.flex_stats()
        info = (p, c, w, mu_cold, mu_warm)
        assert c["esterr"] < 1e-8 and w["esterr"] < 1e-8 and c["stoner"] < 1.0, info
        assert np.abs(s_warm - s_cold).max() < 1e-6 * np.abs(s_cold).max(), info
        assert abs(mu_warm - mu_cold) < 1e-6, info
        gains.append(c["iterations"] - w["iterations"])
    assert np.mean(gains) > 0, gains
