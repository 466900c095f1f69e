"""GPU tests of the learned warm start's network (net.hip) against the FP64 oracle
(oracle/unet_oracle.py, itself pinned by finite differences in tests/test_unet_oracle.py).

The loss weighs each voxel by its stiffness rho^3 (see oracle/unet_oracle.py).

Tolerances: the GPU stores weights, activations and activation gradients in bf16 (8 bits of mantissa)
and accumulates in FP32.  The oracle with bf16=True rounds weights and stored activations the same
way, so the forward pass agrees to a few bf16 ulps (2e-2 of the output scale); gradients pass through
bf16-rounded activation gradients the oracle keeps in FP64, so they agree to 5e-2 in relative L2 norm
per layer, with cosine similarity above 0.998.
"""
import sys

import numpy as np
import pytest

from oracle import unet_oracle as U

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def h3d():
    import homo3d_amd

    return homo3d_amd


def _inputs(g, B, seed):
    rng = np.random.default_rng(seed)
    # smooth periodic densities in [0.05, 1]
    p = np.arange(g) / g
    Z, Y, X = np.meshgrid(p, p, p, indexing="ij")
    rho = []
    for b in range(B):
        a = rng.standard_normal(6)
        f = (a[0] * np.cos(2 * np.pi * X) + a[1] * np.sin(2 * np.pi * Y) + a[2] * np.cos(4 * np.pi * Z) +
             a[3] * np.sin(2 * np.pi * (X + Y)) + a[4] * np.cos(2 * np.pi * (Y - Z)) + a[5])
        rho.append(0.05 + 0.95 / (1 + np.exp(-f)))
    T = rng.standard_normal((B, g, g, g, 18)) * 0.3
    T -= T.mean(axis=(1, 2, 3), keepdims=True)
    return np.array(rho), T


@pytest.mark.parametrize("g,levels,B", [(16, 2, 2), (32, 3, 1)])
def test_forward_matches_oracle(h3d, g, levels, B):
    say("net forward: g=%d levels=%d batch=%d" % (g, levels, B))
    net = h3d.WarmStartNet(grid=g, levels=levels, base_channels=16, batch=B, seed=3)
    theta = U.init_params(7, levels, 16)
    theta[np.arange(theta.size) % 97 == 0] += 0.05  # some non-zero biases
    net.set_params(theta)
    assert net.num_params == U.num_params(levels, 16)
    rho, _ = _inputs(g, B, 1)
    got = net.forward(rho.reshape(B, -1)).reshape(B, g, g, g, 18)
    ref = U.forward(theta.astype(np.float64), rho, levels, 16, bf16=True)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    say("  max abs err / max |P| = %.3e" % err)
    assert err < 2e-2


def test_loss_and_gradient_match_oracle(h3d):
    g, levels, B = 16, 2, 2
    say("net gradient: g=%d levels=%d batch=%d" % (g, levels, B))
    net = h3d.WarmStartNet(grid=g, levels=levels, base_channels=16, batch=B, seed=3)
    theta = U.init_params(11, levels, 16)
    rng = np.random.default_rng(2)
    theta += (0.02 * rng.standard_normal(theta.size)).astype(np.float32)
    net.set_params(theta)
    rho, T = _inputs(g, B, 5)
    T = U.center(T, U.weights(rho, bf16=True))  # targets are mean-free under the loss weights
    loss, grad = net.loss_grad(rho.reshape(B, -1), T.reshape(B, -1, 18))
    Lr, gr = U.loss_and_grad(theta.astype(np.float64), rho, U.bf16_round(T), levels, 16, bf16=True)
    say("  loss gpu %.6e oracle %.6e" % (loss, Lr))
    assert abs(loss - Lr) < 1e-2 * Lr
    o = 0
    for name, ci, co, t, _ in U.layer_specs(levels, 16):
        n_l = co * t * ci + co
        a, b = grad[o:o + n_l].astype(np.float64), gr[o:o + n_l]
        o += n_l
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
        say("  %-8s rel L2 %.3e cos %.6f" % (name, rel, cos))
        assert rel < 5e-2 and cos > 0.998, name
    assert o == grad.size


def test_training_reduces_loss(h3d):
    """Adam steps on a fixed batch with a target the network can represent (a smooth function of the
    density) drive the loss down."""
    g, levels, B = 16, 2, 2
    say("net training: g=%d" % g)
    net = h3d.WarmStartNet(grid=g, levels=levels, base_channels=16, batch=B, seed=4)
    rho, _ = _inputs(g, B, 8)
    T = np.repeat(((rho - rho.mean(axis=(1, 2, 3), keepdims=True)) * 2.0)[..., None], 18, axis=-1)
    T *= np.linspace(-1, 1, 18)
    T = U.center(T, U.weights(rho, bf16=True))
    losses = [net.train_step(rho.reshape(B, -1), T.reshape(B, -1, 18), 2e-3) for _ in range(60)]
    say("  loss %.4e -> %.4e" % (losses[0], losses[-1]))
    assert losses[-1] < 0.2 * losses[0]


def test_solver_pairs_and_warm_start(h3d, orc):
    """Pairs taken from solved handles train; the warm start writes the upsampled prediction (pinned
    vertex at zero) as the FP64 iterate, and the warm solve from it converges to tolerance."""
    n, g = 32, 16
    say("net pairs + warm start: n=%d g=%d" % (n, g))
    net = h3d.WarmStartNet(grid=g, levels=2, base_channels=16, batch=2, seed=5)
    net.dataset_reserve(4)
    hom = h3d.Homogenization(h3d.HomoConfig(reso=n, solver="pcg"))
    xs = []
    for seed in range(4):
        rho = h3d.randc_density(n, seed, 0.3, "reflect6")
        x = (rho.astype(np.float64) ** 3).astype(np.float32)
        hom.update(x)
        st = hom.solve(1e-5, warm=False)
        assert st["status"] == 0
        net.dataset_add(hom)
        xs.append(x)
    assert net.dataset_size() == 4
    l0 = net.eval(0, 4)
    losses = net.train(20, 2e-3, seed=1)
    l1 = net.eval(0, 4)
    say("  dataset loss %.4e -> %.4e (epochs %s)" % (l0, l1, np.array2string(losses[[0, -1]], precision=3)))
    assert l1 < l0
    # warm start on the last design: the FP64 iterate is the trilinear upsampling of the prediction
    f = n // g
    x = xs[-1].astype(np.float64).reshape(n, n, n)
    pooled = np.cbrt(x).reshape(g, f, g, f, g, f).mean(axis=(1, 3, 5))
    P = net.forward(pooled.reshape(1, -1).astype(np.float32))[0].reshape(g, g, g, 18)
    net.warm_start(hom)
    u = hom.get_solution64().reshape(6, 3, n, n, n)  # [case][comp][z][y][x]
    t = (np.arange(n) % f) / f
    i0 = np.arange(n) // f
    i1 = (i0 + 1) % g
    # separable periodic trilinear interpolation along x, y, z (arrays [z][y][x])
    Px = P[:, :, i0, :] * (1 - t)[None, None, :, None] + P[:, :, i1, :] * t[None, None, :, None]
    Py = Px[:, i0] * (1 - t)[None, :, None, None] + Px[:, i1] * t[None, :, None, None]
    Pz = Py[i0] * (1 - t)[:, None, None, None] + Py[i1] * t[:, None, None, None]
    want = (Pz - P[0, 0, 0]) * (n / net.target_scale)
    want = np.moveaxis(want, -1, 0).reshape(6, 3, n, n, n)
    assert np.abs(u[:, :, 0, 0, 0]).max() == 0.0
    err = np.abs(u - want).max() / np.abs(want).max()
    say("  warm start vs numpy upsampling: %.3e" % err)
    assert err < 2e-2
    st = hom.solve(1e-5, warm=True)
    say("  warm solve: iters %s relres %.2e" % (st["iters"], max(st["relres"])))
    assert st["status"] == 0 and max(st["relres"]) <= 1e-5
