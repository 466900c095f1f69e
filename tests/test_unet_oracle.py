"""Pins of the FP64 network oracle (oracle/unet_oracle.py): the analytic backward against central
finite differences for every layer kind, periodic equivariance, translation invariance of the loss,
the bf16 rounding and the parameter count."""
import numpy as np
import pytest

from oracle import unet_oracle as U


def test_param_count_closed_form():
    # c0 = 16, 4 levels: sum over layers of (27 cin cout + cout), head 16*18 + 18
    ch = [16, 32, 64, 128]
    n = 0
    cin = 1
    for c in ch:
        n += 27 * cin * c + c + 27 * c * c + c
        cin = c
    for l in (2, 1, 0):
        n += 27 * (ch[l + 1] + ch[l]) * ch[l] + ch[l] + 27 * ch[l] * ch[l] + ch[l]
    n += 16 * 18 + 18
    assert U.num_params(4, 16) == n


def test_weighted_loss_ignores_void_cells():
    """The loss weights are the stiffness rho^3: errors where rho = 1e-3 (weight 1e-9 of a solid cell)
    change it by nothing measurable, errors in solid cells by their full square."""
    g, B = 8, 1
    rho = np.full((B, g, g, g), 1e-3)
    rho[:, :, :, :4] = 1.0  # half solid
    w = U.weights(rho)
    T = U.center(np.zeros((B, g, g, g, U.NOUT)), w)
    P = np.zeros_like(T)
    assert U.loss(P, T, rho) == 0.0
    Pv = P.copy()
    Pv[:, :, :, 6, :] += 1.0  # a void plane
    Ps = P.copy()
    Ps[:, :, :, 1, :] += 1.0  # a solid plane
    assert U.loss(Pv, T, rho) < 1e-8 * U.loss(Ps, T, rho)
    assert U.loss(Ps, T, rho) > 0.05


def test_bf16_round_known_values():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 1e-40, 3.0e38])
    r = U.bf16_round(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0  # tie -> even mantissa
    assert r[2] == 1.0 + 2 ** -7  # 1.5 ulp rounds up
    assert r[3] == -2.5
    # exactly representable bf16 values are fixed points
    assert np.array_equal(U.bf16_round(r), r)


def _setup(levels=3, c0=4, g=8, B=2, seed=0):
    rng = np.random.default_rng(seed)
    theta = U.init_params(seed, levels, c0).astype(np.float64)
    theta += 0.05 * rng.standard_normal(theta.size)  # non-zero biases
    rho = rng.uniform(0.05, 1.0, (B, g, g, g))
    T = rng.standard_normal((B, g, g, g, U.NOUT)) * 0.1
    T = U.center(T, U.weights(rho))
    return theta, rho, T


def test_backward_matches_finite_differences():
    levels, c0 = 3, 4
    theta, rho, T = _setup(levels, c0)
    L, g = U.loss_and_grad(theta, rho, T, levels, c0)
    # probe parameters in every layer: weights of several taps and the biases
    rng = np.random.default_rng(1)
    o = 0
    idx = []
    for name, ci, co, t, _ in U.layer_specs(levels, c0):
        nW = co * t * ci
        idx += list(o + rng.choice(nW, 4, replace=False))
        idx.append(o + nW + rng.integers(co))
        o += nW + co
    assert o == theta.size
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (U.loss(U.forward(tp, rho, levels, c0), T, rho) - U.loss(U.forward(tm, rho, levels, c0), T, rho)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, np.abs(g).max()) + 1e-4 * abs(fd), (i, fd, g[i])


def test_forward_is_periodic_equivariant_and_loss_translation_invariant():
    levels, c0 = 3, 4
    theta, rho, T = _setup(levels, c0)
    P = U.forward(theta, rho, levels, c0)
    # shifting the input by a multiple of the coarsest cell (4 voxels at 3 levels) shifts the output
    Ps = U.forward(theta, np.roll(rho, 4, axis=3), levels, c0)
    assert np.allclose(Ps, np.roll(P, 4, axis=3), atol=1e-12)
    # adding a constant per channel to the prediction leaves the loss unchanged
    L0 = U.loss(P, T, rho)
    assert abs(U.loss(P + np.arange(18.0), T, rho) - L0) < 1e-12


def test_bf16_forward_close_to_fp64():
    levels, c0 = 3, 8
    theta, rho, T = _setup(levels, c0)
    P = U.forward(theta, rho, levels, c0)
    Pb = U.forward(theta, rho, levels, c0, bf16=True)
    assert np.abs(Pb - P).max() < 5e-2 * np.abs(P).max()
