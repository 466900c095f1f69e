"""FP64 NumPy restatement of the learned warm-start network (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module; the
product library never does.

The network is the project's own (the reference has no learned component: its warm start is the
previous design's stored solutions, homogenization.cu:471-477 / grid.cpp:459-462).  It maps the
density rho on a periodic g^3 grid to the six load cases' displacement fields (18 channels) at the
g^3 vertices, scaled by 1/n and mean-free per channel:

  encoder level l = 0..L-1, channels c_l = C0 2^l, grid g / 2^l:
      E_l1 = relu(conv3(in_l) + b), S_l = relu(conv3(E_l1) + b)
      in_0 = rho (1 channel), in_{l+1} = avgpool2(S_l)
  decoder level l = L-2..0 (D_{L-1} = S_{L-1}):
      CAT_l = [up2(D_{l+1}) | S_l]  (channels c_{l+1} + c_l)
      D_l = relu(conv3(relu(conv3(CAT_l) + b)) + b)
  head: P = conv1(D_0) + b (18 channels, no activation)

conv3 is a 3x3x3 periodic correlation, tap t = (dx+1) + 3 (dy+1) + 9 (dz+1) reads X[v + (dx,dy,dz)];
weights W[co][t][ci].  up2 is nearest-neighbour.  Tensors are channels-last [B, z, y, x, C].

Loss: mean over (B, g^3, 18) of (w / mean_v w) (P - m_w(P) - T)^2 with the stiffness weights
w = rho^3 of the input, m_w the w-weighted mean over a sample's voxels, and T w-mean-free per
(sample, channel): the displacement of near-void cells (stiffness ~1e-9) is barely determined by the
solve and carries no weight, as in the energy norm.

Pins (tests/test_unet_oracle.py): central finite differences of the loss against the analytic
backward on every layer kind; translation invariance of the loss; periodic equivariance of the
forward (a rolled input gives the rolled output); the parameter count in closed form.

bf16=True rounds the weights and every stored activation to bfloat16 (round to nearest even) the way
the GPU path stores them, so that the GPU kernels can be compared with FP32-accumulation tolerances.
"""
import numpy as np

NOUT = 18


def bf16_round(a):
    """Round float64/float32 values to bfloat16 (nearest, ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    rounding = 0x7FFF + ((u >> 16) & 1)
    r = ((u + rounding) & 0xFFFF0000).astype(np.uint32)
    out = r.view(np.float32).astype(np.float64)
    return np.where(np.isfinite(f), out, f.astype(np.float64))


def layer_specs(levels=4, c0=16):
    """[(name, cin, cout, taps, relu)] in parameter order."""
    ch = [c0 << l for l in range(levels)]
    specs = []
    for l in range(levels):
        cin = 1 if l == 0 else ch[l - 1]
        specs.append(("enc%d.c1" % l, cin, ch[l], 27, True))
        specs.append(("enc%d.c2" % l, ch[l], ch[l], 27, True))
    for l in range(levels - 2, -1, -1):
        specs.append(("dec%d.c1" % l, ch[l + 1] + ch[l], ch[l], 27, True))
        specs.append(("dec%d.c2" % l, ch[l], ch[l], 27, True))
    specs.append(("head", ch[0], NOUT, 1, False))
    return specs


def num_params(levels=4, c0=16):
    return sum(co * t * ci + co for _, ci, co, t, _ in layer_specs(levels, c0))


def unpack(theta, levels=4, c0=16):
    """Flat FP32/FP64 parameter vector -> {name: (W[co][t][ci], b[co])}."""
    out, o = {}, 0
    for name, ci, co, t, _ in layer_specs(levels, c0):
        W = np.asarray(theta[o:o + co * t * ci], np.float64).reshape(co, t, ci)
        o += co * t * ci
        b = np.asarray(theta[o:o + co], np.float64)
        o += co
        out[name] = (W, b)
    assert o == len(theta)
    return out


def init_params(seed, levels=4, c0=16):
    """He-normal weights (std sqrt(2 / (taps cin)); head std sqrt(1 / cin) / 4), zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, ci, co, t, relu in layer_specs(levels, c0):
        std = np.sqrt(2.0 / (t * ci)) if relu else np.sqrt(1.0 / ci) / 4
        parts.append(rng.standard_normal(co * t * ci) * std)
        parts.append(np.zeros(co))
    return np.concatenate(parts).astype(np.float32)


def _offsets(taps):
    if taps == 1:
        return [(0, 0, 0)]
    return [(t % 3 - 1, t // 3 % 3 - 1, t // 9 - 1) for t in range(27)]


def _shift(X, d):
    """Xs[v] = X[v + d] on the periodic grid (axes 1..3 = z, y, x)."""
    dx, dy, dz = d
    return np.roll(X, shift=(-dz, -dy, -dx), axis=(1, 2, 3))


def conv(X, W, b):
    co, t, ci = W.shape
    Y = np.zeros(X.shape[:4] + (co,))
    for k, d in enumerate(_offsets(t)):
        Y += _shift(X, d) @ W[:, k, :].T
    return Y + b


def conv_backward(X, W, dZ):
    co, t, ci = W.shape
    dX = np.zeros(X.shape)
    dW = np.zeros(W.shape)
    flat = dZ.reshape(-1, co)
    for k, d in enumerate(_offsets(t)):
        Xs = _shift(X, d)
        dW[:, k, :] = flat.T @ Xs.reshape(-1, ci)
        dX += _shift(dZ @ W[:, k, :], (-d[0], -d[1], -d[2]))
    return dX, dW, flat.sum(axis=0)


def pool(X):
    B, g = X.shape[0], X.shape[1]
    h = g // 2
    return X.reshape(B, h, 2, h, 2, h, 2, -1).mean(axis=(2, 4, 6))


def pool_backward(dP):
    return np.repeat(np.repeat(np.repeat(dP, 2, axis=1), 2, axis=2), 2, axis=3) / 8.0


def up(X):
    return np.repeat(np.repeat(np.repeat(X, 2, axis=1), 2, axis=2), 2, axis=3)


def up_backward(dU):
    B, g = dU.shape[0], dU.shape[1]
    h = g // 2
    return dU.reshape(B, h, 2, h, 2, h, 2, -1).sum(axis=(2, 4, 6))


def forward(theta, rho, levels=4, c0=16, bf16=False, keep=False):
    """rho: [B, g, g, g] (z, y, x).  Returns P [B, g, g, g, 18] (and the stored activations)."""
    rnd = bf16_round if bf16 else (lambda a: a)
    P_ = unpack(theta, levels, c0)
    P_ = {k: (rnd(W), np.asarray(b, np.float64)) for k, (W, b) in P_.items()}
    acts = {}
    x = rnd(np.asarray(rho, np.float64)[..., None])
    acts["in0"] = x
    S = []
    for l in range(levels):
        if l > 0:
            x = rnd(pool(S[l - 1]))
            acts["in%d" % l] = x
        W, b = P_["enc%d.c1" % l]
        e1 = rnd(np.maximum(conv(x, W, b), 0))
        acts["enc%d.c1" % l] = e1
        W, b = P_["enc%d.c2" % l]
        s = rnd(np.maximum(conv(e1, W, b), 0))
        acts["enc%d.c2" % l] = s
        S.append(s)
    D = S[levels - 1]
    for l in range(levels - 2, -1, -1):
        cat = np.concatenate([up(D), S[l]], axis=-1)
        acts["cat%d" % l] = cat
        W, b = P_["dec%d.c1" % l]
        d1 = rnd(np.maximum(conv(cat, W, b), 0))
        acts["dec%d.c1" % l] = d1
        W, b = P_["dec%d.c2" % l]
        D = rnd(np.maximum(conv(d1, W, b), 0))
        acts["dec%d.c2" % l] = D
    W, b = P_["head"]
    P = conv(D, W, b)
    if keep:
        return P, acts, P_
    return P


def weights(rho, bf16=False):
    """Stiffness weights w = rho^3 / mean(rho^3) per sample, [B, g, g, g, 1]."""
    r = bf16_round(rho) if bf16 else np.asarray(rho, np.float64)
    w = r ** 3
    return (w / w.mean(axis=(1, 2, 3), keepdims=True))[..., None]


def center(P, w):
    """P minus its w-weighted mean over each sample's voxels, per channel."""
    return P - (w * P).sum(axis=(1, 2, 3), keepdims=True) / w.sum(axis=(1, 2, 3), keepdims=True)


def loss(P, T, rho, bf16=False):
    w = weights(rho, bf16)
    return float(np.mean(w * (center(P, w) - T) ** 2))


def loss_and_grad(theta, rho, T, levels=4, c0=16, bf16=False):
    """Loss and its gradient with respect to the flat parameter vector (same layout as theta)."""
    P, acts, P_ = forward(theta, rho, levels, c0, bf16, keep=True)
    w = weights(rho, bf16)
    d = center(P, w) - T
    L = float(np.mean(w * d ** 2))
    # d/dP of the weighted mean term vanishes when sum_v w d = 0 (T w-mean-free); keep it general:
    wd = w * d
    dP = 2.0 * (wd - w * wd.sum(axis=(1, 2, 3), keepdims=True) / w.sum(axis=(1, 2, 3), keepdims=True)) / d.size
    grads = {}
    W, b = P_["head"]
    dD, dW, db = conv_backward(acts["dec0.c2"] if levels > 1 else acts["enc0.c2"], W, dP)
    grads["head"] = (dW, db)
    dS = {}
    # decoder, from level 0 upwards (D_l = dec l.c2(dec l.c1(cat_l)))
    for l in range(0, levels - 1):
        d2 = acts["dec%d.c2" % l]
        d1 = acts["dec%d.c1" % l]
        W, b = P_["dec%d.c2" % l]
        dZ = dD * (d2 > 0)
        dd1, dW, db = conv_backward(d1, W, dZ)
        grads["dec%d.c2" % l] = (dW, db)
        W, b = P_["dec%d.c1" % l]
        dZ = dd1 * (d1 > 0)
        dcat, dW, db = conv_backward(acts["cat%d" % l], W, dZ)
        grads["dec%d.c1" % l] = (dW, db)
        cu = P_["dec%d.c1" % l][0].shape[2] - (c0 << l)
        dS[l] = dcat[..., cu:]
        dD = up_backward(dcat[..., :cu])  # gradient of D_{l+1}
    # encoder, from the bottom up: dD is now the gradient of S_{L-1}
    dSl = dD
    for l in range(levels - 1, -1, -1):
        if l < levels - 1:
            dSl = dS[l] + pool_backward(dIn)
        s = acts["enc%d.c2" % l]
        e1 = acts["enc%d.c1" % l]
        W, b = P_["enc%d.c2" % l]
        dZ = dSl * (s > 0)
        de1, dW, db = conv_backward(e1, W, dZ)
        grads["enc%d.c2" % l] = (dW, db)
        W, b = P_["enc%d.c1" % l]
        dZ = de1 * (e1 > 0)
        dIn, dW, db = conv_backward(acts["in%d" % l], W, dZ)
        grads["enc%d.c1" % l] = (dW, db)
    flat = []
    for name, ci, co, t, _ in layer_specs(levels, c0):
        dW, db = grads[name]
        flat.append(dW.ravel())
        flat.append(db)
    return L, np.concatenate(flat)
