"""FP64 NumPy restatement of the density filter and projection of the reference's expression templates
(TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module).

  conv with radial_convker_t<Spline4>(r) (AutoDiff/TensorExpression.h:1007-1046, forward and backward
  kernels TensorExpression.cuh:288-368): half-width h = int(r + 0.5), weight
  w(off) = 1 - 6 q + 8 q sqrt(q) - 3 q^2 for q = |off|^2 / r^2 < 1 (else 0), normalised by the sum
  over the box; out(i) = sum_off w(off) in(i + off) / sum_off w(off) over the neighbours inside the
  grid (all of them when periodic); the backward pass is the exact adjoint,
  din(j) = sum_off w(off) dout(j - off) / W(j - off).
  sgm(k = 30, c = 0.5) (AutoDiff.h:808-823): y = 1 / (1 + exp(-k (x - c))), dy/dx = k y (1 - y).

Arrays are [z][y][x] (x fastest in memory).  Pins (tests/test_filter_oracle.py): the impulse response
is the hand-computed Spline4 table; forward and backward pass the dot-product (adjoint) test; constants
are preserved; sgm's derivative matches central differences.
"""
import numpy as np


def spline4(off2, r):
    q = off2 / (r * r)
    return np.where(q < 1, 1 - 6 * q + 8 * q * np.sqrt(q) - 3 * q * q, 0.0)


def linear(off2, r):
    """WeightFunc<Linear> (TensorExpression.h:1020-1031): 1 - |off| / r inside the radius."""
    q = off2 / (r * r)
    return np.where(q < 1, 1 - np.sqrt(q), 0.0)


def _box(r, kernel="spline4"):
    h = int(r + 0.5)
    f = spline4 if kernel == "spline4" else linear
    offs = [(dx, dy, dz) for dz in range(-h, h + 1) for dy in range(-h, h + 1) for dx in range(-h, h + 1)]
    w = np.array([f(dx * dx + dy * dy + dz * dz, r) for dx, dy, dz in offs], dtype=np.float64)
    return offs, w / w.sum()


def _shifted(a, off, periodic):
    """s[i] = a[i + off] (0 outside the grid unless periodic) and the in-grid mask."""
    dx, dy, dz = off
    n = a.shape
    s = np.roll(a, shift=(-dz, -dy, -dx), axis=(0, 1, 2))
    if periodic:
        return s, np.ones(n, bool)
    idx = [np.arange(k) for k in n]
    mz = (idx[0] + dz >= 0) & (idx[0] + dz < n[0])
    my = (idx[1] + dy >= 0) & (idx[1] + dy < n[1])
    mx = (idx[2] + dx >= 0) & (idx[2] + dx < n[2])
    m = mz[:, None, None] & my[None, :, None] & mx[None, None, :]
    return np.where(m, s, 0.0), m


def conv(a, r, periodic=False, kernel="spline4"):
    """Forward filter; returns (out, W) with W the in-grid weight sums."""
    a = np.asarray(a, np.float64)
    offs, w = _box(r, kernel)
    out = np.zeros_like(a)
    W = np.zeros_like(a)
    for off, wk in zip(offs, w):
        s, m = _shifted(a, off, periodic)
        out += wk * s
        W += wk * m
    return out / W, W


def conv_backward(dout, r, periodic=False, kernel="spline4"):
    dout = np.asarray(dout, np.float64)
    offs, w = _box(r, kernel)
    _, W = conv(np.zeros_like(dout), r, periodic, kernel)
    g = dout / W
    din = np.zeros_like(dout)
    for (dx, dy, dz), wk in zip(offs, w):
        s, _ = _shifted(g, (-dx, -dy, -dz), periodic)
        din += wk * s
    return din


def sgm(x, k=30.0, c=0.5):
    return 1.0 / (1.0 + np.exp(-k * (np.asarray(x, np.float64) - c)))


def sgm_backward(y, dy, k=30.0):
    y = np.asarray(y, np.float64)
    return np.asarray(dy, np.float64) * k * y * (1.0 - y)
