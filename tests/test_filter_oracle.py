"""Pins of the filter oracle (oracle/filter_oracle.py): the Spline4 impulse response against the
hand-computed table, the adjoint (dot-product) test of forward and backward, constant preservation
near the non-periodic boundary, and sgm's derivative by central differences."""
import numpy as np
import pytest

from oracle import filter_oracle as F


def test_spline4_impulse_response_r2():
    # r = 2, half-width int(2.5) = 2; q = |off|^2 / 4; w = 1 - 6q + 8 q^1.5 - 3 q^2 for q < 1
    table = {}
    for d2 in range(0, 13):
        q = d2 / 4.0
        table[d2] = 1 - 6 * q + 8 * q ** 1.5 - 3 * q * q if q < 1 else 0.0
    assert table[0] == 1.0 and abs(table[1] - (1 - 1.5 + 1.0 - 0.1875)) < 1e-15 and table[4] == 0.0
    counts = {0: 1, 1: 6, 2: 12, 3: 8}  # |off|^2 < 4 inside the 5^3 box
    W = sum(counts[k] * table[k] for k in counts)
    n = 11
    a = np.zeros((n, n, n))
    a[5, 5, 5] = 1.0
    out, _ = F.conv(a, 2.0, periodic=False)
    for z in range(n):
        for y in range(n):
            for x in range(n):
                d2 = (x - 5) ** 2 + (y - 5) ** 2 + (z - 5) ** 2
                assert abs(out[z, y, x] - table.get(d2, 0.0) / W) < 1e-15


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("r", [1.5, 2.0, 3.0])
def test_backward_is_the_adjoint(periodic, r):
    rng = np.random.default_rng(3)
    a = rng.standard_normal((9, 10, 12))
    b = rng.standard_normal((9, 10, 12))
    fa, _ = F.conv(a, r, periodic)
    assert abs(np.sum(fa * b) - np.sum(a * F.conv_backward(b, r, periodic))) < 1e-12 * np.abs(fa * b).sum()


def test_constants_preserved_near_boundary():
    a = np.full((7, 8, 9), 0.37)
    out, W = F.conv(a, 2.0, periodic=False)
    assert np.allclose(out, 0.37, atol=1e-15)
    assert W.min() < 1.0 and abs(W.max() - 1.0) < 1e-14  # renormalised at faces, full weight inside


def test_sgm_derivative():
    x = np.linspace(0.3, 0.7, 41)
    y = F.sgm(x)
    h = 1e-6
    fd = (F.sgm(x + h) - F.sgm(x - h)) / (2 * h)
    assert np.allclose(F.sgm_backward(y, np.ones_like(x)), fd, rtol=1e-7, atol=1e-9)
    assert abs(F.sgm(0.5) - 0.5) < 1e-15


def test_linear_impulse_response_r2():
    table = {0: 1.0, 1: 0.5, 2: 1 - np.sqrt(2) / 2, 3: 1 - np.sqrt(3) / 2}
    W = table[0] + 6 * table[1] + 12 * table[2] + 8 * table[3]
    n = 9
    a = np.zeros((n, n, n))
    a[4, 4, 4] = 1.0
    out, _ = F.conv(a, 2.0, periodic=True, kernel="linear")
    for (z, y, x), v in np.ndenumerate(out):
        d2 = (x - 4) ** 2 + (y - 4) ** 2 + (z - 4) ** 2
        assert abs(v - table.get(d2, 0.0) / W) < 1e-15
