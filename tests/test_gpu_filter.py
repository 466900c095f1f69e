"""GPU density filter (conv with radial_convker_t<Spline4|Linear>) and sgm projection, forward and
backward, against the FP64 oracle (oracle/filter_oracle.py).  FP64 sums rounded to FP32: agreement to
2e-6 of the field scale."""
import sys

import numpy as np
import pytest

from oracle import filter_oracle as F

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hom():
    import homo3d_amd

    return homo3d_amd.Homogenization(homo3d_amd.HomoConfig(reso=24, solver="pcg"))


@pytest.mark.parametrize("kernel", ["spline4", "linear"])
@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("r", [2.0, 3.0])
def test_filter_forward_and_adjoint(hom, kernel, periodic, r):
    say("filter: n=24 %s periodic=%s r=%g" % (kernel, periodic, r))
    n = 24
    rng = np.random.default_rng(1)
    a = rng.uniform(0, 1, (n, n, n)).astype(np.float32)
    g = rng.standard_normal((n, n, n)).astype(np.float32)
    got = hom.density_filter(a, r, kernel, periodic).reshape(n, n, n)
    ref, _ = F.conv(a.astype(np.float64), r, periodic, kernel)
    assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()
    gb = hom.density_filter(g, r, kernel, periodic, backward=True).reshape(n, n, n)
    rb = F.conv_backward(g.astype(np.float64), r, periodic, kernel)
    assert np.abs(gb - rb).max() <= 2e-6 * np.abs(rb).max()


def test_projection_forward_and_backward(hom):
    n = 24
    rng = np.random.default_rng(2)
    x = rng.uniform(0.2, 0.8, n ** 3).astype(np.float32)
    y = hom.projection(x, 30.0, 0.5)
    assert np.abs(y - F.sgm(x.astype(np.float64))).max() <= 1e-6
    gy = rng.standard_normal(n ** 3).astype(np.float32)
    gx = hom.projection(gy, 30.0, 0.5, y=y)
    assert np.abs(gx - F.sgm_backward(y.astype(np.float64), gy.astype(np.float64))).max() <= 2e-6 * np.abs(gx).max()
