"""GPU tests of the C ABI's reference-shaped entry points beyond the operators themselves:
MG::v_cycle(w_SOR, pre, post), Sensitivity(..., lexiOrder=false), MMAOptimizer::setBound(float*, float*),
the divergence status of the refined solve and warm starts after level-0 writes."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

E, NU = 10.0, 0.3


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def h3d():
    import homo3d_amd

    return homo3d_amd


def rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)).max() /
                 max(np.abs(np.asarray(b, np.float64)).max(), 1e-300))


def make(h3d, n, solver="mg", seed=0, vol=0.3, sym="reflect6", **kw):
    hom = h3d.Homogenization(h3d.HomoConfig(reso=n, solver=solver, coarse_half=False, **kw))
    rho = h3d.randc_density(n, seed, vol, sym)
    x = (rho.astype(np.float64) ** 3).astype(np.float32)
    hom.update(x)
    return hom, x


@pytest.mark.parametrize("w,pre,post", [(1.0, 1, 1), (0.8, 2, 1), (1.0, 0, 2)])
def test_v_cycle_sweep_counts(h3d, w, pre, post):
    """homo3d_vcycle_sweeps equals the same cycle composed from the level entry points, each of which
    has its own oracle parity test (smooth, residual, restrict, coarse_solve, prolong)."""
    say("v_cycle: n=32 w=%g pre=%d post=%d" % (w, pre, post))
    hom, _ = make(h3d, 32, seed=1)
    hom.macro_rhs()
    zero0 = np.zeros(hom.level_shape(0), np.float32)
    hom.set_field(0, 0, zero0)
    hom.v_cycle(w, pre, post)
    got = hom.get_field(0, 0)
    # the same cycle from the level operations
    hom.set_field(0, 0, zero0)
    nl = len(hom.sizes)
    for l in range(nl - 1):
        if l > 0:
            hom.set_field(l, 0, np.zeros(hom.level_shape(l), np.float32))
        for _ in range(pre):
            hom.smooth(l, w)
        hom.residual(l)
        hom.restrict(l + 1)
    hom.coarse_solve()
    for l in range(nl - 2, -1, -1):
        hom.prolong(l)
        for _ in range(post):
            hom.smooth(l, w)
    ref = hom.get_field(0, 0)
    assert rel(got, ref) < 1e-6
    if (w, pre, post) == (1.0, 1, 1):
        hom.set_field(0, 0, zero0)
        hom.vcycle(False)
        assert rel(hom.get_field(0, 0), got) < 1e-6


def test_sensitivity_storage_order(h3d):
    say("sensitivity order: n=16")
    hom, _ = make(h3d, 16, solver="pcg")
    hom.solve(1e-5, warm=False)
    dC = np.zeros((6, 6), np.float32)
    dC[0, 1] = dC[1, 0] = 1.0
    dC[3, 3] = -0.5
    lex = hom.Sensitivity(dC, lexiOrder=True)
    blk = hom.Sensitivity(dC, lexiOrder=False)
    n, h = 16, 8
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    b = (x & 1) | ((y & 1) << 1) | ((z & 1) << 2)
    sidx = b * h ** 3 + (x >> 1) + h * ((y >> 1) + h * (z >> 1))
    assert np.array_equal(blk[sidx.ravel()], lex)


def test_mma_per_variable_bounds(h3d, orc):
    say("mma bounds: n=64 m=1")
    n, m = 64, 1
    rng = np.random.default_rng(3)
    x = rng.uniform(0.3, 0.6, n)
    lo = rng.uniform(0.01, 0.2, n).astype(np.float32)
    hi = rng.uniform(0.7, 1.0, n).astype(np.float32)
    gpu = h3d.MMAOptimizer(m, n, 1.0, 0.0, 1000.0, 1.0)
    gpu.setBound(lo, hi)
    st = {"x": x.copy(), "xmin": lo.astype(np.float64), "xmax": hi.astype(np.float64), "xold1": x.copy(),
          "xold2": x.copy(), "low": np.zeros(n), "upp": np.zeros(n)}
    w = rng.uniform(0.5, 2.0, (m, n))
    for it in range(4):
        xin = st["x"].astype(np.float32).astype(np.float64)
        st["x"][:] = xin
        df0 = (-(1.0 + np.sin(3 * xin))).astype(np.float32)
        g = np.array([(w[0] * xin).sum() / n - 0.4], dtype=np.float32)
        dg = (w / n).astype(np.float32)
        xg = gpu.update(it, xin.astype(np.float32), df0, g, list(dg))
        orc.mma_update(st, it, df0.astype(np.float64), g.astype(np.float64), dg.astype(np.float64), 1.0,
                       np.zeros(m), np.full(m, 1000.0), np.ones(m), 1.0)
        x64, low, upp, _ = gpu.state()
        assert np.abs(x64 - st["x"]).max() <= 1e-9, it
        assert np.all(xg >= lo) and np.all(xg <= hi)


def test_nonfinite_density_reports_divergence(h3d):
    """A NaN in the density propagates into every residual norm: the refined solve must stop with
    HOMO3D_DIVERGED instead of running to its iteration limit and returning NOT_CONVERGED."""
    say("divergence: n=16 NaN density")
    hom = h3d.Homogenization(h3d.HomoConfig(reso=16, solver="pcg", max_iter=50))
    x = (h3d.randc_density(16, 0, 0.3, "reflect6").astype(np.float64) ** 3).astype(np.float32)
    x[123] = np.nan
    with pytest.raises(h3d.Homo3dError) as ei:
        hom.update(x)
        hom.solve(1e-5, warm=False)
    assert ei.value.code == h3d.HOMO3D_DIVERGED


def test_warm_start_after_level0_prolongation(h3d):
    """A warm refined solve after homo3d_prolong wrote level 0 starts from that field (the handle widens
    it into its FP64 iterate) rather than from the stale FP64 solution of the previous solve."""
    say("warm start after prolongation: n=32")
    hom, _ = make(h3d, 32, solver="pcg", seed=2)
    hom.solve(1e-5, warm=False)
    C0 = hom.elasticMatrix(solve=False)
    st = hom.solve(1e-5, warm=True)
    assert sum(st["iters"]) == 0  # converged iterate: nothing to do
    rng = np.random.default_rng(0)
    hom.set_field(1, 0, rng.standard_normal(hom.level_shape(1)).astype(np.float32))
    hom.prolong(0)  # u[0] += P u[1]: a level-0 write that is not a field upload
    st = hom.solve(1e-5, warm=True)
    assert st["status"] == 0
    assert sum(st["iters"]) > 0  # it started from the perturbed field
    # u itself is barely determined in the near-void cells (stiffness ~1e-9), C^H is not
    assert rel(hom.elasticMatrix(solve=False), C0) < 1e-4


def test_warm_start_worse_than_zero_falls_back(h3d):
    """A warm start whose FP64 residual is not below |f| is dropped for that load case: with every load
    case falling back the solve is the cold one (same CG steps, same C^H); with some of them, those
    restart from zero while the others keep their start."""
    say("warm-start fallback: n=32")
    hom, _ = make(h3d, 32, solver="pcg", seed=5)
    st0 = hom.solve(1e-5, warm=False)
    C0 = hom.elasticMatrix(solve=False)
    u = hom.get_solution64()
    rng = np.random.default_rng(1)
    bad = rng.standard_normal(u.shape) * 100.0 * np.abs(u).max()
    hom.set_solution64(bad)
    st1 = hom.solve(1e-5, warm=True)
    assert st1["status"] == 0 and list(st1["iters"]) == list(st0["iters"]), (st0["iters"], st1["iters"])
    assert rel(hom.elasticMatrix(solve=False), C0) < 1e-6
    mixed = u.copy()
    mixed[3:] = bad[3:]  # load cases 3..5 bad, 0..2 at the converged solution
    hom.set_solution64(mixed)
    st2 = hom.solve(1e-5, warm=True)
    assert st2["status"] == 0
    assert list(st2["iters"])[3:] == list(st0["iters"])[3:]
    assert max(st2["iters"][:3]) == 0
    assert rel(hom.elasticMatrix(solve=False), C0) < 1e-4


def test_pitched_padding_is_left_alone(h3d):
    """Every entry point that takes a row-pitched array writes only the n live columns: the caller's
    padding (pitch > n) keeps its values, on host and on device arrays, after earlier calls have left
    other data in the library's staging memory.  The live columns equal the pitch = n results."""
    import ctypes

    n, pitch, sentinel = 16, 24, np.float32(7.25)
    L = h3d.lib()
    hom, x = make(h3d, n, solver="pcg")
    hom.solve(1e-5, warm=False)
    rng = np.random.default_rng(3)
    rho = (0.2 + 0.6 * rng.random((n, n, n))).astype(np.float32)
    sens = (-rng.random((n, n, n))).astype(np.float32)
    dC = np.eye(6, dtype=np.float32).ravel()
    dCp = dC.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    def pad(a):
        p = np.full((n, n, pitch), sentinel, np.float32)
        p[:, :, :n] = a
        return p

    def calls(pt):
        """name -> function(args dict of pitched arrays as pointers) for pitch pt."""
        return {
            "oc_update": lambda a, dev: L.homo3d_oc_update(hom._h, a["sens"], a["rho"], pt, 0.5, 1e-3, 0.05, 0.5, dev),
            "oc_filter": lambda a, dev: L.homo3d_oc_filter(hom._h, a["sens"], a["rho"], pt, 2.0, 0, dev),
            "filter": lambda a, dev: L.homo3d_density_filter(hom._h, a["rho"], a["out"], pt, 2.0, 0, 0, 0, dev),
            "filter_back": lambda a, dev: L.homo3d_density_filter(hom._h, a["sens"], a["out"], pt, 2.0, 0, 0, 1, dev),
            "sgm": lambda a, dev: L.homo3d_projection(hom._h, a["rho"], None, a["out"], pt, 30.0, 0.5, 0, dev),
            "sgm_back": lambda a, dev: L.homo3d_projection(hom._h, a["sens"], a["rho"], a["out"], pt, 30.0, 0.5, 1,
                                                           dev),
            "symmetrize": lambda a, dev: L.homo3d_symmetrize(hom._h, a["rho"], pt, 2, dev),
            "sensitivity": lambda a, dev: L.homo3d_sensitivity(hom._h, dCp, a["out"], pt, dev),
        }

    outputs = {"oc_update": "rho", "oc_filter": "sens", "filter": "out", "filter_back": "out", "sgm": "out",
               "sgm_back": "out", "symmetrize": "rho", "sensitivity": "out"}
    for name in outputs:
        # the unpitched result
        flat = {"rho": rho.copy(), "sens": sens.copy(), "out": np.zeros_like(rho)}
        assert calls(n)[name]({k: v.ctypes.data_as(ctypes.c_void_p) for k, v in flat.items()}, 0) == 0, name
        want = flat[outputs[name]]
        for dev in (0, 1):
            arrs = {"rho": pad(rho), "sens": pad(sens), "out": np.full((n, n, pitch), sentinel, np.float32)}
            if dev:
                bufs = {k: h3d.DeviceBuffer.from_array(v) for k, v in arrs.items()}
                assert calls(pitch)[name]({k: b.value for k, b in bufs.items()}, 1) == 0, name
                got = bufs[outputs[name]].download(np.float32).reshape(n, n, pitch)
            else:
                assert calls(pitch)[name]({k: v.ctypes.data_as(ctypes.c_void_p) for k, v in arrs.items()}, 0) == 0
                got = arrs[outputs[name]]
            say("  %s on_device=%d: padding %s" % (name, dev, "kept" if np.all(got[:, :, n:] == sentinel) else "WRITTEN"))
            assert np.all(got[:, :, n:] == sentinel), (name, dev)
            assert np.array_equal(got[:, :, :n], want), (name, dev)
    hom.close()


@pytest.mark.parametrize("n", [128, 256])
def test_staged_cg_pass_matches_gather_form(h3d, n):
    """The CG pass with each workgroup's z node box staged in LDS does the same arithmetic in the same
    order as the pass that gathers the neighbours from memory: the refined solves agree bit for bit
    (CG steps, FP64 iterate, C^H).  n = 256 has two tiles per grid row, so it also covers the tile whose
    halo wraps around x = n."""
    res = {}
    for staged in (False, True):
        hom, _ = make(h3d, n, solver="pcg", seed=4, vol=0.25, sym="rotate3")
        hom.stage_tiles(staged)
        st = hom.solve(1e-5, warm=False)
        res[staged] = (st["iters"], hom.get_solution64(), hom.elasticMatrix(solve=False))
        say("  n=%d staged=%d CG steps %s relres %.2e" % (n, staged, st["iters"], max(st["relres"])))
        assert st["status"] == 0 and max(st["relres"]) <= 1e-5
        hom.close()
    assert res[True][0] == res[False][0]
    assert np.array_equal(res[True][1], res[False][1])
    assert np.array_equal(res[True][2], res[False][2])


def test_density_maps_and_device_memory(h3d):
    """x = rho^p, drho = p rho^(p-1) dx, the FP64 field sum and fill (the element maps the optimisation
    drivers keep on the device) against NumPy, on host and device arrays with row padding; the C ABI's
    device allocation and copies round-trip."""
    import ctypes

    n, pitch = 20, 24
    L = h3d.lib()
    rng = np.random.default_rng(11)
    rho = np.zeros((n, n, pitch), np.float32)
    rho[:, :, :n] = (0.001 + rng.random((n, n, n))).astype(np.float32)
    dx = np.zeros_like(rho)
    dx[:, :, :n] = rng.standard_normal((n, n, n)).astype(np.float32)
    ops = h3d.Homogenization(h3d.HomoConfig(reso=n, solver="none"))
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    r, d = rho[:, :, :n], dx[:, :, :n]
    want = {(3.0, "pow"): r * r * r, (3.0, "grad"): 3.0 * r * r * d,
            (2.5, "pow"): np.power(r, np.float32(2.5)), (2.5, "grad"): np.float32(2.5) * np.power(r, np.float32(1.5)) * d}
    for dev in (0, 1):
        for p in (3.0, 2.5):
            x = np.full_like(rho, 7.0)
            g = np.full_like(rho, 7.0)
            if dev:
                bufs = [h3d.DeviceBuffer.from_array(a) for a in (rho, dx, x, g)]
                assert L.homo3d_density_pow(ops._h, bufs[0].value, bufs[2].value, pitch, p, 1) == 0
                assert L.homo3d_density_pow_grad(ops._h, bufs[0].value, bufs[1].value, bufs[3].value, pitch, p, 1) == 0
                x = bufs[2].download(np.float32).reshape(rho.shape)
                g = bufs[3].download(np.float32).reshape(rho.shape)
            else:
                assert L.homo3d_density_pow(ops._h, vp(rho), vp(x), pitch, p, 0) == 0
                assert L.homo3d_density_pow_grad(ops._h, vp(rho), vp(dx), vp(g), pitch, p, 0) == 0
            tol = 0 if p == 3.0 else 4e-7
            assert np.all(x[:, :, n:] == 7.0) and np.all(g[:, :, n:] == 7.0)
            assert np.abs(x[:, :, :n] - want[(p, "pow")]).max() <= tol * np.abs(want[(p, "pow")]).max()
            assert np.abs(g[:, :, :n] - want[(p, "grad")]).max() <= tol * np.abs(want[(p, "grad")]).max()
        s = ctypes.c_double()
        src = h3d.DeviceBuffer.from_array(rho) if dev else None
        assert L.homo3d_field_sum(ops._h, src.value if dev else vp(rho), pitch, ctypes.byref(s), dev) == 0
        assert abs(s.value - r.astype(np.float64).sum()) <= 1e-9 * r.sum()
        f = np.full_like(rho, 7.0)
        if dev:
            fb = h3d.DeviceBuffer.from_array(f)
            assert L.homo3d_field_fill(ops._h, fb.value, pitch, ctypes.c_float(0.25), 1) == 0
            f = fb.download(np.float32).reshape(rho.shape)
        else:
            assert L.homo3d_field_fill(ops._h, vp(f), pitch, ctypes.c_float(0.25), 0) == 0
        assert np.all(f[:, :, :n] == 0.25) and np.all(f[:, :, n:] == 7.0)
        say("  density maps on_device=%d ok" % dev)
    # device memory through the C ABI
    ptr = ctypes.c_void_p()
    a = rng.standard_normal(1000).astype(np.float32)
    b = np.zeros_like(a)
    assert L.homo3d_device_alloc(ctypes.byref(ptr), a.nbytes) == 0 and ptr.value
    ptr2 = ctypes.c_void_p()
    assert L.homo3d_device_alloc(ctypes.byref(ptr2), a.nbytes) == 0
    assert L.homo3d_device_copy(ptr, vp(a), a.nbytes, 0) == 0
    assert L.homo3d_device_copy(ptr2, ptr, a.nbytes, 2) == 0
    assert L.homo3d_device_copy(vp(b), ptr2, a.nbytes, 1) == 0
    assert np.array_equal(a, b)
    assert L.homo3d_device_copy(ptr2, ptr, a.nbytes, 3) == 3  # bad kind
    assert L.homo3d_device_free(ptr) == 0 and L.homo3d_device_free(ptr2) == 0
    ops.close()
