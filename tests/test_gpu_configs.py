"""The benchmark configurations at their real sizes on the GPU.

* 128^3 bulk-modulus maximisation (reflect6, volume 0.3, randc seed 0, OC, reference-mode MG at
  relres 1e-2 with FP16 coarse stencils): five optimisation steps through the drop-in driver
  (homo3d_framework_main, Framework.cu:137-200); the first C^H against the FP64 oracle on the same
  density; volume, bit-exact symmetry and a falling objective (up to solver noise per step).
* 256^3 negative-Poisson-ratio design (rotate3, volume 0.2, randc seed 0, FP64-refined MG-PCG to relres
  1e-5): C^H entry by entry against the FP64 oracle's (the committed fixture
  tests/golden/ch_256_rotate3_seed0.json, written by tests/golden/make_ch256.py); cubic symmetry and
  positive definiteness; the Euler identity sum_e w_e dC_ij/dw_e = C_ij (C^H is homogeneous of degree
  one in the cell stiffnesses w = Emin + x) with the GPU's own sensitivities; one MMA step of the
  driver (testAutoDiff.cu:466-493 plus a volume constraint) that keeps its volume constraint.
"""
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "ch_256_rotate3_seed0.json")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def h3d():
    import homo3d_amd

    return homo3d_amd


def run_main(args):
    """homo3d_framework_main (the `homo3d` executable's entry point) with argv = ["homo3d"] + args."""
    import ctypes

    lib = ctypes.CDLL(os.path.join(ROOT, "homo3d_amd", "libhomo3d_framework.so"))
    fn = lib.homo3d_framework_main
    fn.restype = ctypes.c_int
    argv = [b"homo3d"] + [a.encode() for a in args]
    return fn(len(argv), (ctypes.c_char_p * len(argv))(*argv))


def read_clog(path):
    rows = []
    for line in open(path):
        parts = line.split()
        rows.append(np.array([float(v) for v in parts[2:38]]).reshape(6, 6))
    return rows


def test_config2_bulk_128_oc(tmp_path, orc):
    say("config 2: 128^3 bulk, reflect6, vol 0.3, 5 OC steps, MG relres 1e-2")
    pre = str(tmp_path) + "/"
    rc = run_main(["-reso=128", "-obj=bulk", "-init=randc", "-sym=reflect6", "-vol=0.3", "-N=5", "-solver=mg",
                   "-relthres=1e-2", "-logc=1", "-logobj=1", "-prefix=" + pre])
    assert rc == 0
    rho0 = np.fromfile(pre + "initRho", dtype=np.float32)
    assert rho0.size == 128 ** 3
    Cs = read_clog(pre + "Clog")
    assert len(Cs) == 5
    say("  oracle: 128^3 MG-PCG to 1e-6 on the initial density")
    mg = orc.MG(128)
    mg.set_density(rho0.astype(np.float64) ** 3)
    Cref, _, code, its = mg.homogenize(tol=1e-6, maxit=400, mode="pcg")
    del mg
    err = np.abs(Cs[0] - Cref).max() / np.abs(Cref).max()
    say("  first C^H vs oracle: %.3e (reference stop rule relres 1e-2)" % err)
    # the reference's stopping rule (stale residual <= 1e-2); C^H is quadratic in the displacement error
    assert err < 2e-2
    objs = np.loadtxt(pre + "objlist")
    say("  objective %s" % np.array2string(objs, precision=5))
    # -bulk modulus falls; OC on relres-1e-2 solves is not a strict descent method, so single steps may
    # rise by solver noise (a few 1e-4 relative) once the design settles
    assert objs[-1] < objs[0]
    assert np.all(np.diff(objs) < 1e-3 * np.abs(objs).max())
    rho = np.fromfile(pre + "rho", dtype=np.float32)
    assert abs(rho.astype(np.float64).mean() - 0.3) < 2e-4
    assert np.array_equal(orc.symmetrize(128, rho, "reflect6"), rho)
    assert rho.min() >= 1e-3 - 1e-7 and rho.max() <= 1.0 + 1e-7


@pytest.fixture(scope="module")
def design256(h3d):
    n = 256
    say("config 3: 256^3 rotate3 vol 0.2 seed 0, MG-PCG to 1e-5")
    hom = h3d.Homogenization(h3d.HomoConfig(reso=n, solver="pcg", max_iter=400))
    rho = h3d.randc_density(n, 0, 0.2, "rotate3")
    x = (rho.astype(np.float64) ** 3).astype(np.float32)
    hom.update(x)
    st = hom.solve(1e-5, warm=False)
    C = hom.elasticMatrix(solve=False)
    say("  CG steps %s relres %s" % (st["iters"], np.array2string(np.array(st["relres"]), precision=2)))
    return hom, x, st, C


def test_config3_relres_and_ch_vs_oracle(design256):
    hom, x, st, C = design256
    assert st["status"] == 0 and max(st["relres"]) <= 1e-5
    ref = json.load(open(FIXTURE))
    assert ref["design"] == {"n": 256, "seed": 0, "sym": "rotate3", "vol": 0.2, "E": 10.0, "nu": 0.3, "emin": 1e-9,
                             "p": 3}
    Cref = np.array(ref["C"]).reshape(6, 6)
    big = np.abs(Cref) > 1e-6 * np.abs(Cref).max()
    rel = np.abs(C - Cref)[big] / np.abs(Cref)[big]
    say("  C^H per-entry relative error vs FP64 oracle: max %.2e over %d entries" % (rel.max(), big.sum()))
    assert rel.max() <= 1e-4
    assert np.abs(C - Cref)[~big].max() <= 1e-6 * np.abs(Cref).max()


def test_config3_cubic_symmetry_and_spd(design256):
    _, _, _, C = design256
    s = np.abs(C).max()
    # rotate3 (the 24 rotations of the cube) leaves a cubic tensor
    assert np.ptp(np.diag(C)[:3]) <= 1e-6 * s
    assert np.ptp(np.diag(C)[3:]) <= 1e-6 * s
    assert np.ptp([C[0, 1], C[0, 2], C[1, 2]]) <= 1e-6 * s
    mask = np.ones((6, 6), bool)
    mask[:3, :3] = False
    mask[[3, 4, 5], [3, 4, 5]] = False
    assert np.abs(C[mask]).max() <= 1e-6 * s
    assert np.allclose(C, C.T, rtol=0, atol=1e-9 * s)
    assert np.linalg.eigvalsh(C).min() > 0


@pytest.mark.parametrize("ij", [(0, 0), (0, 1), (3, 3)])
def test_config3_euler_identity(design256, ij):
    hom, x, _, C = design256
    i, j = ij
    dC = np.zeros((6, 6), np.float32)
    dC[i, j] = 1.0
    s = hom.Sensitivity(dC).astype(np.float64)
    w = x.astype(np.float64) + 1e-9
    lhs = float(np.dot(w, s))
    say("  sum_e w_e dC_%d%d/dw_e = %.8e, C_%d%d = %.8e" % (i, j, lhs, i, j, C[i, j]))
    assert abs(lhs - C[i, j]) <= 1e-4 * np.abs(C).max()


def test_config3_mma_step_keeps_volume(tmp_path):
    say("config 3: 256^3 NPR by MMA, one step through the driver")
    pre = str(tmp_path) + "/"
    rc = run_main(["-reso=256", "-obj=npr", "-optimizer=mma", "-init=randc", "-sym=rotate3", "-vol=0.2", "-N=1",
                   "-solver=pcg", "-relthres=1e-5", "-logobj=1", "-prefix=" + pre])
    assert rc == 0
    rho0 = np.fromfile(pre + "initRho", dtype=np.float32).astype(np.float64)
    rho = np.fromfile(pre + "rho", dtype=np.float32).astype(np.float64)
    say("  volume %.6f -> %.6f, max change %.3e" % (rho0.mean(), rho.mean(), np.abs(rho - rho0).max()))
    assert abs(rho0.mean() - 0.2) < 1e-3
    assert rho.mean() <= 0.2 + 1e-3
    assert rho.min() >= 1e-3 - 1e-7 and rho.max() <= 1.0 + 1e-7
    assert np.abs(rho - rho0).max() > 0
