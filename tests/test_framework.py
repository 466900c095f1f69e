"""The C++ drivers (cfg::HomoConfig::parse, initDensity, example_opti_*, runInstance) through their
C entry point homo3d_framework_main, which the `homo3d` executable wraps."""
import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FW = os.path.join(ROOT, "homo3d_amd", "libhomo3d_framework.so")


def run_main(args):
    lib = ctypes.CDLL(FW)
    fn = lib.homo3d_framework_main
    fn.restype = ctypes.c_int
    argv = [b"homo3d"] + [a.encode() for a in args]
    arr = (ctypes.c_char_p * len(argv))(*argv)
    return fn(len(argv), arr)


def test_flag_errors_are_reported_without_gpu(tmp_path):
    pre = str(tmp_path) + "/"
    assert run_main(["-reso=32", "-obj=bogus", "-prefix=" + pre]) == 1
    assert run_main(["-nosuchflag=3", "-prefix=" + pre]) == 1
    assert run_main(["-sym", "sideways", "-prefix=" + pre]) == 1


def read_clog(path):
    rows = []
    for line in open(path):
        parts = line.split()
        rows.append(np.array([float(v) for v in parts[2:38]]).reshape(6, 6))
    return rows


@pytest.mark.gpu
def test_example_opti_bulk_oc(tmp_path, orc):
    pre = str(tmp_path) + "/"
    rc = run_main(["-reso=32", "-obj=bulk", "-init=randc", "-sym=reflect6", "-vol=0.3", "-N=6", "-solver=pcg",
                   "-relthres=1e-5", "-logc=1", "-logobj=1", "-prefix=" + pre])
    assert rc == 0
    rho0 = np.fromfile(pre + "initRho", dtype=np.float32)
    Cs = read_clog(pre + "Clog")
    assert len(Cs) == 6
    mg = orc.MG(32)
    mg.set_density(rho0.astype(np.float64) ** 3)
    Cref, _, _, _ = mg.homogenize(tol=1e-10, maxit=400, mode="pcg")
    assert np.abs(Cs[0] - Cref).max() <= 1e-4 * np.abs(Cref).max()
    objs = np.loadtxt(pre + "objlist")
    assert objs[-1] < objs[0]  # bulk modulus grows (objective is its negative)
    rho = np.fromfile(pre + "rho", dtype=np.float32)
    assert abs(rho.astype(np.float64).mean() - 0.3) < 2e-4
    assert np.array_equal(orc.symmetrize(32, rho, "reflect6"), rho)
    C = np.fromfile(pre + "C", dtype=np.float32).reshape(6, 6)
    assert np.allclose(C, C.T, rtol=1e-5, atol=1e-7)


@pytest.mark.gpu
def test_example_opti_npr_mma_and_shear(tmp_path):
    pre = str(tmp_path) + "/npr_"
    rc = run_main(["-reso=32", "-obj=npr", "-optimizer=mma", "-sym=rotate3", "-vol=0.2", "-N=4", "-solver=pcg",
                   "-relthres=1e-5", "-prefix=" + pre])
    assert rc == 0
    rho = np.fromfile(pre + "rho", dtype=np.float32)
    assert rho.min() >= 1e-3 - 1e-7 and rho.max() <= 1.0 + 1e-7
    pre2 = str(tmp_path) + "/shear_"
    rc = run_main(["-reso=32", "-obj=shear", "-sym=reflect6", "-vol=0.3", "-N=3", "-solver=pcg", "-relthres=1e-4",
                   "-prefix=" + pre2])
    assert rc == 0
    assert np.isfinite(np.loadtxt(pre2 + "objlist")).all()
