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
def test_example_opti_npr_mma_and_shear(tmp_path, orc):
    pre = str(tmp_path) + "/npr_"
    rc = run_main(["-reso=32", "-obj=npr", "-optimizer=mma", "-sym=rotate3", "-vol=0.2", "-N=4", "-solver=pcg",
                   "-relthres=1e-5", "-prefix=" + pre])
    assert rc == 0
    rho = np.fromfile(pre + "rho", dtype=np.float32)
    assert rho.min() >= 1e-3 - 1e-7 and rho.max() <= 1.0 + 1e-7
    pre2 = str(tmp_path) + "/shear_"
    rc = run_main(["-reso=32", "-obj=shear", "-sym=reflect6", "-vol=0.3", "-N=3", "-solver=pcg", "-relthres=1e-6",
                   "-logc=1", "-prefix=" + pre2])
    assert rc == 0
    assert np.isfinite(np.loadtxt(pre2 + "objlist")).all()
    # the solved density is sgm(conv(rho, Spline4(2)))^3 of the design (Framework.cu:288): its C^H
    # from the FP64 oracle on that pipeline equals the driver's first C^H
    from oracle import filter_oracle as F

    rho0 = np.fromfile(pre2 + "initRho", dtype=np.float32).astype(np.float64).reshape(32, 32, 32)
    rhat = F.sgm(F.conv(rho0, 2.0, periodic=False)[0])
    mg = orc.MG(32)
    mg.set_density((rhat ** 3).ravel())
    Cref, _, _, _ = mg.homogenize(tol=1e-10, maxit=400, mode="pcg")
    C0 = read_clog(pre2 + "Clog")[0]
    assert np.abs(C0 - Cref).max() <= 1e-4 * np.abs(Cref).max()


def test_shear_isotropy_terms_follow_reference_scaling():
    """example_opti_shear_isotropy's constraint (Framework.cu:317-323): the value is
    1000 * (1000 (A/B - 1)^2) - 0.1, its gradient that of 1000 (A/B - 1)^2, with
    A = 2 (C33 + C44 + C55), B = C00 + C11 + C22 - C01 - C02 - C12."""
    lib = ctypes.CDLL(FW)
    fn = lib.homo3d_shear_isotropy_terms
    fn.restype = ctypes.c_int
    rng = np.random.default_rng(5)
    C = np.diag([5.0, 5.2, 4.9, 1.1, 1.3, 1.2])
    C[0, 1] = C[1, 0] = 2.0
    C[0, 2] = C[2, 0] = 2.1
    C[1, 2] = C[2, 1] = 1.9
    C += 0.01 * rng.standard_normal((6, 6))
    Cf = np.ascontiguousarray(C.ravel())
    obj, g = ctypes.c_double(), ctypes.c_double()
    dobj = (ctypes.c_float * 36)()
    dg = (ctypes.c_float * 36)()
    assert fn(Cf.ctypes.data_as(ctypes.c_void_p), ctypes.byref(obj), dobj, ctypes.byref(g), dg) == 0

    def terms(M):
        A = 2 * (M[3, 3] + M[4, 4] + M[5, 5])
        B = M[0, 0] + M[1, 1] + M[2, 2] - M[0, 1] - M[0, 2] - M[1, 2]
        return -(M[3, 3] + M[4, 4] + M[5, 5]) / 3, 1000.0 * (A / B - 1) ** 2

    o_ref, k_ref = terms(C)
    assert abs(obj.value - o_ref) < 1e-12
    assert abs(g.value - (k_ref * 1000.0 - 0.1)) < 1e-9 * max(1.0, abs(k_ref * 1000))
    # gradients by central differences of the back-propagated expressions, entry by entry
    for idx in range(36):
        i, j = divmod(idx, 6)
        h = 1e-6
        Cp, Cm = C.copy(), C.copy()
        Cp[i, j] += h
        Cm[i, j] -= h
        do = (terms(Cp)[0] - terms(Cm)[0]) / (2 * h)
        dk = (terms(Cp)[1] - terms(Cm)[1]) / (2 * h)
        assert abs(dobj[idx] - do) < 1e-6, (i, j)
        assert abs(dg[idx] - dk) < 1e-4 * max(1.0, abs(dk)), (i, j, dg[idx], dk)
