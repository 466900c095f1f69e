"""GPU parity: every operation of the HIP path, called through the C ABI, against the FP64 oracle
on the same seeded inputs.

Tolerances (written per test) follow from FP32 storage and FP32 accumulation against FP64:
single operator applications agree to ~1e-6 relative to the operand scale; iterates after a
sweep or a V-cycle to ~1e-5; converged quantities (C^H) to the solve tolerance; integer-like and
projection outputs bit-exactly.
"""
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


def make(h3d, n, solver="mg", seed=0, vol=0.3, sym="reflect6", coarse_half=False):
    """Handle on a seeded randc density.  The operator-level parity tests use FP32 coarse stencils
    (coarse_half=False), the exact counterpart of the oracle; the FP16 stencils of the default
    configuration have their own test."""
    hom = h3d.Homogenization(h3d.HomoConfig(reso=n, solver=solver, coarse_half=coarse_half))
    rho = h3d.randc_density(n, seed, vol, sym)
    x = (rho.astype(np.float64) ** 3).astype(np.float32)
    hom.update(x)
    return hom, x


@pytest.fixture(scope="module")
def case16(h3d, orc):
    say("case16: n=16 randc seed 0")
    hom, x = make(h3d, 16)
    mg = orc.MG(16, E, NU)
    mg.set_density(x.astype(np.float64))
    return hom, mg, x


def test_fine_operator_and_residual(case16, orc):
    hom, mg, x = case16
    rng = np.random.default_rng(0)
    u = rng.standard_normal(hom.level_shape(0)).astype(np.float32)
    f = rng.standard_normal(hom.level_shape(0)).astype(np.float32)
    hom.set_field(0, 0, u)
    hom.set_field(0, 1, f)
    hom.apply_K()
    y = hom.get_field(0, 2)
    hom.residual(0)
    r = hom.get_field(0, 2)
    for s in range(6):
        ky = orc.apply_K(16, mg.KE, mg.w, u[s].astype(np.float64), pin=True)
        assert rel(y[s], ky) < 2e-6
        rr = orc.residual(16, mg.KE, mg.w, u[s].astype(np.float64), f[s].astype(np.float64))
        assert rel(r[s], rr) < 2e-6


def test_macro_rhs(case16, orc):
    hom, mg, x = case16
    hom.macro_rhs()
    f = hom.get_field(0, 1)
    for s in range(6):
        assert rel(f[s], orc.macro_rhs(16, mg.KE, mg.w, s)) < 2e-6


@pytest.mark.parametrize("omega,reverse", [(1.0, False), (1.0, True), (1.4, False)])
def test_fine_gauss_seidel_sweep(case16, orc, omega, reverse):
    hom, mg, x = case16
    rng = np.random.default_rng(1)
    u = rng.standard_normal(hom.level_shape(0)).astype(np.float32) * 0.1
    f = rng.standard_normal(hom.level_shape(0)).astype(np.float32)
    hom.set_field(0, 0, u)
    hom.set_field(0, 1, f)
    hom.smooth(0, omega, reverse)
    got = hom.get_field(0, 0)
    for s in range(6):
        ref = orc.gs_fine(16, mg.KE, mg.w, u[s].astype(np.float64), f[s].astype(np.float64), omega, reverse)
        assert rel(got[s], ref) < 2e-5


def fp16_case(h3d, orc, n, seed, mfma=True, nu=NU):
    hom = h3d.Homogenization(h3d.HomoConfig(reso=n, solver="pcg", coarse_half=False, fine_mfma=mfma,
                                            poissonRatio=nu))
    rho = h3d.randc_density(n, seed, 0.3, "reflect6")
    x = (rho.astype(np.float64) ** 3).astype(np.float32)
    hom.update(x)
    mg = orc.MG(n, E, nu)
    mg.set_density(x.astype(np.float64))
    return hom, mg


@pytest.mark.parametrize("n,mfma,nu", [(32, True, NU), (64, True, NU), (32, True, 0.35), (32, False, NU)])
def test_fine_operator_fp16(h3d, orc, n, mfma, nu):
    """Packed-FP16 finest-level operator: element form on the matrix cores (integer Lame tables, exact in
    FP16: one combined table when lambda/mu is a short binary fraction, as for nu = 0.3, two tables
    otherwise, as for nu = 0.35; exact products, FP32 sums) or the FP32 row form reading FP16 u.  Both
    against the FP64 oracle on the same FP16-rounded u, to the FP32-field tolerance."""
    say("fp16 operator: n=%d mfma=%s nu=%g" % (n, mfma, nu))
    hom, mg = fp16_case(h3d, orc, n, 6, mfma, nu)
    rng = np.random.default_rng(n)
    u = rng.standard_normal(hom.level_shape(0)).astype(np.float32)
    f = rng.standard_normal(hom.level_shape(0)).astype(np.float32)
    u16 = u.astype(np.float16).astype(np.float64)
    hom.set_field(0, 0, u)
    hom.set_field(0, 1, f)
    hom.apply_K16(residual=False)
    y = hom.get_field(0, 2)
    hom.apply_K16(residual=True)
    r = hom.get_field(0, 2)
    for s in range(6):
        ky = orc.apply_K(n, mg.KE, mg.w, u16[s], pin=True)
        rr = orc.residual(n, mg.KE, mg.w, u16[s], f[s].astype(np.float64))
        say("  case %d: K u %.2e, f - K u %.2e" % (s, rel(y[s], ky), rel(r[s], rr)))
        assert rel(y[s], ky) < 2e-6
        assert rel(r[s], rr) < 2e-6


@pytest.mark.parametrize("n,reverse", [(32, False), (32, True), (24, True)])
def test_fine_gauss_seidel_fp16(h3d, orc, n, reverse):
    """One sweep with FP16 storage of u (colour pairs at n = 32, one colour per launch at n = 24)
    against the FP64 sweep from the same FP16-rounded start.  Every colour's new values are rounded to
    FP16 before later colours read them, so the tolerance is FP16's (2^-11) with some propagation."""
    say("fp16 sweep: n=%d reverse=%s" % (n, reverse))
    hom, mg = fp16_case(h3d, orc, n, 7)
    rng = np.random.default_rng(n + 1)
    u = (rng.standard_normal(hom.level_shape(0)) * 0.1).astype(np.float32)
    f = rng.standard_normal(hom.level_shape(0)).astype(np.float32)
    u16 = u.astype(np.float16).astype(np.float64)
    hom.set_field(0, 0, u)
    hom.set_field(0, 1, f)
    hom.smooth16(1.0, reverse)
    got = hom.get_field(0, 0)
    for s in range(6):
        ref = orc.gs_fine(n, mg.KE, mg.w, u16[s], f[s].astype(np.float64), 1.0, reverse)
        say("  case %d: %.2e" % (s, rel(got[s], ref)))
        assert rel(got[s], ref) < 3e-3


@pytest.mark.parametrize("n,omega", [(32, 1.0), (64, 1.0), (32, 0.8)])
def test_presmooth_residual_fp16(h3d, orc, n, omega):
    """The preconditioner's level-0 pre-smoothing: a forward sweep from u = 0 whose colour passes leave
    their own-colour residuals behind, completed by the later-colour terms.  u against the FP64 sweep
    from zero (FP16 storage tolerance, as for the sweep test; whatever u held before is ignored), and r
    against the FP64 residual f - K u of the returned FP16 u, to the tolerance of the full residual pass
    on the same u (the same terms, summed in another order)."""
    say("presmooth residual: n=%d omega=%g" % (n, omega))
    hom, mg = fp16_case(h3d, orc, n, 11, mfma=False)
    rng = np.random.default_rng(n + 5)
    # loads small enough that the sweep's u (large in the near-void cells) stays within FP16 range
    f = (rng.standard_normal(hom.level_shape(0)) * 1e-3).astype(np.float32)
    hom.set_field(0, 1, f)
    hom.set_field(0, 0, rng.standard_normal(hom.level_shape(0)).astype(np.float32))  # ignored
    hom.presmooth16(omega)
    u = hom.get_field(0, 0)
    r = hom.get_field(0, 2)
    hom.apply_K16(residual=True)  # full residual pass on the same u
    r_full = hom.get_field(0, 2)
    zero = np.zeros(u.shape[1:], np.float64)
    for s in range(6):
        ref = orc.gs_fine(n, mg.KE, mg.w, zero, f[s].astype(np.float64), omega, False)
        rr = orc.residual(n, mg.KE, mg.w, u[s].astype(np.float64), f[s].astype(np.float64))
        say("  case %d: u %.2e, r %.2e, full pass %.2e" % (s, rel(u[s], ref), rel(r[s], rr), rel(r_full[s], rr)))
        assert rel(u[s], ref) < 3e-3
        assert rel(r_full[s], rr) < 2e-6
        assert rel(r[s], rr) < 2e-6


@pytest.mark.parametrize("n,reverse", [(32, False), (32, True), (128, False), (128, True)])
def test_coarse_gauss_seidel_sweep(h3d, orc, n, reverse):
    """One level-1 sweep against the FP64 oracle sweep on the same stored stencil: 16^3 runs one colour
    and one load case per thread, 64^3 colour pairs (whole grid rows per workgroup) with six."""
    say("coarse sweep: n=%d reverse=%s" % (n, reverse))
    hom, x = make(h3d, n, seed=3)
    nc = hom.sizes[1]
    st = hom.stencil(1)
    rng = np.random.default_rng(n + 9)
    u = rng.standard_normal(hom.level_shape(1)).astype(np.float32)
    f = rng.standard_normal(hom.level_shape(1)).astype(np.float32)
    hom.set_field(1, 0, u)
    hom.set_field(1, 1, f)
    hom.smooth(1, 1.0, reverse)
    got = hom.get_field(1, 0)
    for s in range(6):
        ref = orc.stencil_gs(nc, st, u[s].astype(np.float64), f[s].astype(np.float64), 1.0, reverse)
        say("  case %d: %.2e" % (s, rel(got[s], ref)))
        assert rel(got[s], ref) < 1e-5


@pytest.mark.parametrize("omega", [1.0, 0.8])
def test_coarse_presmooth_residual(h3d, orc, omega):
    """A coarse level's zero-start pre-smoothing (64^3, colour pairs, six load cases per thread) with
    the residual formed from its own-colour residuals plus the later-colour terms: u against the FP64
    oracle sweep from zero and r against the FP64 stencil residual of the returned u, both on the same
    stored stencil."""
    say("coarse presmooth residual: level 1 of 128^3, omega=%g" % omega)
    hom, x = make(h3d, 128, seed=5)
    nc = hom.sizes[1]
    st = hom.stencil(1)
    rng = np.random.default_rng(17)
    f = rng.standard_normal(hom.level_shape(1)).astype(np.float32)
    hom.set_field(1, 1, f)
    hom.set_field(1, 0, rng.standard_normal(hom.level_shape(1)).astype(np.float32))  # ignored
    hom.presmooth(1, omega)
    u = hom.get_field(1, 0)
    r = hom.get_field(1, 2)
    zero = np.zeros(u.shape[1:], np.float64)
    for s in range(6):
        ref = orc.stencil_gs(nc, st, zero, f[s].astype(np.float64), omega, False)
        rr = orc.stencil_residual(nc, st, u[s].astype(np.float64), f[s].astype(np.float64))
        say("  case %d: u %.2e, r %.2e" % (s, rel(u[s], ref), rel(r[s], rr)))
        assert rel(u[s], ref) < 1e-5
        assert rel(r[s], rr) < 1e-5


def test_galerkin_stencils(h3d, orc):
    say("galerkin: n=32")
    hom, x = make(h3d, 32, seed=4)
    mg = orc.MG(32, E, NU)
    mg.set_density(x.astype(np.float64))
    for level in range(1, len(hom.sizes)):
        got = hom.stencil(level)
        ref = mg.stencil(level)
        assert rel(got, ref) < 1e-5, level


def test_galerkin_stencils_fp16(h3d, orc):
    """FP16 storage of the intermediate levels' stencils (the default, as in the reference): each
    level agrees with the FP64 Galerkin product to FP16 rounding compounded over the levels above it
    (2^-11 relative per level, scale = the level's largest entry); the coarsest stays FP32."""
    say("galerkin fp16: n=32")
    hom, x = make(h3d, 32, seed=4, coarse_half=True)
    mg = orc.MG(32, E, NU)
    mg.set_density(x.astype(np.float64))
    nl = len(hom.sizes)
    for level in range(1, nl):
        got = hom.stencil(level)
        ref = mg.stencil(level)
        assert rel(got, ref) < 2.0 ** -11 * (level + 1), level
        if level < nl - 1:  # values are FP16-representable
            assert np.array_equal(got.astype(np.float16).astype(np.float32), got), level


def test_transfers_and_coarse_solve(case16, orc):
    hom, mg, x = case16
    rng = np.random.default_rng(2)
    r0 = rng.standard_normal(hom.level_shape(0)).astype(np.float32)
    # residual field of level 0 is the restriction input: write it through u/f/residual
    hom.set_field(0, 0, np.zeros(hom.level_shape(0), np.float32))
    hom.set_field(0, 1, r0)
    hom.residual(0)  # r = f - K*0 = f (pinned row zeroed)
    r0p = hom.get_field(0, 2)
    hom.restrict(1)
    fc = hom.get_field(1, 1)
    for s in range(6):
        assert rel(fc[s], orc.restrict(16, r0p[s].astype(np.float64))) < 1e-6
    uc = rng.standard_normal(hom.level_shape(1)).astype(np.float32)
    uf = rng.standard_normal(hom.level_shape(0)).astype(np.float32)
    hom.set_field(1, 0, uc)
    hom.set_field(0, 0, uf)
    hom.prolong(0)
    got = hom.get_field(0, 0)
    for s in range(6):
        ref = orc.prolong_add(16, uc[s].astype(np.float64), uf[s].astype(np.float64), zero_pin=True)
        assert rel(got[s], ref) < 1e-6
    # coarsest solve against the oracle's dense Cholesky of the same Galerkin operator
    lc = len(hom.sizes) - 1
    nc = hom.sizes[lc]
    fcs = rng.standard_normal(hom.level_shape(lc)).astype(np.float32)
    hom.set_field(lc, 1, fcs)
    hom.coarse_solve()
    uc = hom.get_field(lc, 0)
    A = orc.dense_from_stencil(nc, mg.stencil(lc))
    for s in range(6):
        ref = np.linalg.solve(A, fcs[s].astype(np.float64).T.reshape(-1)).reshape(-1, 3).T
        assert rel(uc[s], ref) < 1e-4


@pytest.mark.parametrize("n", [20, 24, 28])
def test_coarse_solve_every_coarsest_size(h3d, orc, n):
    """The coarsest level's inverse is assembled, Cholesky-factorised and inverted on the GPU at the
    density update.  Coarsest grids are 4^3 to 7^3 (m = 192 .. 1029 unknowns); 4^3 is covered above.
    Against a dense FP64 solve of the oracle's Galerkin operator of the same design."""
    say("coarse solve: n=%d" % n)
    hom, x = make(h3d, n, seed=3)
    mg = orc.MG(n, E, NU)
    mg.set_density(x.astype(np.float64))
    lc = len(hom.sizes) - 1
    nc = hom.sizes[lc]
    assert nc == n // 4
    rng = np.random.default_rng(n)
    fcs = rng.standard_normal(hom.level_shape(lc)).astype(np.float32)
    hom.set_field(lc, 1, fcs)
    hom.coarse_solve()
    uc = hom.get_field(lc, 0)
    A = orc.dense_from_stencil(nc, mg.stencil(lc))
    for s in range(6):
        ref = np.linalg.solve(A, fcs[s].astype(np.float64).T.reshape(-1)).reshape(-1, 3).T
        assert rel(uc[s], ref) < 1e-4, (n, s)


@pytest.mark.parametrize("symmetric", [False, True])
def test_vcycle(h3d, orc, symmetric):
    say("vcycle: n=32 symmetric=%s" % symmetric)
    hom, x = make(h3d, 32, seed=1)
    mg = orc.MG(32, E, NU)
    mg.set_density(x.astype(np.float64))
    hom.macro_rhs()
    f = hom.get_field(0, 1)
    hom.set_field(0, 0, np.zeros(hom.level_shape(0), np.float32))
    hom.vcycle(symmetric)
    hom.vcycle(symmetric)
    got = hom.get_field(0, 0)
    for s in range(6):
        u = np.zeros((3, 32 ** 3))
        u, _ = mg.vcycle(u, f[s].astype(np.float64), 1.0, symmetric)
        u, _ = mg.vcycle(u, f[s].astype(np.float64), 1.0, symmetric)
        assert rel(got[s], u) < 1e-4


def test_mg_solve_and_elastic_matrix_32(h3d, orc):
    say("homogenize: n=32 MG relres 1e-2 and PCG 1e-6")
    hom, x = make(h3d, 32, solver="mg", seed=0)
    C = hom.elasticMatrix(tol=1e-2, warm=False)
    st = hom.last_stats
    assert st["status"] == 0
    mg = orc.MG(32, E, NU)
    mg.set_density(x.astype(np.float64))
    Cref, chi, code, its = mg.homogenize(tol=1e-9, maxit=400, mode="pcg")
    # reference stop rule: stale residual <= 1e-2; C^H is quadratic in the displacement error
    assert rel(C, Cref) < 2e-2
    # sensitivities from the GPU's own solution fields match the oracle's formula on the same fields
    chi_gpu = hom.get_field(0, 0).astype(np.float64)
    dC = np.zeros((6, 6), np.float32)
    dC[0, 0] = dC[1, 1] = dC[2, 2] = 1.0
    dC[0, 1] = dC[1, 2] = dC[0, 2] = 2.0
    s_gpu = hom.Sensitivity(dC)
    s_ref = orc.sensitivity(32, mg.KE, chi_gpu, dC.astype(np.float64))
    assert rel(s_gpu, s_ref) < 1e-5
    C_same = orc.elastic_matrix(32, mg.KE, mg.w, chi_gpu)
    assert rel(C, C_same) < 1e-5


@pytest.mark.parametrize("mixed,mfma", [(True, False), (False, False), (True, True)])
def test_pcg_tight_solve_matches_oracle(h3d, orc, mixed, mfma):
    """mixed: the production configuration (FP16 z and intermediate-level stencils); otherwise all inner
    storage FP32.  mfma: the finest-level operator on FP16 z in element form on the matrix cores.  Either
    way the FP64-refined relres meets the tolerance."""
    say("pcg: n=32 relres 1e-5 mixed=%s mfma=%s" % (mixed, mfma))
    hom = h3d.Homogenization(h3d.HomoConfig(reso=32, solver="pcg", coarse_half=mixed, precond_half=mixed,
                                            fine_mfma=mfma))
    rho = h3d.randc_density(32, 2, 0.2, "rotate3")
    x = (rho.astype(np.float64) ** 3).astype(np.float32)
    hom.update(x)
    C = hom.elasticMatrix(tol=1e-5, warm=False)
    st = hom.last_stats
    assert st["status"] == 0 and max(st["relres"]) <= 1e-5
    mg = orc.MG(32, E, NU)
    mg.set_density(x.astype(np.float64))
    Cref, chi, code, its = mg.homogenize(tol=1e-10, maxit=400, mode="pcg")
    assert rel(C, Cref) < 1e-4
    say("pcg iters gpu", st["iters"], "oracle", its)
    # the reported relres is the FP64 residual of the FP64 refinement iterate: recompute it in the oracle
    U = hom.get_solution64()
    for s in range(6):
        f = orc.macro_rhs(32, mg.KE, mg.w, s)
        r = orc.residual(32, mg.KE, mg.w, U[s], f)
        ref_rel = np.linalg.norm(r) / np.linalg.norm(f)
        assert ref_rel <= 1e-5 * 1.01
        # same FP64 operator and weights: the two residual evaluations agree to rounding
        assert abs(ref_rel - st["relres"][s]) <= 1e-6 * st["relres"][s] + 1e-14
    # warm start from the stored solution converges immediately
    st2 = hom.solve(1e-5, warm=True)
    assert sum(st2["iters"]) == 0


def test_pcg_mfma_matches_row_form(h3d):
    """The two forms of the finest-level operator are the same operator to FP32 rounding, so a 64^3 solve
    takes the same CG steps (to one) and gives the same C^H with either."""
    say("pcg: n=64 mfma vs row form")
    rho = h3d.randc_density(64, 3, 0.2, "rotate3")
    x = (rho.astype(np.float64) ** 3).astype(np.float32)
    out = []
    for mfma in (False, True):
        hom = h3d.Homogenization(h3d.HomoConfig(reso=64, solver="pcg", fine_mfma=mfma))
        hom.update(x)
        C = hom.elasticMatrix(tol=1e-6, warm=False)
        st = hom.last_stats
        say("  mfma=%s iters %s relres %.2e" % (mfma, st["iters"], max(st["relres"])))
        assert st["status"] == 0 and max(st["relres"]) <= 1e-6
        out.append((np.asarray(st["iters"]), C))
    assert np.abs(out[0][0] - out[1][0]).max() <= 1
    assert rel(out[1][1], out[0][1]) < 1e-5


def test_solid_and_laminate(h3d, orc):
    say("closed forms: n=16 solid and laminate")
    from test_oracle_solver import C0, laminate_C

    hom = h3d.Homogenization(h3d.HomoConfig(reso=16, solver="pcg"))
    hom.update(np.ones(16 ** 3, np.float32))
    C = hom.elasticMatrix(tol=1e-6, warm=False)
    assert rel(C, (1 + 1e-9) * C0()) < 1e-5
    layers = np.array([1.0] * 6 + [0.2] * 4 + [0.05] * 6, dtype=np.float32)
    x = np.repeat(layers, 16 * 16)
    hom.update(x)
    C = hom.elasticMatrix(tol=1e-7, warm=False)
    w = 1e-9 + np.array([1.0, 0.2, 0.05])
    ref = laminate_C([wk * C0() for wk in w], np.array([6, 4, 6]) / 16.0)
    assert rel(C, ref) < 1e-4


def test_symmetrize_bit_exact(case16, orc):
    hom, mg, x = case16
    rng = np.random.default_rng(3)
    f = rng.standard_normal(16 ** 3).astype(np.float32)
    for sym in ("reflect3", "reflect6", "rotate3"):
        got = hom.symmetrize(f, sym)
        ref = orc.symmetrize(16, f, sym)
        assert np.array_equal(got, ref), sym


def test_oc_filter_and_update(case16, orc, h3d):
    hom, mg, x = case16
    rng = np.random.default_rng(5)
    rho = h3d.randc_density(16, 3, 0.3, "reflect6")
    sens = (-rng.uniform(0.1, 1.0, 16 ** 3)).astype(np.float32)
    oc = h3d.OCOptimizer(hom, 0.001, 0.05, 0.5)
    fs = oc.filterSens(sens, rho, 2.0)
    ref = orc.oc_filter(16, sens.astype(np.float64), rho.astype(np.float64), 2.0, False)
    assert rel(fs, ref) < 1e-6
    new = oc.update(fs, rho, 0.3)
    refn, _ = orc.oc_update(fs, rho, 0.3, 0.001, 0.05, 0.5)
    # same bisection path; per-cell FP64 arithmetic rounded to FP32 -> at most 1 ulp apart
    ulp = np.abs(new.view(np.int32).astype(np.int64) - refn.view(np.int32).astype(np.int64))
    assert ulp.max() <= 1
    assert abs(new.astype(np.float64).mean() - 0.3) <= 1.01e-4


@pytest.mark.parametrize("n,m,seed", [(20, 1, 0), (4096, 2, 1)])
def test_mma_matches_oracle(h3d, orc, n, m, seed):
    say("mma: n=%d m=%d" % (n, m))
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.2, 0.8, n)
    gpu = h3d.MMAOptimizer(m, n, 1.0, 0.0, 1000.0, 1.0)
    gpu.setBound(0.001, 1.0)
    st = {"x": x.copy(), "xmin": np.full(n, 0.001), "xmax": np.ones(n), "xold1": x.copy(), "xold2": x.copy(),
          "low": np.zeros(n), "upp": np.zeros(n)}
    w = rng.uniform(0.5, 2.0, (m, n))
    for it in range(6):
        xin = st["x"].astype(np.float32).astype(np.float64)
        st["x"][:] = xin
        df0 = (-(1.0 + np.sin(3 * xin))).astype(np.float32)
        g = np.array([(w[i] * xin).sum() / n - 0.4 - 0.1 * i for i in range(m)], dtype=np.float32)
        dg = (w / n).astype(np.float32)
        xg = gpu.update(it, xin.astype(np.float32), df0, g, list(dg))
        orc.mma_update(st, it, df0.astype(np.float64), g.astype(np.float64), dg.astype(np.float64), 1.0,
                       np.zeros(m), np.full(m, 1000.0), np.ones(m), 1.0)
        x64, lo, up, nit = gpu.state()
        assert np.abs(x64 - st["x"]).max() <= 1e-9 * max(1.0, np.abs(st["x"]).max()), it
        assert np.abs(lo - st["low"]).max() <= 1e-12 and np.abs(up - st["upp"]).max() <= 1e-12
        assert np.array_equal(xg, x64.astype(np.float32))
