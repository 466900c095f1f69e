"""Pins for the FP64 oracle's stateless kernels (CPU only).

Each test checks an oracle function against something the oracle does not share code with:
the reference's embedded tables (golden vector), dense brute-force assembly, closed forms or
invariants.
"""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def lame(E, nu):
    return nu * E / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def test_ke_matches_reference_integer_tables(orc):
    d = json.load(open(os.path.join(GOLDEN, "ke_tables.json")))
    kl = np.array(d["kelam72"], dtype=float).reshape(24, 24)
    km = np.array(d["kemu72"], dtype=float).reshape(24, 24)
    assert np.abs(orc.ke_lame(1, 0) * 72 - kl).max() < 1e-12
    assert np.abs(orc.ke_lame(0, 1) * 72 - km).max() < 1e-12
    lam, mu = lame(10.0, 0.3)
    assert np.abs(orc.ke(10.0, 0.3) - (lam / 72 * kl + mu / 72 * km)).max() < 1e-12


def test_ke_symmetric_with_rigid_null_space(orc):
    K = orc.ke(10.0, 0.3)
    assert np.abs(K - K.T).max() < 1e-14
    ev = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(ev) < 1e-10) == 6
    assert ev.min() > -1e-12
    # rigid translations and infinitesimal rotations are in the kernel
    for c in range(3):
        t = np.zeros(24)
        t[c::3] = 1
        assert np.abs(K @ t).max() < 1e-12
    pos = np.array([[i % 2, i // 2 % 2, i // 4] for i in range(8)], dtype=float)
    for axis in range(3):
        om = np.zeros(3)
        om[axis] = 1
        rot = np.cross(om, pos).ravel()
        assert np.abs(K @ rot).max() < 1e-12


def test_level_sizes_closed_form(orc):
    assert orc.level_sizes(32) == [32, 16, 8, 4]
    assert orc.level_sizes(48) == [48, 24, 12, 6]
    assert orc.level_sizes(64) == [64, 32, 16, 8, 4]
    assert orc.level_sizes(80) == [80, 40, 20, 10, 5]
    assert orc.level_sizes(96) == [96, 48, 24, 12, 6]
    assert orc.level_sizes(128) == [128, 64, 32, 16, 8, 4]
    assert orc.level_sizes(256) == [256, 128, 64, 32, 16, 8, 4]
    assert orc.level_sizes(100) == [112, 56, 28, 14, 7]


def test_philox_known_answers(orc):
    # Random123 known-answer vectors for philox4x32-10
    assert orc.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert orc.philox4x32_10([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert orc.philox4x32_10([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def dense_assembly(n, K, w, pin_diag=None):
    """Brute-force global matrix by direct element loops in Python (independent of orc)."""
    N = n ** 3
    A = np.zeros((3 * N, 3 * N))
    for e in range(N):
        cx, cy, cz = e % n, e // n % n, e // (n * n)
        vs = [((cx + b % 2) % n) + n * (((cy + b // 2 % 2) % n) + n * ((cz + b // 4) % n)) for b in range(8)]
        for a in range(8):
            for b in range(8):
                A[3 * vs[a]:3 * vs[a] + 3, 3 * vs[b]:3 * vs[b] + 3] += w[e] * K[3 * a:3 * a + 3, 3 * b:3 * b + 3]
    if pin_diag is not None:
        A[0:3, :] = 0
        A[:, 0:3] = 0
        A[0:3, 0:3] = pin_diag * np.eye(3)
    return A


def to_dof(u):  # (3,N) -> dof order 3*v+c
    return u.T.reshape(-1)


def from_dof(x):
    return x.reshape(-1, 3).T.copy()


@pytest.fixture(scope="module")
def small(orc):
    rng = np.random.default_rng(1)
    n = 4
    K = orc.ke(10.0, 0.3)
    w = 1e-9 + rng.uniform(0.001, 1.0, n ** 3) ** 3
    return n, K, w, rng


def test_apply_K_matches_dense(orc, small):
    n, K, w, rng = small
    A = dense_assembly(n, K, w)
    u = rng.standard_normal((3, n ** 3))
    y = orc.apply_K(n, K, w, u, pin=False)
    assert np.allclose(to_dof(y), A @ to_dof(u), rtol=1e-12, atol=1e-12)
    Ap = dense_assembly(n, K, w, pin_diag=0.0)
    yp = orc.apply_K(n, K, w, u, pin=True)
    assert np.allclose(to_dof(yp), Ap @ to_dof(u), rtol=1e-12, atol=1e-12)
    # symmetric, translations in the kernel before pinning
    v = rng.standard_normal((3, n ** 3))
    assert abs(np.vdot(orc.apply_K(n, K, w, u, False), v) - np.vdot(u, orc.apply_K(n, K, w, v, False))) < 1e-9
    const = np.ones((3, n ** 3))
    assert np.abs(orc.apply_K(n, K, w, const, False)).max() < 1e-12


def test_residual_and_rhs(orc, small):
    n, K, w, rng = small
    u = rng.standard_normal((3, n ** 3))
    f = rng.standard_normal((3, n ** 3))
    r = orc.residual(n, K, w, u, f)
    y = orc.apply_K(n, K, w, u, pin=True)
    assert np.allclose(r[:, 1:], (f - y)[:, 1:]) and np.all(r[:, 0] == 0)
    # uniform density: interior cancellation gives a zero load for every strain
    voigt = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    for s in range(6):
        f0 = orc.macro_rhs(n, K, np.full(n ** 3, 0.7), s)
        assert np.abs(f0).max() < 1e-12
        fs = orc.macro_rhs(n, K, w, s)
        # f = sum_e w_e KE u_e for the unwrapped affine field u(p) = eps p (unit Voigt strain s,
        # engineering shear split in halves)
        eps = np.zeros((3, 3))
        i, j = voigt[s]
        eps[i, j] = eps[j, i] = 1.0 if i == j else 0.5
        ref = np.zeros((3, n ** 3))
        for e in range(n ** 3):
            c = np.array([e % n, e // n % n, e // (n * n)])
            ue = np.concatenate([eps @ (c + [b % 2, b // 2 % 2, b // 4]) for b in range(8)])
            fe = w[e] * (K @ ue)
            for b in range(8):
                q = (c + [b % 2, b // 2 % 2, b // 4]) % n
                ref[:, q[0] + n * (q[1] + n * q[2])] += fe[3 * b:3 * b + 3]
        ref[:, 0] = 0
        assert np.allclose(fs, ref, rtol=1e-12, atol=1e-12)
        fs2 = orc.macro_rhs(n, K, 2 * w, s)
        assert np.allclose(fs2, 2 * fs)


def test_gs_matches_dense_block_gauss_seidel(orc, small):
    n, K, w, rng = small
    A = dense_assembly(n, K, w, pin_diag=1.0)
    u0 = rng.standard_normal((3, n ** 3))
    u0[:, 0] = 0
    f = rng.standard_normal((3, n ** 3))
    f[:, 0] = 0
    for omega, reverse in [(1.0, False), (1.3, False), (1.0, True)]:
        got = orc.gs_fine(n, K, w, u0, f, omega, reverse)
        x = to_dof(u0).copy()
        b = to_dof(f)
        colours = range(7, -1, -1) if reverse else range(8)
        for c in colours:
            par = (1 - c % 2, 1 - c // 2 % 2, 1 - c // 4)
            verts = [v for v in range(n ** 3) if (v % n % 2, v // n % n % 2, v // (n * n) % 2) == par]
            newx = x.copy()
            for v in verts:
                comps = (2, 1, 0) if reverse else (0, 1, 2)
                for i in comps:
                    row = 3 * v + i
                    s = b[row] - A[row] @ newx + A[row, row] * newx[row]
                    newx[row] = omega * s / A[row, row] + (1 - omega) * newx[row]
            x = newx
        x[0:3] = 0
        assert np.allclose(to_dof(got), x, rtol=1e-11, atol=1e-11)


def dense_P(orc, nf):
    Nc = (nf // 2) ** 3
    P = np.zeros((3 * nf ** 3, 3 * Nc))
    for V in range(Nc):
        for c in range(3):
            uc = np.zeros((3, Nc))
            uc[c, V] = 1
            P[:, 3 * V + c] = to_dof(orc.prolong_add(nf, uc, np.zeros((3, nf ** 3)), zero_pin=False))
    return P


def test_restriction_is_prolongation_transpose(orc):
    nf = 8
    P = dense_P(orc, nf)
    R = np.zeros_like(P.T)
    for v in range(nf ** 3):
        for c in range(3):
            rf = np.zeros((3, nf ** 3))
            rf[c, v] = 1
            R[:, 3 * v + c] = to_dof(orc.restrict(nf, rf))
    assert np.abs(R - P.T).max() == 0
    # trilinear hat: a coarse spike prolongates to weights 1, 1/2, 1/4, 1/8
    col = from_dof(P[:, 3 * 9])  # coarse vertex (1,2,0)
    vals = sorted(set(np.round(col[0][col[0] > 0], 12)))
    assert vals == [0.125, 0.25, 0.5, 1.0]


def test_galerkin_matches_dense_triple_product(orc):
    rng = np.random.default_rng(7)
    nf = 8
    K = orc.ke(10.0, 0.3)
    w = 1e-9 + rng.uniform(0.001, 1.0, nf ** 3) ** 3
    P = dense_P(orc, nf)
    A = dense_assembly(nf, K, w, pin_diag=8.0)  # the reference puts 8*I at the pinned node
    Ac = P.T @ A @ P
    st = orc.galerkin_fine(nf, K, w, clamp=0.0)
    got = orc.dense_from_stencil(nf // 2, st)
    assert np.abs(got - Ac).max() < 1e-12 * np.abs(Ac).max()
    # coarser level from a stored stencil: 16 -> 8 -> 4
    n2 = 16
    w2 = 1e-9 + rng.uniform(0.001, 1.0, n2 ** 3) ** 3
    st8 = orc.galerkin_fine(n2, K, w2, clamp=0.0)
    A8 = orc.dense_from_stencil(8, st8)
    st4 = orc.galerkin_coarse(8, st8, clamp=0.0)
    got4 = orc.dense_from_stencil(4, st4)
    ref4 = P.T @ A8 @ P
    assert np.abs(got4 - ref4).max() < 1e-12 * np.abs(ref4).max()
    # the clamp raises small diagonals only
    stc = orc.galerkin_fine(nf, K, np.full(nf ** 3, 1e-9), clamp=1e-4)
    diag = stc[:, 13, [0, 1, 2], [0, 1, 2]]
    assert diag.min() >= 1e-4


def test_stencil_ops_match_dense(orc):
    rng = np.random.default_rng(3)
    n = 4
    st = rng.standard_normal((n ** 3, 27, 3, 3))
    A = orc.dense_from_stencil(n, st)
    # make a symmetric stencil from the symmetrised dense matrix by re-reading A
    u = rng.standard_normal((3, n ** 3))
    y = orc.stencil_apply(n, st, u)
    Araw = np.zeros_like(A)
    for v in range(n ** 3):
        x, yy, z = v % n, v // n % n, v // (n * n)
        for d in range(27):
            nb = ((x + d % 3 - 1) % n) + n * (((yy + d // 3 % 3 - 1) % n) + n * ((z + d // 9 - 1) % n))
            Araw[3 * v:3 * v + 3, 3 * nb:3 * nb + 3] += st[v, d]
    assert np.allclose(to_dof(y), Araw @ to_dof(u))
    f = rng.standard_normal((3, n ** 3))
    assert np.allclose(orc.stencil_residual(n, st, u, f), f - y)
