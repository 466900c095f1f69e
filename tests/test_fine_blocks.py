"""CPU check of the structure the fine-operator kernels build their 3x3 blocks from (fine_block and
stage_ke in homo3d_amd/csrc/kernels.hip).

The block between vertex v and its neighbour v + d is S_d = sum over the cells e holding both of
w_e KE[7 - e][b(e, d)].  The kernels form it as C_d,ij times a signed sum of the weights of those
cells, with C_d,ij the entry of one cell (the one with e_m = 1 exactly where d_m = +1).  That holds
when, over the cells of d, entry (i, i) of KE[a][b] is one value and entry (i, j), i != j, is one
value times s_i s_j (s_m = +1 at the upper node of axis m).  The isotropic element matrix is
lambda/72 Kl + mu/72 Km with the reference's integer tables Kl, Km (golden/ke_tables.json,
reference homogenization/templateMatrix.cpp:8-61), so checking both tables covers every isotropic
material; the block build is then compared with the cell-by-cell sum on random weights.
"""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def cells_of(d):
    """(e, a, b, signs) for the cells holding v and v + d: cell e at offset (e&1, e>>1&1, e>>2) - 1."""
    dm = (d % 3 - 1, d // 3 % 3 - 1, d // 9 - 1)
    out = []
    for e in range(8):
        em = (e & 1, (e >> 1) & 1, e >> 2)
        b = [dm[m] + 1 - em[m] for m in range(3)]
        if all(0 <= x <= 1 for x in b):
            out.append((e, 7 - e, b[0] + 2 * b[1] + 4 * b[2], [(-1 if em[m] else 1) for m in range(3)]))
    return dm, out


def block_direct(K, d, w):
    S = np.zeros((3, 3))
    for e, a, b, _ in cells_of(d)[1]:
        S += w[e] * K[3 * a:3 * a + 3, 3 * b:3 * b + 3]
    return S


def block_constants(K):
    """C_d,ij as stage_ke sets them."""
    C = np.zeros((27, 3, 3))
    for d in range(27):
        dm = (d % 3 - 1, d // 3 % 3 - 1, d // 9 - 1)
        e0 = sum(1 << m for m in range(3) if dm[m] > 0)
        b0 = sum((dm[m] + 1 - ((e0 >> m) & 1)) << m for m in range(3))
        a0 = 7 - e0
        C[d] = K[3 * a0:3 * a0 + 3, 3 * b0:3 * b0 + 3]
    return C


def block_from_forms(C, d, w):
    """fine_block: four signed sums of the cell weights, nine products."""
    dm, cells = cells_of(d)
    free = [dm[m] == 0 for m in range(3)]

    def form(axes):
        s = 0.0
        for e, _, _, sg in cells:
            f = 1.0
            for m in axes:
                if free[m]:
                    f *= sg[m]
            s += f * w[e]
        return s

    F = np.empty((3, 3))
    for i in range(3):
        for j in range(3):
            F[i, j] = form(()) if i == j else form((i, j))
    return C[d] * F


def tables():
    t = json.load(open(os.path.join(GOLDEN, "ke_tables.json")))
    return (np.array(t["kelam72"], dtype=float).reshape(24, 24),
            np.array(t["kemu72"], dtype=float).reshape(24, 24))


def test_cell_entries_share_one_value_up_to_the_sign_pattern():
    for K in tables():
        for d in range(27):
            _, cells = cells_of(d)
            for i in range(3):
                for j in range(3):
                    vals = [K[3 * a + i, 3 * b + j] * (1 if i == j else s[i] * s[j]) for _, a, b, s in cells]
                    assert max(vals) - min(vals) == 0.0, (d, i, j, vals)


def test_block_build_matches_cell_sum():
    kl, km = tables()
    rng = np.random.default_rng(7)
    for E, nu in ((10.0, 0.3), (1e6, 0.45), (1.0, 0.0), (3.0, -0.5)):
        lam, mu = nu * E / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
        K = lam / 72 * kl + mu / 72 * km
        C = block_constants(K)
        scale = np.abs(K).max()
        for _ in range(20):
            w = rng.uniform(1e-9, 1.0, 8)
            for d in range(27):
                err = np.abs(block_from_forms(C, d, w) - block_direct(K, d, w)).max() / scale
                assert err < 1e-14, (E, nu, d, err)


def test_a_wrong_sign_pattern_is_caught():
    """Planting a sign error in the forms (the first free axis unsigned) breaks the match."""
    kl, km = tables()
    K = 0.5 * kl + km
    C = block_constants(K)
    w = np.random.default_rng(3).uniform(0.1, 1.0, 8)
    d = 13  # the vertex itself: all eight cells, every axis free
    dm, cells = cells_of(d)
    bad = np.empty((3, 3))
    for i in range(3):
        for j in range(3):
            bad[i, j] = sum(w[e] * (1 if i == j else s[j]) for e, _, _, s in cells)
    assert np.abs(C[d] * bad - block_direct(K, d, w)).max() > 1e-3 * np.abs(K).max()
