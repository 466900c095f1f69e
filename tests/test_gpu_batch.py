"""Batched uneven designs (homo3d_batch_*) and the operator-form autotuner on the GPU.

Designs of 32^3 .. 96^3 (48^3, 80^3 and 96^3 take the one-colour-per-launch smoother path, 256 % (n/2) != 0)
solved concurrently, each to FP64 relres 1e-5: the batch gives the same C^H and CG steps as solving
the designs one by one on fresh handles (each design runs the same kernels on its own stream), and the
48^3 and 80^3 tensors agree with the FP64 oracle to 1e-4 (the solve tolerance bounds C^H's error).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def h3d():
    import homo3d_amd

    return homo3d_amd


def designs(h3d, sizes, seed0=0):
    xs = []
    for i, n in enumerate(sizes):
        rng = np.random.default_rng(seed0 + i)
        vol = float(rng.uniform(0.1, 0.5))
        rho = h3d.randc_density(n, seed0 + i, vol, ("reflect6", "rotate3", "reflect3", "none")[i % 4])
        xs.append((rho.astype(np.float64) ** 3).astype(np.float32))
    return xs


def test_batch_matches_single_designs_and_oracle(h3d, orc):
    sizes = [48, 80, 32, 96, 64, 48, 80, 32]
    say("batch: sizes %s" % sizes)
    xs = designs(h3d, sizes)
    b = h3d.Batch(sizes, h3d.HomoConfig(solver="pcg", max_iter=400), max_inflight=4)
    C, stats = b.solve(xs, tol=1e-5)
    for i, (n, st) in enumerate(zip(sizes, stats)):
        say("  design %d n=%d CG steps %s relres %.2e" % (i, n, st["iters"], max(st["relres"])))
        assert st["status"] == 0 and max(st["relres"]) <= 1e-5
    # one by one on fresh handles: identical results
    for i in (0, 1, 3):
        hom = h3d.Homogenization(h3d.HomoConfig(reso=sizes[i], solver="pcg", max_iter=400))
        hom.update(xs[i])
        st = hom.solve(1e-5, warm=False)
        assert st["iters"] == stats[i]["iters"]
        assert np.array_equal(hom.elasticMatrix(solve=False), C[i])
        hom.close()
    # the one-colour smoother sizes against the FP64 oracle
    for i in (0, 1):
        mg = orc.MG(sizes[i])
        mg.set_density(xs[i].astype(np.float64))
        Cref, _, _, _ = mg.homogenize(tol=1e-9, maxit=400, mode="pcg")
        err = np.abs(C[i] - Cref).max() / np.abs(Cref).max()
        say("  n=%d C^H vs oracle %.2e" % (sizes[i], err))
        assert err <= 1e-4


def test_autotune_picks_a_form_and_keeps_results(h3d):
    n = 64
    say("autotune: n=%d" % n)
    x = designs(h3d, [n], seed0=7)[0]
    hom = h3d.Homogenization(h3d.HomoConfig(reso=n, solver="pcg", max_iter=400))
    hom.update(x)
    st0 = hom.solve(1e-5, warm=False)
    C0 = hom.elasticMatrix(solve=False)
    rep = hom.autotune()
    say("  %s" % rep)
    assert rep["row_ms"] > 0 and rep["mfma_ms"] > 0
    assert rep["chosen"] == ("mfma" if rep["mfma_ms"] < rep["row_ms"] else "row")
    st1 = hom.solve(1e-5, warm=False)
    assert st1["status"] == 0 and max(st1["relres"]) <= 1e-5
    assert np.abs(hom.elasticMatrix(solve=False) - C0).max() <= 1e-5 * np.abs(C0).max()
    assert abs(sum(st1["iters"]) - sum(st0["iters"])) <= 6
