"""Writes tests/golden/ch_256_rotate3_seed0.json: the FP64 oracle's homogenized tensor of the 256^3
benchmark design (randc seed 0, rotate3, volume 0.2, x = rho^3, E = 10, nu = 0.3, Emin = 1e-9), solved
by the oracle's MG-PCG to relres 1e-5, with the CPU time it took.

    python tests/golden/make_ch256.py          (about 12 minutes on 8 cores)

The GPU's C^H for the same design is checked against it entry by entry (tests/test_gpu_configs.py and
bench.py)."""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ch_256_rotate3_seed0.json")


def main():
    n, tol = 256, 1e-5
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    rho = oracle.randc_density(n, 0, 0.2, "rotate3").astype(np.float64)
    t0 = time.perf_counter()
    mg = oracle.MG(n)
    mg.set_density(rho ** 3)
    t1 = time.perf_counter()
    C, chi, code, its = mg.homogenize(tol=tol, maxit=400, mode="pcg")
    t2 = time.perf_counter()
    rec = {"design": {"n": n, "seed": 0, "sym": "rotate3", "vol": 0.2, "E": 10.0, "nu": 0.3, "emin": 1e-9, "p": 3},
           "tol": tol, "C": np.asarray(C).reshape(-1).tolist(), "cg_steps": list(map(int, its)), "status": int(code),
           "setup_s": round(t1 - t0, 2), "solve_s": round(t2 - t1, 2), "threads": threads}
    with open(OUT, "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps({k: v for k, v in rec.items() if k != "C"}))


if __name__ == "__main__":
    main()
