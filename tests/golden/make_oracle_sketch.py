"""Sketches of the CPU oracle's FP64 solutions of the large workloads, written
to oracle_sketch.json, for rel-L2 checks of GPU solutions whose full oracle
vectors (8.7-34 MB each) are too large to commit.

For each system the oracle assembles the element matrices
(orc_assemble_omp, bitwise the serial przest3d loop), solves with the
reference's Jacobi-BiCG (dbcg_simple, VANADIS.FOR:977-1055; OpenMP products,
orc_dbcg_simple_omp) to rtol 1e-12, cross-checks the solution against the
oracle's Jacobi-BiCGSTAB run to 1e-13, and stores K Gaussian projections
s_k = g_k . x of the solution, g_k = numpy default_rng(SEED0 + k)
.standard_normal(N).  For any vector y, sqrt(mean_k (g_k . y - s_k)^2)
estimates |y - x| (Johnson-Lindenstrauss: E (g . v)^2 = |v|^2; with K = 32
the estimate is within about +-25 % of the true norm with high probability).
Systems whose BiCG stalls before 1e-12 (plain BiCG loses its recurrence on
some lambda ~ 1e-4 systems) take the BiCGSTAB solution as the reference,
recorded per entry.

    python tests/golden/make_oracle_sketch.py c1 c2 c3 c4a c4b c5   (minutes; 8 threads)
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import __graft_entry__ as ge  # noqa: E402

SEED0 = 910000
OUT = os.path.join(HERE, "oracle_sketch.json")


def projections(x, k, seed0=SEED0):
    return [float(np.dot(np.random.default_rng(seed0 + i).standard_normal(x.size), x))
            for i in range(k)]


def solve_entry(po, p, k, threads, bicg_itmax=20000):
    t0 = time.time()
    bb, b, d = po.assemble(p, threads=threads)
    b, d, x0 = po.impose_mask(p.mask, b, d, np.zeros(p.N))
    args = (bb, p.iss, d, p.nax, p.nay, p.naz, p.mask)
    x, it, err, rc = po.dbcg_simple_omp(b, x0, 1e-12, bicg_itmax, *args, threads)
    xs, its, errs, sts = po.bicgstab_jacobi(b, x0, 1e-13, 20000, *args, threads=threads)
    n = np.linalg.norm
    ref = "bicg"
    if rc != 0:
        x, ref = xs, "bicgstab"
    r = b - po.matvec_0_colored(x, p.iss, bb, p.nax, p.nay, p.naz, p.mask, threads)
    e = {"N": int(p.N), "E": int(p.E), "rtol": 1e-12, "reference_solver": ref,
         "bicg_iterations": int(it), "bicg_converged": rc == 0,
         "bicgstab_iterations": int(its), "bicgstab_status": int(sts),
         "bicg_vs_bicgstab_rel_l2": float(n(x - xs) / n(xs)) if rc == 0 else None,
         "true_rel_residual": float(n(r) / n(b)), "b_norm2": float(n(b)),
         "x_norm2": float(n(x)), "x_sum": float(x.sum()),
         "mass": po.mass(p.iss, p.xx, p.yy, p.zz, x, p.E),
         "sketch_seed0": SEED0, "sketch": projections(x, k),
         "oracle_seconds": round(time.time() - t0, 1)}
    print(f"  N {p.N}: bicg it {it} rc {rc}, bicgstab it {its} st {sts}, "
          f"rel {e['bicg_vs_bicgstab_rel_l2']}, true res {e['true_rel_residual']:.2e}, "
          f"{e['oracle_seconds']} s", flush=True)
    return e


def c5_entry(pkg, po, threads, steps=200, k=32):
    """The implicit-Euler workload (128x128x64, dt 0.01, lambda 1e-3,
    velocity seed 0, source on the helix of seed 0 laid over `steps` steps):
    `steps` steps from x_0 = 0, each solved by the oracle's BiCG to 1e-12 from
    the previous state; the sketch is of the last state."""
    dt = 0.01
    p = pkg.problem.unit_box(128, 128, 64, vel_seed=0, lam=1e-3, scheme=2, dt=dt)
    centres, sigma = pkg.learned.helix(0, steps)
    ex, ey, ez = p.nax - 1, p.nay - 1, p.naz - 1
    ax, ay, az = (np.arange(n + 1) / n for n in (ex, ey, ez))
    e = np.arange(p.E)
    i, j, kk = e % ex, (e // ex) % ey, e // (ex * ey)
    cx, cy, cz = (0.5 * (a[t] + a[t + 1]) for a, t in ((ax, i), (ay, j), (az, kk)))
    vol = (ax[i + 1] - ax[i]) * (ay[j + 1] - ay[j]) * (az[kk + 1] - az[kk])
    x = np.zeros(p.N)
    its = []
    t0 = time.time()
    for n in range(steps):
        c = centres[n]
        d2 = (cx - c[0]) ** 2 + (cy - c[1]) ** 2 + (cz - c[2]) ** 2
        p.q_in = np.exp(-d2 / (2.0 * sigma ** 2)) * vol
        bb, b, d = po.assemble(p, t_old=x, threads=threads)
        b, d, x = po.impose_mask(p.mask, b, d, x)
        x, it, err, rc = po.dbcg_simple_omp(b, x, 1e-12, 20000, bb, p.iss, d, p.nax, p.nay,
                                            p.naz, p.mask, threads)
        assert rc == 0, (n, it, err)
        its.append(int(it))
        if (n + 1) % 20 == 0:
            print(f"  step {n + 1}: {it} iterations", flush=True)
    return {"workload": f"implicit Euler 128x128x64, dt {dt}, helix seed 0 over {steps} "
                        f"steps, x_0 = 0; sketch of x_{steps}",
            "N": int(p.N), "E": int(p.E), "steps": steps, "rtol": 1e-12,
            "reference_solver": "bicg", "bicg_iterations_per_step": its,
            "x_norm2": float(np.linalg.norm(x)), "x_sum": float(x.sum()),
            "mass": po.mass(p.iss, p.xx, p.yy, p.zz, x, p.E),
            "sketch_seed0": SEED0, "sketch": projections(x, k),
            "oracle_seconds": round(time.time() - t0, 1)}


def main(names, threads=8):
    pkg = ge.load_package()
    po = ge.load_oracle()
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for name in names:
        print(name, flush=True)
        if name in ("c1", "c2", "c3"):
            data[name] = solve_entry(po, pkg.problem.config_problem(name), 32, threads)
            data[name]["workload"] = f"config_problem('{name}')"
        elif name == "c5":
            data[name] = c5_entry(pkg, po, threads)
        elif name == "c4a":
            data[name] = {"workload": "batch_rhs_problems()", "members": [
                solve_entry(po, q, 16, threads, bicg_itmax=4000)
                for q in pkg.problem.batch_rhs_problems()]}
        elif name == "c4b":
            data[name] = {"workload": "batch_subdomain_problems()", "members": [
                solve_entry(po, q, 16, threads, bicg_itmax=4000)
                for q in pkg.problem.batch_subdomain_problems()]}
        else:
            raise KeyError(name)
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["c2", "c3"])
