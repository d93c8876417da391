"""The learned-warm-start workload at its full size: implicit Euler on the
128x128x64 box (dt 0.01, lambda 1e-3, velocity seed 0, source on the helix of
seed 0), 200 steps from x_0 = 0 with the committed network parameters
(data/learned_guess_128x128x64.npz, trained by learned.main on seeds 1-10).

Every step is solved to rtol 1e-8 whatever the guess:
* the true residual of every step's solution is <= 1e-8;
* steps 1, 100 and 200 are within 1e-8 rel-L2 of the CPU oracle's FP64 BiCG
  (to 1e-12) of the same step system from the same previous state;
* the final states of every guess x method agree with each other and with the
  oracle's own 200-step trajectory (each step to 1e-12; its sketch in
  tests/golden/oracle_sketch.json "c5") to 1e-7: every step's solver error
  (up to 1e-8 of the step's solution) enters the next step's right-hand side,
  and 200 of them accumulate to ~3e-8 between guesses.
The iterations of the network's guess against the plain method (previous-step
guess, BiCG) and against the best non-learned guess are printed: at this size
the network does not gain."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 200
DT = 0.01
MODES = ("previous", "net+jacobi", "extrap3+jacobi8@bicgstab", "net+jacobi@bicgstab")
CHECK_STEPS = (1, 100, 200)
CHECK_MODES = ("previous", "net+jacobi@bicgstab")


def step_sources(pkg, p, steps):
    """q_in of every step (the device's vgpu_gaussian_dev on the helix)."""
    centres, sigma = pkg.learned.helix(0, steps)
    ex, ey, ez = p.nax - 1, p.nay - 1, p.naz - 1
    ax, ay, az = (np.arange(n + 1) / n for n in (ex, ey, ez))
    e = np.arange(p.E)
    i, j, k = e % ex, (e // ex) % ey, e // (ex * ey)
    cx, cy, cz = (0.5 * (a[t] + a[t + 1]) for a, t in ((ax, i), (ay, j), (az, k)))
    vol = (ax[i + 1] - ax[i]) * (ay[j + 1] - ay[j]) * (az[k + 1] - az[k])

    def q(n):
        c = centres[n - 1]
        d2 = (cx - c[0]) ** 2 + (cy - c[1]) ** 2 + (cz - c[2]) ** 2
        return np.exp(-d2 / (2.0 * sigma ** 2)) * vol
    return q


def test_implicit_euler_128x128x64_guesses(pkg, po, gpu):
    path = os.path.join(os.path.dirname(pkg.__file__), "data", "learned_guess_128x128x64.npz")
    params = pkg.learned.load_params(path)
    p = pkg.problem.unit_box(128, 128, 64, vel_seed=0, lam=1e-3, scheme=pkg.SCHEME_EULER,
                             dt=DT)
    sk = po.load_sketches().get("c5")
    assert sk is None or sk["steps"] == STEPS
    q_of = step_sources(pkg, p, STEPS)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    ctx = pkg.Context.from_problem(p)
    st = pkg.learned.Stepper(ctx, p, DT)
    net = pkg.Net(ctx)
    buf = st.alloc_states(STEPS + 1)
    try:
        vx, vy, vz = pkg.learned.nodal_velocity(p)
        net.set_params(params["w"], params["b"])
        net.set_static(vx, vy, vz, params["in_scale"], float(params["out_scale"]),
                       float(params["extrap"]))
        net.set_smoothing(int(params["smoothing"]))
        finals, its = {}, {}
        for mode in MODES:
            print(f"c5 {STEPS} steps: {mode}", flush=True)
            r = pkg.learned.run(st, 0, STEPS, mode, net, states=buf, true_residuals=True)
            finals[mode] = ctx.d2h(st.ptr(buf, STEPS), p.N)
            its[mode] = sum(r["iterations"])
            worst = max(r["true_residuals"])
            est = po.sketch_rel_l2(finals[mode], sk) if sk else float("nan")
            print(f"  {its[mode]} iterations, fallbacks {r['fallbacks']}, worst true "
                  f"residual {worst:.2e}, final state rel-L2 vs the oracle's trajectory "
                  f"(sketch) {est:.2e}", flush=True)
            assert worst <= 1e-8, (mode, worst)
            if sk:
                assert est <= 1e-7, (mode, est)
            if mode not in CHECK_MODES:
                continue
            for n in CHECK_STEPS:
                x_prev = ctx.d2h(st.ptr(buf, n - 1), p.N)
                x_n = ctx.d2h(st.ptr(buf, n), p.N)
                p.q_in = q_of(n)
                bb, b, d = po.assemble(p, t_old=x_prev, threads=threads)
                b, d, x0 = po.impose_mask(p.mask, b, d, x_prev)
                xo, ito, _, rco = po.dbcg_simple_omp(b, x0, 1e-12, 20000, bb, p.iss, d,
                                                     p.nax, p.nay, p.naz, p.mask, threads)
                assert rco == 0
                rel = np.linalg.norm(x_n - xo) / np.linalg.norm(xo)
                print(f"  step {n}: rel-L2 vs the oracle's solve of the step {rel:.2e}",
                      flush=True)
                assert rel <= 1e-8, (mode, n, rel)
        ref = finals["previous"]
        for mode in MODES:
            assert np.linalg.norm(finals[mode] - ref) <= 1e-7 * np.linalg.norm(ref), mode
        print(f"  network vs previous-step guess (BiCG): {its['net+jacobi']} vs "
              f"{its['previous']} iterations; network vs cubic extrapolation + 8 sweeps "
              f"(BiCGSTAB): {its['net+jacobi@bicgstab']} vs "
              f"{its['extrap3+jacobi8@bicgstab']}", flush=True)
    finally:
        ctx.dev_free(buf)
        net.close()
        ctx.close()
