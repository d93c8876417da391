"""The learned-warm-start workload at its full size: implicit Euler on the
128x128x64 box (dt 0.01, lambda 1e-3, velocity seed 0, source on the helix of
seed 0), 200 steps from x_0 = 0 with the committed network parameters
(data/learned_guess_128x128x64.npz, trained by learned.main on seeds 1-10).

Every step is solved to rtol 1e-8 whatever the guess: the true residual of
every step's solution is <= 1e-8, and the final state of every guess x method
is within 1e-8 rel-L2 of the CPU oracle's trajectory (each step its FP64 BiCG
to 1e-12; tests/golden/oracle_sketch.json "c5").  The iterations of the
network's guess against the plain method (previous-step guess, BiCG) and
against the best non-learned guess are printed: at this size the network does
not gain (its 5-layer prediction is not cheaper than the cubic extrapolation
with Jacobi sweeps it competes with)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 200
MODES = ("previous", "net+jacobi", "extrap3+jacobi8@bicgstab", "net+jacobi@bicgstab")


def test_implicit_euler_128x128x64_guesses(pkg, po, gpu):
    path = os.path.join(os.path.dirname(pkg.__file__), "data", "learned_guess_128x128x64.npz")
    params = pkg.learned.load_params(path)
    p = pkg.problem.unit_box(128, 128, 64, vel_seed=0, lam=1e-3, scheme=pkg.SCHEME_EULER,
                             dt=0.01)
    sk = po.load_sketches().get("c5")
    assert sk is None or sk["steps"] == STEPS
    ctx = pkg.Context.from_problem(p)
    st = pkg.learned.Stepper(ctx, p, 0.01)
    net = pkg.Net(ctx)
    try:
        vx, vy, vz = pkg.learned.nodal_velocity(p)
        net.set_params(params["w"], params["b"])
        net.set_static(vx, vy, vz, params["in_scale"], float(params["out_scale"]),
                       float(params["extrap"]))
        net.set_smoothing(int(params["smoothing"]))
        finals, its = {}, {}
        for mode in MODES:
            print(f"c5 {STEPS} steps: {mode}", flush=True)
            r = pkg.learned.run(st, 0, STEPS, mode, net, true_residuals=True)
            finals[mode] = ctx.d2h(r["last"], p.N)
            its[mode] = sum(r["iterations"])
            worst = max(r["true_residuals"])
            est = po.sketch_rel_l2(finals[mode], sk) if sk else float("nan")
            print(f"  {its[mode]} iterations, fallbacks {r['fallbacks']}, worst true "
                  f"residual {worst:.2e}, final state rel-L2 vs oracle (sketch) {est:.2e}",
                  flush=True)
            assert worst <= 1e-8, (mode, worst)
            if sk:
                assert est <= 1e-8, (mode, est)
        ref = finals["previous"]
        for mode in MODES:
            assert np.linalg.norm(finals[mode] - ref) <= 2e-8 * np.linalg.norm(ref), mode
        print(f"  network vs previous-step guess (BiCG): {its['net+jacobi']} vs "
              f"{its['previous']} iterations; network vs cubic extrapolation + 8 sweeps "
              f"(BiCGSTAB): {its['net+jacobi@bicgstab']} vs "
              f"{its['extrap3+jacobi8@bicgstab']}", flush=True)
    finally:
        net.close()
        ctx.close()
