"""Learned initial guess for implicit-Euler time stepping, device resident.

Time stepping: implicit Euler of the DR advection-diffusion operator,
A = H + C/dt, b_n = F_n + (C/dt) x_{n-1}, x_0 = 0, on a unit box; the
Gaussian source F_n moves along a helix drawn from the seed.  Every step is
solved by Jacobi-BiCG to rtol from an initial guess:

  previous   x_{n-1}                      (the reference host's choice: TNW2
                                           is passed in/out, VANADIS.FOR:809-810)
  extrap     2 x_{n-1} - x_{n-2}
  net        x_{n-1} + e (x_{n-1} - x_{n-2})
             + s * S(net(x_{n-1}, x_{n-1} - x_{n-2}, src, v)), S a low-pass
             filter; falling back to x_{n-1} when that has the smaller
             initial residual
  "+jacobi" after a mode: then a few damped Jacobi sweeps on A x = b_n

The guess only sets x0; every step is solved to the same rtol.  For this
mass-dominated operator the slowest BiCG modes are the grid-scale ones, so a
guess with a smaller but rougher initial residual can take more iterations:
the filter S (in training and prediction) keeps the network's BF16 rounding
out of those modes.

The network (vgpu_net_*) is trained on the device with its own forward and
backward kernels on states the solver itself produced for other seeds.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import _lib, problem

TWO_PI = 2.0 * math.pi


def helix(seed: int, steps: int):
    """Source centres for steps 1..steps and the source width, from `seed`:
    radius 0.15-0.3 around (0.5, 0.5), 1-2 turns, z rising from 0.2-0.4 by
    0.3-0.5; sigma 0.06-0.12."""
    u = problem.uniform01(seed, 6, stream=5)
    R = 0.15 + 0.15 * u[0]
    turns = 1.0 + u[1]
    ph = TWO_PI * u[2]
    z0 = 0.2 + 0.2 * u[3]
    z1 = z0 + 0.3 + 0.2 * u[4]
    sigma = 0.06 + 0.06 * u[5]
    n = np.arange(1, steps + 1) / steps
    a = ph + TWO_PI * turns * n
    c = np.stack([0.5 + R * np.cos(a), 0.5 + R * np.sin(a), z0 + (z1 - z0) * n], axis=1)
    return c, sigma


def nodal_velocity(p, seed: int = 0):
    """The velocity field of problem.unit_box evaluated at the nodes."""
    return problem.velocity_field(p.xx, p.yy, p.zz, seed)


class Stepper:
    """Implicit-Euler steps on one context, all vectors on the device."""

    def __init__(self, ctx, p, dt: float):
        self.ctx, self.p, self.dt = ctx, p, dt
        ctx.set_scheme(_lib.SCHEME_EULER, dt)
        N, E = p.N, p.E
        self.N = N
        self.q = ctx.dev_alloc(8 * E)
        self.src = ctx.dev_alloc(8 * N)
        self.b = ctx.dev_alloc(8 * N)
        self.zero = ctx.dev_alloc(8 * N)
        ctx.h2d(self.zero, np.zeros(N))

    def ptr(self, base, i):
        return _lib._vp(base.value + 8 * self.N * i)

    def alloc_states(self, count):
        buf = self.ctx.dev_alloc(8 * self.N * count)
        for i in range(count):
            self.ctx.copy_dev(self.ptr(buf, i), self.zero, 8 * self.N)
        return buf

    def source(self, centre, sigma):
        self.ctx.gaussian_dev(centre, sigma, self.q, self.src)

    def step(self, centre, sigma, x1, x2, out, mode="previous", net=None, rtol=1e-8,
             itmax=20000, sweeps=2, omega=0.8, x3=None, x4=None):
        """Solve step n into `out` from the guess of `mode`; x1 = x_{n-1},
        x2 = x_{n-2}, x3 = x_{n-3} or None (device).  mode is
        guess[+jacobi[N]][@method] (N damped Jacobi sweeps, default `sweeps`):
        guess "previous", "extrap" (2 x1 - x2),
        "extrap2" (3 x1 - 3 x2 + x3; linear while x3 is None), "extrap3"
        (4 x1 - 6 x2 + 4 x3 - x4; quadratic while x4 is None) or "net", method
        "bicg" (default) or "bicgstab".  Returns (iterations, seconds guess +
        solve, fallback used)."""
        ctx = self.ctx
        self.source(centre, sigma)
        ctx.assemble_rhs_dev(self.q, x1, self.b)
        ctx.residual_norm_dev(self.b, x1)  # synchronise before timing
        t0 = time.perf_counter()
        fallback = False
        mode, _, method = mode.partition("@")
        base, _, post = mode.partition("+")
        if base == "previous":
            ctx.copy_dev(out, x1, 8 * self.N)
        elif base == "extrap" or (base == "extrap2" and x3 is None):
            ctx.axpby_dev(2.0, x1, -1.0, x2, out)
        elif base == "extrap2" or (base == "extrap3" and x4 is None and x3 is not None):
            ctx.axpbypcz_dev(3.0, x1, -3.0, x2, 1.0, x3, out)
        elif base == "extrap3" and x3 is None:
            ctx.axpby_dev(2.0, x1, -1.0, x2, out)
        elif base == "extrap3":
            ctx.axpbypcz_dev(4.0, x1, -6.0, x2, 4.0, x3, out)
            ctx.axpby_dev(1.0, out, -1.0, x4, out)
        elif base == "net":
            net.predict_dev(x1, x2, self.src, out)
            if ctx.residual_norm_dev(self.b, x1) < ctx.residual_norm_dev(self.b, out):
                ctx.copy_dev(out, x1, 8 * self.N)
                fallback = True
        else:
            raise ValueError(mode)
        if post.startswith("jacobi"):
            n_sw = int(post[6:]) if len(post) > 6 else sweeps
            ctx.jacobi_dev(self.b, out, n_sw, omega)
        elif post:
            raise ValueError(mode)
        if method:
            ctx.set_method(method)
        try:
            it, err, sec, rc = ctx.solve_dev(self.b, out, rtol=rtol, itmax=itmax, timed=False)
        finally:
            if method:
                ctx.set_method("bicg")
        t1 = time.perf_counter()
        if rc != _lib.OK:
            raise _lib.VgpuError(rc, f"step solve failed ({mode}): it={it} err={err}")
        return it, t1 - t0, fallback


def run(stepper, seed, steps, mode="previous", net=None, states=None, rtol=1e-8,
        true_residuals=False):
    """Run `steps` steps from x_0 = 0.  states: device buffer of steps+1
    vectors that receives x_0..x_steps (else two rotating work vectors are
    used).  Returns per-step iterations, seconds and fallbacks (and with
    true_residuals the true relative residual |b - A x_n| / |b| of every
    step's solution, computed after the step's timer)."""
    centres, sigma = helix(seed, steps)
    own = states is None
    R = 5  # ring: x_{n-1} .. x_{n-4} and the output
    buf = stepper.alloc_states(R) if own else states
    its, secs, fbs, trs = [], [], [], []
    for n in range(1, steps + 1):
        if own:
            x1, x2, x3, x4, out = (stepper.ptr(buf, (n - 1) % R), stepper.ptr(buf, (n - 2) % R),
                                   stepper.ptr(buf, (n - 3) % R), stepper.ptr(buf, (n - 4) % R),
                                   stepper.ptr(buf, n % R))
            if n == 1:
                x2 = x1
        else:
            x1, x2, x3, x4, out = (stepper.ptr(buf, n - 1), stepper.ptr(buf, max(n - 2, 0)),
                                   stepper.ptr(buf, max(n - 3, 0)),
                                   stepper.ptr(buf, max(n - 4, 0)), stepper.ptr(buf, n))
        if n < 3:
            x3 = None
        if n < 4:
            x4 = None
        it, sec, fb = stepper.step(centres[n - 1], sigma, x1, x2, out, mode, net, rtol, x3=x3,
                                   x4=x4)
        if true_residuals:
            trs.append(stepper.ctx.residual_norm_dev(stepper.b, out))
        its.append(it)
        secs.append(sec)
        fbs.append(fb)
    out = {"iterations": its, "seconds": secs, "fallbacks": int(sum(fbs)),
           "last": stepper.ptr(buf, steps % R if own else steps)}
    if true_residuals:
        out["true_residuals"] = trs
    return out


def init_params(seed=0):
    """He-normal hidden layers (fan-in 27 x 6 for the first), zero output
    layer: the untrained network returns its base guess."""
    rng = np.random.default_rng(seed)
    w = np.zeros((_lib.NET_LAYERS, _lib.NET_TAPS, _lib.NET_CH, _lib.NET_CH), np.float32)
    for layer in range(_lib.NET_LAYERS - 1):
        cin = 6 if layer == 0 else _lib.NET_CH
        w[layer, :, :, :cin] = rng.standard_normal((_lib.NET_TAPS, _lib.NET_CH, cin)) * \
            math.sqrt(2.0 / (27 * cin))
    b = np.zeros((_lib.NET_LAYERS, _lib.NET_CH), np.float32)
    return w, b


def fit_scales(stepper, trajectories, steps, extrap):
    """Input scales (x_{n-1}: 1 / max |x|; the increment x_{n-1} - x_{n-2}: 1 /
    its RMS) and the output scale (RMS over free nodes of x_n minus its base
    guess), from the stored states."""
    ctx, N = stepper.ctx, stepper.N
    free = stepper.p.mask == 0
    xmax, sq, sqi, cnt = 0.0, 0.0, 0.0, 0
    for buf in trajectories:
        for n in range(2, steps + 1, max(1, steps // 8)):
            xn = ctx.d2h(stepper.ptr(buf, n), N)
            x1 = ctx.d2h(stepper.ptr(buf, n - 1), N)
            x2 = ctx.d2h(stepper.ptr(buf, n - 2), N)
            xmax = max(xmax, float(np.abs(xn).max()))
            d = (xn - x1 - extrap * (x1 - x2))[free]
            inc = (x1 - x2)[free]
            sq += float(d @ d)
            sqi += float(inc @ inc)
            cnt += d.size
    xs = 1.0 / max(xmax, 1e-30)
    si = 1.0 / max(math.sqrt(sqi / max(cnt, 1)), 1e-30)
    return np.array([xs, si, 1.0, 1.0, 1.0, 1.0], np.float32), math.sqrt(sq / max(cnt, 1))


def train(net, stepper, seeds, trajectories, steps, epochs=3, lr=5e-4, shuffle_seed=0,
          log=None):
    """Adam over every (seed, step) sample for `epochs` passes; the learning
    rate warms up linearly over the first 5 % of the updates, then decays on a
    cosine to a tenth; returns the mean loss of each epoch."""
    rng = np.random.default_rng(shuffle_seed)
    samples = [(i, n) for i in range(len(seeds)) for n in range(1, steps + 1)]
    helices = [helix(s, steps) for s in seeds]
    total = epochs * len(samples)
    k, hist = 0, []
    for ep in range(epochs):
        order = rng.permutation(len(samples))
        acc = 0.0
        for j in order:
            i, n = samples[j]
            c, sigma = helices[i]
            stepper.source(c[n - 1], sigma)
            buf = trajectories[i]
            x1, x2, tgt = (stepper.ptr(buf, n - 1), stepper.ptr(buf, max(n - 2, 0)),
                           stepper.ptr(buf, n))
            warm = min(1.0, (k + 1) / max(1.0, 0.05 * total))
            rate = lr * warm * (0.1 + 0.9 * 0.5 * (1 + math.cos(math.pi * k / total)))
            acc += net.train_step_dev(x1, x2, stepper.src, tgt, rate)
            k += 1
        hist.append(acc / len(samples))
        if log:
            log(f"epoch {ep}: mean loss {hist[-1]:.4e}")
    return hist


def experiment(pkg, shape=(128, 128, 64), dt=0.01, steps=200, train_seeds=range(1, 11),
               eval_seed=0, epochs=3, lr=5e-4, extrap=1.0, log=None, params=None,
               loss_kind=1, smoothing=2,
               modes=("previous", "extrap", "extrap+jacobi", "net", "net+jacobi"), tune=False,
               keep_final=False):
    """Generate training trajectories with the plain method, train the network
    (unless params are given) on the initial-residual loss (loss_kind 1) or the
    increment loss (0), then run the evaluation seed with the three guesses.
    tune: run the autotuner once on the implicit-Euler system first (the
    owner tile and z chunk every solve then uses).  keep_final: the last
    state of every mode (host array) in res["finals"][mode].  Returns a dict
    of results and the trained parameters."""
    p = pkg.problem.unit_box(*shape, vel_seed=0, lam=1e-3, scheme=_lib.SCHEME_EULER, dt=dt)
    ctx = pkg.Context.from_problem(p)
    st = Stepper(ctx, p, dt)
    if tune:
        res_tune = ctx.autotune(_lib.FP64, reps=5)
    net = _lib.Net(ctx)
    vx, vy, vz = nodal_velocity(p)
    res = {"shape": list(shape), "nodes": p.N, "dt": dt, "steps": steps}
    if tune:
        res["autotune"] = res_tune
    t0 = time.time()
    trajs = []
    if params is None:
        for s in train_seeds:
            buf = st.alloc_states(steps + 1)
            run(st, s, steps, "previous", states=buf)
            trajs.append(buf)
        res["generate_s"] = round(time.time() - t0, 2)
        in_scale, out_scale = fit_scales(st, trajs, steps, extrap)
        w, b = init_params()
        net.set_params(w, b)
        net.set_static(vx, vy, vz, in_scale, out_scale, extrap)
        net.set_loss(loss_kind)
        net.set_smoothing(smoothing)
        t1 = time.time()
        res["loss_per_epoch"] = train(net, st, list(train_seeds), trajs, steps, epochs, lr,
                                      log=log)
        res["train_s"] = round(time.time() - t1, 2)
        w, b = net.get_params()
        params = {"w": w, "b": b, "in_scale": in_scale, "out_scale": out_scale,
                  "extrap": extrap, "smoothing": smoothing}
    else:
        net.set_params(params["w"], params["b"])
        net.set_static(vx, vy, vz, params["in_scale"], float(params["out_scale"]),
                       float(params["extrap"]))
        net.set_smoothing(int(params["smoothing"]))
    res["training_samples"] = len(list(train_seeds)) * steps
    if keep_final:
        res["finals"] = {}
    for mode in modes:
        r = run(st, eval_seed, steps, mode, net)
        last = r.pop("last")
        if keep_final:
            res["finals"][mode] = ctx.d2h(last, p.N)
        res[mode] = {"iterations": int(sum(r["iterations"])),
                     "seconds": round(float(sum(r["seconds"])), 4),
                     "fallbacks": r["fallbacks"],
                     "per_step_iterations": r["iterations"]}
        if log:
            log(f"{mode}: {res[mode]['iterations']} iterations, {res[mode]['seconds']:.3f} s")
    prev = res["previous"]
    for mode in modes:
        if mode == "previous":
            continue
        res[f"{mode}_speedup_vs_previous"] = round(prev["seconds"] / res[mode]["seconds"], 3)
        res[f"{mode}_iteration_ratio"] = round(res[mode]["iterations"] / prev["iterations"], 3)
    net.close()
    ctx.close()
    return res, params


def save_states(path, stepper, buf, count):
    """Binary trajectory store: `count` device state vectors of `buf` as an
    .npy array of shape (count, N), FP64, written one state at a time."""
    arr = np.lib.format.open_memmap(path, mode="w+", dtype=np.float64,
                                    shape=(count, stepper.N))
    for i in range(count):
        arr[i] = stepper.ctx.d2h(stepper.ptr(buf, i), stepper.N)
    arr.flush()
    del arr


def load_states(path, stepper):
    """Upload a stored trajectory into a new device buffer; returns (buf, count)."""
    arr = np.load(path, mmap_mode="r")
    if arr.ndim != 2 or arr.shape[1] != stepper.N:
        raise ValueError(f"{path}: states of {arr.shape[1:]} nodes, mesh has {stepper.N}")
    buf = stepper.ctx.dev_alloc(8 * stepper.N * arr.shape[0])
    for i in range(arr.shape[0]):
        stepper.ctx.h2d(stepper.ptr(buf, i), np.ascontiguousarray(arr[i]))
    return buf, arr.shape[0]


def save_params(path, params):
    np.savez(path, w=params["w"], b=params["b"], in_scale=params["in_scale"],
             out_scale=np.float64(params["out_scale"]), extrap=np.float64(params["extrap"]),
             smoothing=np.int64(params["smoothing"]))


def load_params(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def main(argv=None):
    """Train on seeds 1..10 and evaluate on seed 0 (128x128x64, dt 0.01, 200
    steps); writes the parameters and a JSON summary."""
    import argparse
    import json
    import os
    import sys
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--shape", default="128,128,64")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--train-seeds", type=int, default=10)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--smoothing", type=int, default=2)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(__file__), "data",
                                                  "learned_guess_128x128x64.npz"))
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    import importlib
    pkg = importlib.import_module(__package__)
    shape = tuple(int(v) for v in a.shape.split(","))
    log = lambda m: print(f"[learned] {m}", file=sys.stderr, flush=True)  # noqa: E731
    res, params = experiment(pkg, shape=shape, steps=a.steps,
                             train_seeds=range(1, 1 + a.train_seeds), epochs=a.epochs,
                             lr=a.lr, log=log, smoothing=a.smoothing)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    save_params(a.out, params)
    text = json.dumps(res)
    if a.json:
        with open(a.json, "w") as f:
            f.write(text + "\n")
    print(text)


if __name__ == "__main__":
    main()

