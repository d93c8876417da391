"""Learned-guess network on the GPU against the CPU oracle (oracle/netoracle.c,
pinned by finite differences in test_net_oracle.py), and the gain of the
trained network over the previous-step guess on implicit-Euler stepping."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L, T, C = 5, 27, 16


def log(msg):
    print(msg, flush=True)


def _setup(pkg, shape, seed):
    p = pkg.problem.unit_box(*shape, lam=1e-3)
    ctx = pkg.Context.from_problem(p)
    net = pkg.Net(ctx)
    rng = np.random.default_rng(seed)
    vec = {k: rng.standard_normal(p.N) for k in ("x1", "x2", "src", "vx", "vy", "vz", "tgt")}
    dev = {}
    for k in ("x1", "x2", "src", "tgt"):
        dev[k] = ctx.dev_alloc(8 * p.N)
        ctx.h2d(dev[k], vec[k])
    dev["guess"] = ctx.dev_alloc(8 * p.N)
    return p, ctx, net, vec, dev


def test_forward_known_answer_shift(pkg, po):
    p, ctx, net, vec, dev = _setup(pkg, (19, 9, 6), 1)
    x1 = np.abs(vec["x1"]) + 0.1
    ctx.h2d(dev["x1"], x1)
    w = np.zeros((L, T, C, C), np.float32)
    for layer in range(L - 1):
        w[layer, 13] = np.eye(C, dtype=np.float32)
    w[L - 1, 14, 0, 0] = 1.0
    b = np.zeros((L, C), np.float32)
    scale = np.array([2.0, 1, 1, 1, 1, 1], np.float32)
    net.set_params(w, b)
    net.set_static(vec["vx"], vec["vy"], vec["vz"], scale, 1.0, 0.0)
    net.predict_dev(dev["x1"], dev["x2"], dev["src"], dev["guess"])
    out = ctx.d2h(dev["guess"], p.N) - x1
    _, _, _, ref = po.net_grads(p.nax, p.nay, p.naz, x1, vec["x2"], vec["src"], vec["vx"],
                                vec["vy"], vec["vz"], vec["tgt"], p.mask, scale, 1.0, w, b,
                                quant=True)
    assert np.max(np.abs(out - ref)) <= 1e-14 * max(1.0, np.abs(ref).max())
    net.close()
    ctx.close()


@pytest.mark.parametrize("extrap", [0.0, 1.0])
def test_gradients_match_oracle(pkg, po, extrap):
    p, ctx, net, vec, dev = _setup(pkg, (20, 17, 9), 2)
    rng = np.random.default_rng(3)
    w = (rng.standard_normal((L, T, C, C)) * 0.12).astype(np.float32)
    b = (rng.standard_normal((L, C)) * 0.05).astype(np.float32)
    scale = np.array([1.0, 0.8, 1.2, 1, 1, 1], np.float32)
    net.set_params(w, b)
    net.set_static(vec["vx"], vec["vy"], vec["vz"], scale, 0.7, extrap)
    loss = net.train_step_dev(dev["x1"], dev["x2"], dev["src"], dev["tgt"], 0.0)
    gw, gb = net.get_grads()
    lo, dw, db, out = po.net_grads(p.nax, p.nay, p.naz, vec["x1"], vec["x2"], vec["src"],
                                   vec["vx"], vec["vy"], vec["vz"], vec["tgt"], p.mask, scale,
                                   0.7, w, b, quant=True, extrap=extrap)
    log(f"loss gpu {loss:.6e} oracle {lo:.6e}")
    assert loss == pytest.approx(lo, rel=2e-3)
    gw = gw.reshape(L, -1)
    dw = dw.reshape(L, -1)
    for layer in range(L):
        r = np.linalg.norm(gw[layer] - dw[layer]) / np.linalg.norm(dw[layer])
        log(f"layer {layer}: |dW gpu - oracle| / |dW| = {r:.2e}")
        assert r <= 5e-3, (layer, r)  # measured 5e-5 (layer 4) .. 2.0e-3 (layer 0)
    rb = np.linalg.norm(gb - db) / np.linalg.norm(db)
    assert rb <= 1e-3, rb  # measured 2.1e-4
    # the prediction is the oracle's forward pass
    net.predict_dev(dev["x1"], dev["x2"], dev["src"], dev["guess"])
    g = ctx.d2h(dev["guess"], p.N)
    base = vec["x1"] + extrap * (vec["x1"] - vec["x2"])
    pred = (g - base) / 0.7
    rp = np.linalg.norm(pred - out) / np.linalg.norm(out)
    log(f"prediction: |out gpu - oracle| / |out| = {rp:.2e}")
    assert rp <= 2e-3  # measured 4.6e-4
    # deterministic: a second pass gives the same gradients bitwise
    net.train_step_dev(dev["x1"], dev["x2"], dev["src"], dev["tgt"], 0.0)
    gw2, gb2 = net.get_grads()
    assert np.array_equal(gw2.reshape(L, -1), gw) and np.array_equal(gb2, gb)
    net.close()
    ctx.close()


def test_layer_chunking_is_bitwise_invisible(pkg, po):
    """The layer kernels march up z over chunks of node planes; the
    prediction, the loss and the gradients are bitwise the same for one plane
    per wave, several (uneven last chunk) and the whole grid in one march,
    and match the oracle's forward pass."""
    p, ctx, net, vec, dev = _setup(pkg, (20, 17, 9), 4)
    rng = np.random.default_rng(6)
    w = (rng.standard_normal((L, T, C, C)) * 0.12).astype(np.float32)
    b = (rng.standard_normal((L, C)) * 0.05).astype(np.float32)
    scale = np.array([1.0, 0.8, 1.2, 1, 1, 1], np.float32)
    net.set_params(w, b)
    net.set_static(vec["vx"], vec["vy"], vec["vz"], scale, 0.7, 1.0)
    res = []
    for zc in (1, 3, 4, 10, 0):
        log(f"net z chunk {zc}")
        net.set_zchunk(zc)
        net.predict_dev(dev["x1"], dev["x2"], dev["src"], dev["guess"])
        g = ctx.d2h(dev["guess"], p.N)
        loss = net.train_step_dev(dev["x1"], dev["x2"], dev["src"], dev["tgt"], 0.0)
        gw, gb = net.get_grads()
        res.append((g, loss, gw, gb))
    for g, loss, gw, gb in res[1:]:
        assert np.array_equal(g, res[0][0]) and loss == res[0][1]
        assert np.array_equal(gw, res[0][2]) and np.array_equal(gb, res[0][3])
    _, _, _, out = po.net_grads(p.nax, p.nay, p.naz, vec["x1"], vec["x2"], vec["src"],
                                vec["vx"], vec["vy"], vec["vz"], vec["tgt"], p.mask, scale,
                                0.7, w, b, quant=True, extrap=1.0)
    base = vec["x1"] + (vec["x1"] - vec["x2"])
    pred = (res[0][0] - base) / 0.7
    assert np.linalg.norm(pred - out) <= 1e-2 * np.linalg.norm(out)
    with pytest.raises(pkg.VgpuError):
        net.set_zchunk(-1)
    for what in range(4):
        assert net.time_layers(what, reps=2) > 0.0
    with pytest.raises(pkg.VgpuError):
        net.time_layers(4, reps=2)
    net.close()
    ctx.close()


def test_residual_loss_gradient_matches_finite_difference(pkg):
    """Initial-residual loss |A (t - g)|^2 / |A (t - base)|^2: its gradient
    w.r.t. the output bias (the sum of dL/dout) against central differences."""
    p = pkg.problem.unit_box(18, 14, 8, lam=1e-3, scheme=pkg.SCHEME_EULER, dt=0.01)
    ctx = pkg.Context.from_problem(p)
    net = pkg.Net(ctx)
    rng = np.random.default_rng(7)
    dev = {}
    for k in ("x1", "x2", "src", "tgt"):
        dev[k] = ctx.dev_alloc(8 * p.N)
        ctx.h2d(dev[k], rng.standard_normal(p.N) * (1 - p.mask))
    w, b = pkg.learned.init_params(2)
    w[L - 1] = rng.standard_normal(w[L - 1].shape).astype(np.float32) * 0.05
    v = rng.standard_normal(p.N)
    net.set_static(v, v, v, np.ones(6, np.float32), 0.3, 1.0)
    net.set_loss(1)
    net.set_smoothing(2)
    net.set_params(w, b)
    l0 = net.train_step_dev(dev["x1"], dev["x2"], dev["src"], dev["tgt"], 0.0)
    _, gb = net.get_grads()
    h = 1e-3
    lp = []
    for sgn in (1, -1):
        bb = b.copy()
        bb[L - 1, 0] += sgn * h
        net.set_params(w, bb)
        lp.append(net.train_step_dev(dev["x1"], dev["x2"], dev["src"], dev["tgt"], 0.0))
    fd = (lp[0] - lp[1]) / (2 * h)
    log(f"residual loss {l0:.5e}: db fd {fd:.5e} kernel {gb[(L - 1) * C]:.5e}")
    assert gb[(L - 1) * C] == pytest.approx(fd, rel=2e-2)
    net.close()
    ctx.close()


def test_adam_steps_reduce_loss(pkg):
    p, ctx, net, vec, dev = _setup(pkg, (16, 16, 8), 4)
    w, b = pkg.learned.init_params(1)
    w[L - 1] = np.random.default_rng(5).standard_normal(w[L - 1].shape) * 0.01
    net.set_params(w, b)
    net.set_static(vec["vx"], vec["vy"], vec["vz"], np.ones(6, np.float32), 1.0, 0.0)
    losses = [net.train_step_dev(dev["x1"], dev["x2"], dev["src"], dev["tgt"], 3e-3)
              for _ in range(40)]
    log(f"loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    assert losses[-1] < 0.9 * losses[0]
    net.close()
    ctx.close()


def test_learned_guess_beats_previous_step(pkg):
    """Small implicit-Euler run: train on 8 seeds x 40 steps, evaluate on the
    unseen seed 0; every step is solved to rtol 1e-8 whatever the guess.  The
    trained guess (filtered, then 2 Jacobi sweeps) needs fewer BiCG
    iterations than the previous-step guess of the plain method."""
    res, params = pkg.learned.experiment(pkg, shape=(32, 32, 16), dt=0.01, steps=40,
                                         train_seeds=range(1, 9), eval_seed=0, epochs=10,
                                         lr=5e-4, log=log)
    for mode in ("previous", "extrap", "extrap+jacobi", "net", "net+jacobi"):
        log(f"{mode}: {res[mode]['iterations']} iterations")
    assert res["loss_per_epoch"][-1] < 0.2 * res["loss_per_epoch"][0]
    assert res["net+jacobi"]["iterations"] < 0.9 * res["previous"]["iterations"]


def test_trajectory_store_round_trip(pkg, tmp_path):
    p = pkg.problem.unit_box(12, 10, 8, lam=1e-3, scheme=pkg.SCHEME_EULER, dt=0.01)
    ctx = pkg.Context.from_problem(p)
    st = pkg.learned.Stepper(ctx, p, 0.01)
    buf = st.alloc_states(6)
    pkg.learned.run(st, 3, 5, "previous", states=buf)
    path = str(tmp_path / "traj.npy")
    pkg.learned.save_states(path, st, buf, 6)
    buf2, n = pkg.learned.load_states(path, st)
    assert n == 6
    for i in range(6):
        a = ctx.d2h(st.ptr(buf, i), p.N)
        assert np.array_equal(a, ctx.d2h(st.ptr(buf2, i), p.N))
        assert np.array_equal(a, np.load(path, mmap_mode="r")[i])
    assert np.abs(ctx.d2h(st.ptr(buf, 5), p.N)).max() > 0
    ctx.close()



def test_warm_starts_and_methods_reach_the_same_states(pkg):
    """20 implicit-Euler steps on a small box from every guess (previous,
    linear, quadratic and cubic extrapolation, with Jacobi sweeps) and both methods:
    each step is solved to rtol 1e-8 of its own right-hand side, so the final
    states agree to the solver tolerance; the extrapolated guesses need fewer
    iterations than the previous-step guess, and quadratic extrapolation
    fewer than linear with BiCGSTAB."""
    p = pkg.problem.unit_box(24, 20, 12, vel_seed=0, lam=1e-3, scheme=pkg.SCHEME_EULER,
                             dt=0.01)
    ctx = pkg.Context.from_problem(p)
    st = pkg.learned.Stepper(ctx, p, 0.01)
    finals, its = {}, {}
    modes = ("previous", "extrap+jacobi", "extrap2+jacobi", "previous@bicgstab",
             "extrap@bicgstab", "extrap2@bicgstab", "extrap2+jacobi@bicgstab",
             "extrap3@bicgstab", "extrap3+jacobi8@bicgstab")
    for mode in modes:
        r = pkg.learned.run(st, 0, 20, mode)
        finals[mode] = ctx.d2h(r["last"], p.N)
        its[mode] = sum(r["iterations"])
        log(f"{mode}: {its[mode]} iterations")
    ref = finals["previous"]
    for mode in modes:
        assert np.linalg.norm(finals[mode] - ref) <= 1e-6 * np.linalg.norm(ref), mode
    assert its["extrap+jacobi"] < its["previous"]
    assert its["extrap2+jacobi"] < its["previous"]
    assert its["extrap@bicgstab"] < its["previous@bicgstab"]
    assert its["extrap2@bicgstab"] <= its["extrap@bicgstab"]
    assert its["extrap3+jacobi8@bicgstab"] < its["extrap2+jacobi@bicgstab"]
    ctx.close()


def test_net_destroy_after_its_context(pkg, gpu):
    """vgpu_net_destroy after vgpu_destroy of the network's context (the order a
    garbage collector may pick): it waits for the device and frees the
    network's own buffers only.  Closing a Context closes its networks first."""
    p = pkg.problem.unit_box(8, 6, 5, lam=1e-3)
    ctx = pkg.Context.from_problem(p)
    net = pkg.Net(ctx)
    assert gpu.vgpu_destroy(ctx.h) == 0
    ctx.h = None
    assert gpu.vgpu_net_destroy(net.h) == 0
    net.h = None
    ctx2 = pkg.Context.from_problem(p)
    net2 = pkg.Net(ctx2)
    ctx2.close()
    assert net2.h is None and ctx2.h is None


@pytest.mark.parametrize("shape", [(1, 1, 1), (3, 1, 2), (1, 17, 1), (33, 2, 1)])
def test_degenerate_grids_match_oracle(pkg, po, shape):
    """The layer kernels on node grids smaller than one tile in one, two or
    three directions (2 nodes in a direction; a 34-node row over three tiles):
    loss, gradients and prediction against the oracle, as on the interior
    grids above."""
    p, ctx, net, vec, dev = _setup(pkg, shape, 7)
    rng = np.random.default_rng(8)
    w = (rng.standard_normal((L, T, C, C)) * 0.12).astype(np.float32)
    b = (rng.standard_normal((L, C)) * 0.05).astype(np.float32)
    scale = np.array([1.0, 0.8, 1.2, 1, 1, 1], np.float32)
    net.set_params(w, b)
    net.set_static(vec["vx"], vec["vy"], vec["vz"], scale, 0.7, 1.0)
    log(f"net on {shape}: N {p.N}")
    loss = net.train_step_dev(dev["x1"], dev["x2"], dev["src"], dev["tgt"], 0.0)
    gw, gb = net.get_grads()
    lo, dw, db, out = po.net_grads(p.nax, p.nay, p.naz, vec["x1"], vec["x2"], vec["src"],
                                   vec["vx"], vec["vy"], vec["vz"], vec["tgt"], p.mask, scale,
                                   0.7, w, b, quant=True, extrap=1.0)
    log(f"  loss gpu {loss:.6e} oracle {lo:.6e}")
    assert loss == pytest.approx(lo, rel=2e-3)
    gw = gw.reshape(L, -1)
    dw = dw.reshape(L, -1)
    for layer in range(L):
        r = np.linalg.norm(gw[layer] - dw[layer]) / max(np.linalg.norm(dw[layer]), 1e-30)
        assert r <= 5e-3, (layer, r)
    assert np.linalg.norm(gb - db) <= 1e-3 * np.linalg.norm(db)
    net.predict_dev(dev["x1"], dev["x2"], dev["src"], dev["guess"])
    g = ctx.d2h(dev["guess"], p.N)
    pred = (g - (2.0 * vec["x1"] - vec["x2"])) / 0.7
    assert np.linalg.norm(pred - out) <= 2e-3 * np.linalg.norm(out)
    net.close()
    ctx.close()
