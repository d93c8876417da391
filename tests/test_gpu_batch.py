"""Autotuner and batched solves on the GPU.

* Owner-mode results do not depend on the tile shape or z chunk (bitwise):
  every owner node is summed by one thread in a fixed order, and the dot
  partials are indexed by (tile, plane) whatever the traversal order.
* The autotuner keeps results bitwise and its JSON cache round-trips.
* A batch member's solution, iteration count and recursive error are bitwise
  those of its context's own solve (same tile shape), for uneven subdomains
  and for many right-hand sides sharing one mesh, in FP64 and FP32.
"""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def _ctx(pkg, p):
    return pkg.Context.from_problem(p)


def test_owner_apply_bitwise_across_tiles_and_chunks(pkg):
    p = pkg.problem.unit_box(37, 21, 13, lam=2e-3, vel_seed=3)
    ctx = _ctx(pkg, p)
    x = np.random.default_rng(5).standard_normal(p.N)
    for prec in (pkg.FP64, pkg.FP32):
        ref = None
        for tile in (0, 1, 2):
            for zc in (0, 1, 3, 16):
                log(f"apply prec {prec} tile {tile} zc {zc}")
                ctx.set_options("owner", zc, tile)
                y = (ctx.apply(x, precision=prec), ctx.apply(x, precision=prec, transpose=True))
                if ref is None:
                    ref = y
                else:
                    assert np.array_equal(y[0], ref[0]) and np.array_equal(y[1], ref[1]), \
                        (prec, tile, zc)
    ctx.close()


def test_autotune_keeps_results_and_cache_roundtrips(pkg, tmp_path):
    p = pkg.problem.unit_box(40, 40, 20, lam=1e-3)
    ctx = _ctx(pkg, p)
    x = np.random.default_rng(6).standard_normal(p.N)
    ctx.set_options("owner", 0, 0)
    y0 = ctx.apply(x)
    cache = str(tmp_path / "tune.json")
    log("autotune")
    r1 = pkg.autotune.tune(ctx, pkg.FP64, cache_path=cache, reps=5)
    cand = {0, 4, 8, 16} | {-(-p.naz // k) for k in range(1, 9)}
    assert not r1["cached"] and r1["tile"] in (0, 1, 2) and r1["zchunk"] in cand
    assert r1["ms"] > 0
    assert np.array_equal(ctx.apply(x), y0)
    text = open(cache).read()
    entry = json.loads(text)
    assert list(entry) == [pkg.autotune.cache_key(ctx, pkg.FP64)]
    ctx.set_options("owner", 8, 2)
    r2 = pkg.autotune.tune(ctx, pkg.FP64, cache_path=cache)
    assert r2["cached"] and (r2["tile"], r2["zchunk"]) == (r1["tile"], r1["zchunk"])
    assert open(cache).read() == text
    assert np.array_equal(ctx.apply(x), y0)
    ctx.close()


def _single_solves(pkg, ctxs, bs, prec, rtol=1e-8):
    out = []
    for c, b in zip(ctxs, bs):
        x, it, err, rc = c.solve(b, rtol=rtol, itmax=5000, precision=prec,
                                 raise_on_fail=False)
        out.append((x, it, err, rc))
    return out


@pytest.mark.parametrize("prec,zc", [("fp64", 0), ("fp32", 0), ("fp64", 5), ("fp32", 7)])
def test_batch_uneven_subdomains_bitwise(pkg, prec, zc):
    """zc > 0: z-chunk launches (one workgroup per (chunk, tile) group, the
    batch's groups found through the members' group prefix sums)."""
    precision = pkg.FP64 if prec == "fp64" else pkg.FP32
    rtol = 1e-8 if prec == "fp64" else 1e-5
    probs = pkg.problem.batch_subdomain_problems(count=5, lo=8, hi=22)
    ctxs = [_ctx(pkg, p) for p in probs]
    for c in ctxs:
        c.set_options("owner", zc, 1)
    bs = [c.assemble_rhs(p.q_in) for c, p in zip(ctxs, probs)]
    log("single solves " + " ".join(p.name for p in probs))
    ref = _single_solves(pkg, ctxs, bs, precision, rtol)
    log("batch solve")
    with pkg.Batch(ctxs) as bt:
        xs, its, errs, st, rc = bt.solve(bs, rtol=rtol, itmax=5000, precision=precision,
                                         raise_on_fail=False)
    if prec == "fp64":
        assert rc == pkg.OK and list(st) == [1] * len(ctxs)
    for m, (x, it, err, rcm) in enumerate(ref):
        assert (st[m] == 1) == (rcm == pkg.OK), (m, st[m], rcm)
        assert its[m] == it, (m, its[m], it)
        assert errs[m] == err, m
        assert np.array_equal(xs[m], x), (m, np.abs(xs[m] - x).max())
    assert len(set(int(i) for i in its)) > 1  # systems stopped independently
    for c in ctxs:
        c.close()


def test_batch_many_rhs_one_mesh_bitwise(pkg):
    probs = pkg.problem.batch_rhs_problems(shape=(24, 24, 12), count=8)
    ctxs = [_ctx(pkg, p) for p in probs]
    bs = [c.assemble_rhs(p.q_in) for c, p in zip(ctxs, probs)]
    ref = _single_solves(pkg, ctxs, bs, pkg.FP64)
    log("batch of 8 right-hand sides")
    with pkg.Batch(ctxs) as bt:
        off, total = bt.layout()
        assert off[0] == 0 and all(o % 256 == 0 for o in off) and total >= sum(c.N for c in ctxs)
        for rep in range(2):  # second solve replays the captured graph
            xs, its, errs, st, rc = bt.solve(bs, rtol=1e-8, itmax=5000)
            for m, (x, it, err, _) in enumerate(ref):
                assert its[m] == it and np.array_equal(xs[m], x), (rep, m)
        # coefficient change of one member is picked up at the next solve
        p0 = probs[0]
        ctxs[0].set_coeffs(p0.lam * 2.0, p0.vel, p0.P, p0.alpha, p0.variant, p0.scheme, p0.dt)
        x0, it0, err0, _ = ctxs[0].solve(bs[0], rtol=1e-8, itmax=5000)
        xs, its, errs, st, rc = bt.solve(bs, rtol=1e-8, itmax=5000)
        assert its[0] == it0 and np.array_equal(xs[0], x0)
    for c in ctxs:
        c.close()


def test_batch_matches_oracle_solution(pkg, po):
    probs = pkg.problem.batch_subdomain_problems(count=3, lo=6, hi=10)
    ctxs = [_ctx(pkg, p) for p in probs]
    bs, refs = [], []
    for c, p in zip(ctxs, probs):
        bb, f, d = po.assemble(p)
        b, d, x0 = po.impose_mask(p.mask, f, d, np.zeros(p.N))
        bs.append(c.assemble_rhs(p.q_in))
        assert np.linalg.norm(bs[-1] - b) <= 1e-13 * np.linalg.norm(b)
        xo, ito, erro, rco = po.dbcg_simple(b, x0, 1e-10, 5000, bb, p.iss, d, p.nax, p.nay,
                                            p.naz, p.mask)
        refs.append(xo)
    with pkg.Batch(ctxs) as bt:
        xs, its, errs, st, rc = bt.solve(bs, rtol=1e-10, itmax=5000)
    for m in range(len(ctxs)):
        r = np.linalg.norm(xs[m] - refs[m]) / np.linalg.norm(refs[m])
        assert r <= 1e-7, (m, r)
    for c in ctxs:
        c.close()


def test_fallback_finishes_unconverged_members(pkg, gpu):
    """vgpu_batch_set_fallback: members that the batched BiCG leaves at itmax
    are finished by BiCGSTAB on their own contexts from their current
    iterate and reach the solution; members that converge in the batch with a
    true residual within rtol are untouched (bitwise the result without the
    fallback), the others end with a true residual within rtol."""
    probs = [pkg.problem.unit_box(14, 11, 9, vel_seed=s, lam=2e-3) for s in (1, 2, 3)]
    ctxs = [_ctx(pkg, q) for q in probs]
    bs = [c.assemble_rhs(q.q_in) for c, q in zip(ctxs, probs)]
    with pkg.Batch(ctxs) as bt:
        log("batch fallback: itmax 3 without the fallback")
        _, it0, _, st0, rc0 = bt.solve(bs, rtol=1e-10, itmax=3, raise_on_fail=False)
        assert (st0 == 2).all() and (it0 == 3).all() and rc0 == 5  # VGPU_NOT_CONVERGED
        bt.set_fallback("bicgstab", itmax=5000)
        log("batch fallback: itmax 3 with the BiCGSTAB fallback")
        xs, its, ers, st, rc = bt.solve(bs, rtol=1e-10, itmax=3)
        assert rc == 0 and (st == 1).all() and (its > 3).all() and (ers <= 1e-10).all()
        for c, b, x in zip(ctxs, bs, xs):
            xr, _, _, _ = c.solve(b, rtol=1e-12, itmax=20000)
            assert c.true_residual(b, x) <= 1e-8
            assert np.linalg.norm(x - xr) <= 1e-7 * np.linalg.norm(xr)
        log("batch fallback: converged members untouched")
        xa, ita, era, sta, _ = bt.solve(bs, rtol=1e-8, itmax=20000)
        bt.set_fallback(None)
        xb, itb, erb, stb, _ = bt.solve(bs, rtol=1e-8, itmax=20000)
        assert (sta == 1).all() and (stb == 1).all()
        for m, (c, b) in enumerate(zip(ctxs, bs)):
            tb = c.true_residual(b, xb[m])
            log(f"  member {m}: BiCG true residual {tb:.2e}")
            assert c.true_residual(b, xa[m]) <= 1e-8
            if tb <= 1e-8:
                assert ita[m] == itb[m] and era[m] == erb[m] and np.array_equal(xa[m], xb[m])
        with pytest.raises(pkg.VgpuError):
            bt.set_fallback("bicgstab", itmax=-1)
    for c in ctxs:
        c.close()


@pytest.mark.parametrize("scatter", ["colored", "atomic"])
def test_brick_autotune_keeps_parity_and_cache_roundtrips(pkg, po, tmp_path, scatter):
    """The brick tuner of the colored / atomic scatter picks a brick shape and
    depth among the candidates; the operator with the choice still matches
    the oracle to 1e-13, the context then solves with that scatter, colored
    results are bitwise repeatable, and the JSON cache round-trips."""
    p = pkg.problem.unit_box(40, 24, 20, lam=1e-3, vel_seed=2)
    bb, b, d = po.assemble(p)
    ctx = _ctx(pkg, p)
    x = np.random.default_rng(8).standard_normal(p.N)
    cache = str(tmp_path / "tune.json")
    log(f"brick autotune {scatter}")
    r = pkg.autotune.tune(ctx, pkg.FP64, cache_path=cache, reps=3, scatter=scatter)
    assert not r["cached"] and r["tile"] in (0, 1, 2)
    assert r["zchunk"] in (0, 1, 2, 3, 4, 6, 8, 12, 16) and r["ms"] > 0
    for tr in (False, True):
        ref = po.matvec_0(x, p.iss, bb, p.nax, p.nay, p.naz, p.mask, tr)
        y = ctx.apply(x, scatter=scatter, transpose=tr)
        assert np.linalg.norm(y - ref) <= 1e-13 * np.linalg.norm(ref)
    if scatter == "colored":
        assert np.array_equal(ctx.apply(x, scatter=scatter), ctx.apply(x, scatter=scatter))
    bo, do, x0 = po.impose_mask(p.mask, b, d, np.zeros(p.N))
    xs, it, err, rc = ctx.solve(bo, rtol=1e-10, itmax=5000)  # the tuned scatter
    assert rc == 0 and err <= 1e-10
    r2 = pkg.autotune.tune(ctx, pkg.FP64, cache_path=cache, scatter=scatter)
    assert r2["cached"] and (r2["tile"], r2["zchunk"]) == (r["tile"], r["zchunk"])
    entry = json.loads(open(cache).read())
    assert any(k.endswith("|" + scatter) for k in entry)
    ctx.close()
