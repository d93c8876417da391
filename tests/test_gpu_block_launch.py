"""The solver launches in their two addressing forms.  The BiCG and BiCGSTAB
engines carve their vectors from one allocation and their owner launches
address them through one buffer resource (the form every engine below 2^31
bytes of vectors takes); larger engines use a resource per vector.  Both
forms do the same arithmetic in the same order, so their solves are bitwise
equal; VGPU_NO_BLOCK_LAUNCH=1 forces the per-vector form on these small
meshes.  Captured graphs are reused per context, so each form gets its own
contexts."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def _solves(pkg, p, zchunk, tile):
    out = {}
    with pkg.Context.from_problem(p) as ctx:
        ctx.set_options("owner", zchunk, tile)
        b = ctx.assemble_rhs(p.q_in)
        for method, precision in (("bicg", pkg.FP64), ("bicgstab", pkg.FP64),
                                  ("bicg", pkg.FP32), ("bicg", pkg.MIXED)):
            ctx.set_method(method)
            rtol = 1e-5 if precision == pkg.FP32 else 1e-9
            x, it, err, rc = ctx.solve(b, rtol=rtol, itmax=5000, precision=precision)
            out[(method, precision)] = (x, it, err, rc)
        ctx.set_method("bicg")
    return out


@pytest.mark.parametrize("zchunk,tile", [(0, 0), (5, 1)])
def test_block_and_per_vector_launches_agree(pkg, gpu, monkeypatch, zchunk, tile):
    p = pkg.problem.unit_box(30, 22, 14, vel_seed=3, lam=2e-3)
    log(f"block launches zchunk {zchunk} tile {tile}")
    blk = _solves(pkg, p, zchunk, tile)
    monkeypatch.setenv("VGPU_NO_BLOCK_LAUNCH", "1")
    log("per-vector launches")
    per = _solves(pkg, p, zchunk, tile)
    for key, (x, it, err, rc) in blk.items():
        x2, it2, err2, rc2 = per[key]
        log(f"  {key}: iterations {it} / {it2}, err {err:.3e} / {err2:.3e}")
        assert rc == 0 and rc2 == 0, key
        assert it == it2 and err == err2 and np.array_equal(x, x2), key


def test_block_and_per_vector_batches_agree(pkg, gpu, monkeypatch):
    probs = [pkg.problem.unit_box(14, 11, 9, vel_seed=s, lam=2e-3) for s in (1, 2)] + \
            [pkg.problem.unit_box(12, 10, 8, vel_seed=4, lam_range=(1e-4, 1e-2), lam_seed=3)]

    def run():
        ctxs = [pkg.Context.from_problem(q) for q in probs]
        try:
            for c in ctxs:
                c.set_options("owner", 4, 1)
            bs = [c.assemble_rhs(q.q_in) for c, q in zip(ctxs, probs)]
            res = {}
            with pkg.Batch(ctxs) as bt:
                for prec in (pkg.FP64, pkg.MIXED):
                    res[prec] = bt.solve(bs, rtol=1e-9, itmax=5000, precision=prec)
            return res
        finally:
            for c in ctxs:
                c.close()

    log("batch, block launches")
    a = run()
    monkeypatch.setenv("VGPU_NO_BLOCK_LAUNCH", "1")
    log("batch, per-vector launches")
    b = run()
    for prec in a:
        xs, its, errs, st, rc = a[prec]
        xs2, its2, errs2, st2, rc2 = b[prec]
        log(f"  precision {prec}: iterations {list(its)} / {list(its2)}")
        assert rc == 0 and rc2 == 0
        assert np.array_equal(its, its2) and np.array_equal(errs, errs2)
        assert all(np.array_equal(x, y) for x, y in zip(xs, xs2))
