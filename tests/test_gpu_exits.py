"""Exits of the solvers that a time stepper meets, through the C ABI:

* b = 0: x = 0 solves A x = b exactly from any start, while the relative
  residual |r| / |b| is 0/0 or |r|/0 (the reference's dbcg_simple,
  VANADIS.FOR:977-1055, divides by |b| = 0 and its loop leaves NaN in x).
  Every solve path returns x = 0, 0 iterations, err 0 and success: BiCG and
  BiCGSTAB in FP64 and FP32, mixed precision, batch members, the legacy
  drop-in and its allocation-free form;
* itmax = 0: not converged, x the masked start;
* a start that already meets rtol: 0 iterations, x bitwise the start;
* a system on which plain Jacobi-BiCG diverges (a 256x256x8 unit box, 32:1
  element aspect; the oracle's dbcg_simple reaches err 4.5e11 after 4000
  iterations): BiCG reports it has not converged, BiCGSTAB converges (the
  oracle's in 499 iterations)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def sys1(pkg, gpu):
    p = pkg.problem.unit_box(12, 10, 8, vel_seed=3, lam=3e-3)
    ctx = pkg.Context.from_problem(p)
    b = ctx.assemble_rhs(p.q_in)
    yield p, ctx, b
    ctx.close()


MODES = [("bicg", "FP64"), ("bicg", "FP32"), ("bicgstab", "FP64"), ("bicgstab", "FP32"),
         ("bicg", "MIXED")]


@pytest.mark.parametrize("method,prec", MODES)
@pytest.mark.parametrize("start", ["zero", "random"])
def test_zero_rhs_gives_zero_solution(pkg, sys1, method, prec, start):
    p, ctx, _ = sys1
    x0 = None if start == "zero" else np.random.default_rng(5).standard_normal(p.N)
    ctx.set_method(method)
    try:
        x, it, err, rc = ctx.solve(np.zeros(p.N), x0, rtol=1e-8, itmax=500,
                                   precision=getattr(pkg, prec), raise_on_fail=False)
    finally:
        ctx.set_method("bicg")
    log(f"b = 0, {method} {prec}, start {start}: rc {rc}, iterations {it}, err {err}")
    assert rc == pkg.OK and it == 0 and err == 0.0
    assert np.all(x == 0.0)


@pytest.mark.parametrize("prec", ["FP64", "FP32", "MIXED"])
def test_zero_rhs_batch_member(pkg, gpu, prec):
    """A batch with one zero right-hand side: that member x = 0, status 1,
    0 iterations, err 0; the other member bitwise its single solve."""
    probs = [pkg.problem.unit_box(12, 10, 8, vel_seed=3, lam=3e-3),
             pkg.problem.unit_box(10, 9, 7, vel_seed=4, lam=3e-3)]
    ctxs = [pkg.Context.from_problem(q) for q in probs]
    try:
        for c in ctxs:
            c.set_options("owner", 3, 1)
        precision = getattr(pkg, prec)
        rtol = 1e-5 if prec == "FP32" else 1e-8
        bs = [np.zeros(probs[0].N), ctxs[1].assemble_rhs(probs[1].q_in)]
        x0s = [np.random.default_rng(6).standard_normal(probs[0].N), np.zeros(probs[1].N)]
        x1, it1, e1, rc1 = ctxs[1].solve(bs[1], rtol=rtol, itmax=2000, precision=precision)
        with pkg.Batch(ctxs) as bt:
            xs, its, errs, st, rc = bt.solve(bs, x0s, rtol=rtol, itmax=2000,
                                             precision=precision, raise_on_fail=False)
        log(f"batch b = 0 member, {prec}: rc {rc}, status {list(st)}, iterations {list(its)}")
        assert rc == pkg.OK and list(st) == [1, 1]
        assert its[0] == 0 and errs[0] == 0.0 and np.all(xs[0] == 0.0)
        assert its[1] == it1 and errs[1] == e1 and np.array_equal(xs[1], x1)
    finally:
        for c in ctxs:
            c.close()


@pytest.mark.parametrize("scatter", ["colored", "atomic"])
def test_zero_rhs_legacy_drop_in(pkg, po, gpu, scatter):
    p = pkg.problem.unit_box(12, 10, 8, vel_seed=3, lam=3e-3)
    bb, b, d = po.assemble(p)
    _, d, _ = po.impose_mask(p.mask, b, d, np.zeros(p.N))
    x0 = np.random.default_rng(7).standard_normal(p.N)
    x, it, err = pkg.kernel_dbcg_simple(p.N, np.zeros(p.N), x0, 1e-8, 500, bb, p.iss, d, p.nax,
                                        p.nay, p.naz, p.mask, scatter=scatter)
    log(f"legacy b = 0 {scatter}: iterations {it}, err {err}")
    assert it == 0 and err == 0.0 and np.all(x == 0.0)
    # the allocation-free form (device buffers of a context used as the
    # caller's allocator)
    ctx = pkg.Context.from_problem(p, coeffs=False)
    try:
        t = {}
        for k, v in (("b", np.zeros(p.N)), ("x", x0), ("bb", bb),
                     ("iss", p.iss.astype(np.int32)), ("d", d),
                     ("bc", p.mask.astype(np.int32))):
            v = np.ascontiguousarray(v)
            t[k] = ctx.dev_alloc(v.nbytes)
            ctx.h2d(t[k], v)
        itn, errn = C.c_int(-1), C.c_double(-1.0)
        code = {"colored": pkg.SCATTER_COLORED, "atomic": pkg.SCATTER_ATOMIC}[scatter]
        rc = pkg.load().kernel_dbcg_simple_noalloc(
            p.N, t["b"], t["x"], 1e-8, 500, C.byref(itn), C.byref(errn), t["bb"], t["iss"],
            t["d"], p.nax, p.nay, p.naz, t["bc"], code, None)
        xn = ctx.d2h(t["x"], p.N)
    finally:
        ctx.close()
    log(f"noalloc b = 0 {scatter}: rc {rc}, iterations {itn.value}, err {errn.value}")
    assert rc == pkg.OK and itn.value == 0 and errn.value == 0.0 and np.all(xn == 0.0)


@pytest.mark.parametrize("method,prec", MODES)
def test_itmax_zero_and_converged_start(pkg, sys1, method, prec):
    p, ctx, b = sys1
    precision = getattr(pkg, prec)
    rtol = 1e-5 if prec == "FP32" else 1e-8
    x0 = np.random.default_rng(8).standard_normal(p.N)
    ctx.set_method(method)
    try:
        x, it, err, rc = ctx.solve(b, x0, rtol=rtol, itmax=0, precision=precision,
                                   raise_on_fail=False)
        log(f"itmax 0, {method} {prec}: rc {rc}, iterations {it}, err {err:.3e}")
        assert rc == pkg.NOT_CONVERGED and it == 0 and err > rtol
        xm = x0 * (1 - p.mask)
        if prec == "FP32":  # the start rounded to FP32 and back
            assert np.array_equal(x, xm.astype(np.float32).astype(np.float64))
        else:
            assert np.array_equal(x, xm)
        # a start that meets rtol (the FP64 solution to 1e-12): no iteration
        xs, _, _, _ = ctx.solve(b, rtol=1e-12, itmax=5000, precision=pkg.FP64)
        x2, it2, err2, rc2 = ctx.solve(b, xs, rtol=rtol, itmax=100, precision=precision)
        log(f"  from the solution: iterations {it2}, err {err2:.3e}")
        assert rc2 == pkg.OK and it2 == 0 and err2 <= rtol
        if prec != "FP32":
            assert np.array_equal(x2, xs)
    finally:
        ctx.set_method("bicg")


def test_diverging_bicg_reports_and_bicgstab_converges(pkg, gpu):
    p = pkg.problem.unit_box(256, 256, 8, vel_seed=0, lam=1e-3)
    ctx = pkg.Context.from_problem(p)
    try:
        b = ctx.assemble_rhs(p.q_in)
        x, it, err, rc = ctx.solve(b, rtol=1e-8, itmax=4000, precision=pkg.FP64,
                                   raise_on_fail=False)
        log(f"256x256x8 BiCG: rc {rc}, iterations {it}, err {err:.3e}")
        assert rc in (pkg.NOT_CONVERGED, pkg.BREAKDOWN) and not err <= 1e-8
        ctx.set_method("bicgstab")
        x, it, err, rc = ctx.solve(b, rtol=1e-8, itmax=4000, precision=pkg.FP64)
        tr = ctx.true_residual(b, x)
        log(f"256x256x8 BiCGSTAB: rc {rc}, iterations {it}, err {err:.3e}, true {tr:.3e}")
        assert rc == pkg.OK and err <= 1e-8 and tr <= 1e-7
    finally:
        ctx.set_method("bicg")
        ctx.close()
