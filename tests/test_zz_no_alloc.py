"""The allocation-free drop-in (kernel_dbcg_simple_noalloc) against the legacy
drop-in kernel_dbcg_simple_ on the same inputs: bitwise the same solution,
iterations and error with the colored scatter, within the solver tolerance
with the atomic one; two calls in a row, the second on a smaller system (the
fixed scratch array keeps the first call's contents); sizes above the array's
limit are refused on the host."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def inputs(pkg, po, p):
    bb, b, d = po.assemble(p)
    b, d, x0 = po.impose_mask(p.mask, b, d, np.zeros(p.N))
    return bb, b, d, x0


def run_noalloc(pkg, p, bb, b, d, x0, scatter, tol=1e-8, itmax=3000):
    """Device copies of the inputs (buffers of a context used only as the
    caller's allocator), the no-alloc call on the null stream, x back."""
    ctx = pkg.Context.from_problem(p, coeffs=False)
    try:
        arrs = (("b", b), ("x", x0.copy()), ("bb", bb), ("iss", p.iss.astype(np.int32)),
                ("d", d), ("bc", p.mask.astype(np.int32)))
        t = {}
        for k, v in arrs:
            v = np.ascontiguousarray(v)
            t[k] = ctx.dev_alloc(v.nbytes)
            ctx.h2d(t[k], v)
        it, err = C.c_int(-1), C.c_double(-1.0)
        code = {"colored": pkg.SCATTER_COLORED, "atomic": pkg.SCATTER_ATOMIC}[scatter]
        rc = pkg.load().kernel_dbcg_simple_noalloc(
            p.N, t["b"], t["x"], tol, itmax, C.byref(it), C.byref(err), t["bb"], t["iss"],
            t["d"], p.nax, p.nay, p.naz, t["bc"], code, None)
        x = ctx.d2h(t["x"], p.N)
        return x, it.value, err.value, rc
    finally:
        ctx.close()


@pytest.mark.parametrize("scatter", ["colored", "atomic"])
def test_noalloc_matches_legacy_two_calls(pkg, po, gpu, scatter):
    mx = C.c_int64(0)
    assert pkg.load().vgpu_noalloc_max_nodes(C.byref(mx)) == 0 and mx.value == 1 << 18
    for dims in ((24, 20, 12), (10, 8, 6)):  # the second call smaller than the first
        p = pkg.problem.unit_box(*dims, vel_seed=4, lam=2e-3)
        bb, b, d, x0 = inputs(pkg, po, p)
        print(f"noalloc {scatter} {dims}: N {p.N}", flush=True)
        xr, itr, errr = pkg.kernel_dbcg_simple(p.N, b, x0, 1e-8, 3000, bb, p.iss, d, p.nax,
                                               p.nay, p.naz, p.mask, scatter=scatter)
        xn, itn, errn, rc = run_noalloc(pkg, p, bb, b, d, x0, scatter)
        print(f"  legacy {itr} iterations err {errr:.3e}; noalloc rc {rc} {itn} iterations "
              f"err {errn:.3e}", flush=True)
        assert rc == pkg.OK and errn <= 1e-8
        if scatter == "colored":  # deterministic: the same bits
            assert itn == itr and errn == errr
            assert np.array_equal(xn, xr)
        else:  # atomic sum order varies run to run
            assert abs(itn - itr) <= max(3, itr // 20)
            assert np.linalg.norm(xn - xr) <= 1e-8 * np.linalg.norm(xr)


def test_noalloc_refuses_sizes_above_its_array(pkg, gpu):
    it, err = C.c_int(0), C.c_double(0.0)
    nax, nay, naz = 65, 65, 65  # 274,625 nodes > 2^18
    rc = pkg.load().kernel_dbcg_simple_noalloc(
        nax * nay * naz, 1, 1, 1e-8, 10, C.byref(it), C.byref(err), 1, 1, 1, nax, nay, naz, 1,
        pkg.SCATTER_COLORED, None)
    assert rc == pkg.ERR_ARG
    assert b"exceed" in pkg.load().vgpu_last_error()
