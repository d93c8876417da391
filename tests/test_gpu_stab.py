"""BiCGSTAB method and the mixed-precision solve on the GPU, through the C ABI,
against the oracle's Jacobi-BiCGSTAB and BiCG (same inputs).

Bars: FP64 BiCGSTAB reaches rtol on its recursive residual, its true residual
stays within 1e-8 at rtol 1e-10, its solution matches the oracle's FP64 BiCG
solution to 1e-7 relative and its iteration count the oracle BiCGSTAB's
within 10 % (BiCGSTAB amplifies summation-order differences, so the iterates
are not bitwise those of the oracle); every tile shape and z chunk; the BiCG
method is unaffected by a BiCGSTAB solve on the same context (bitwise)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


CASES = {
    "box37x11x9_steady_v1": dict(kind="box", dims=(37, 11, 9), scheme=1, variant=1, lam=3e-3),
    "box20x17x13_theta_v2": dict(kind="box", dims=(20, 17, 13), scheme=0, variant=2, lam=1e-2,
                                 dt=0.05, P=0.3),
    "reference_default_theta_v1": dict(kind="ref"),
}


def make_problem(pkg, name):
    c = CASES[name]
    if c["kind"] == "ref":
        return pkg.problem.reference_default_case()
    ex, ey, ez = c["dims"]
    return pkg.problem.unit_box(ex, ey, ez, vel_seed=7, lam=c["lam"], scheme=c["scheme"],
                                dt=c.get("dt", 1.0), variant=c["variant"], P=c.get("P", 0.0))


@pytest.fixture(scope="module", params=list(CASES))
def case(request, pkg, po, gpu):
    name = request.param
    p = make_problem(pkg, name)
    bb, b, d = po.assemble(p)
    b, d, x0 = po.impose_mask(p.mask, b, d, np.zeros(p.N))
    if np.linalg.norm(b) == 0.0:
        b = np.random.default_rng(16).standard_normal(p.N) * (1 - p.mask)
    xo, ito, _, _ = po.dbcg_simple(b, x0, 1e-12, 20000, bb, p.iss, d, p.nax, p.nay, p.naz,
                                   p.mask)
    xs, its, _, sts = po.bicgstab_jacobi(b, x0, 1e-10, 20000, bb, p.iss, d, p.nax, p.nay,
                                         p.naz, p.mask)
    assert sts == 1
    ctx = pkg.Context.from_problem(p)
    yield name, p, b, x0, xo, its, ctx
    ctx.close()


@pytest.mark.parametrize("tile,zchunk", [(0, 0), (1, 0), (2, 0), (0, 3), (1, 5), (2, 64)])
def test_bicgstab_fp64_matches_oracle(case, tile, zchunk):
    name, p, b, x0, xo, its, ctx = case
    log(f"bicgstab fp64 {name} tile {tile} zchunk {zchunk}")
    ctx.set_options(scatter="owner", zchunk=zchunk, tile=tile)
    ctx.set_method("bicgstab")
    try:
        x, it, err, rc = ctx.solve(b, x0, rtol=1e-10, itmax=5000)
    finally:
        ctx.set_method("bicg")
        ctx.set_options(scatter="owner", zchunk=0, tile=0)
    log(f"  iters gpu {it} oracle {its}  err {err:.2e}")
    assert rc == 0 and err <= 1e-10
    assert abs(it - its) <= max(2, its // 10)
    assert ctx.true_residual(b, x) <= 1e-8
    assert rel(x, xo) <= 1e-7
    assert np.all(x[p.mask == 1] == 0.0)


@pytest.mark.parametrize("scatter", ["colored", "atomic"])
def test_bicgstab_runs_on_owner_launches_for_any_scatter(case, scatter):
    name, p, b, x0, xo, its, ctx = case
    log(f"bicgstab fp64 {name} scatter {scatter}")
    ctx.set_options(scatter=scatter)
    ctx.set_method("bicgstab")
    try:
        x, it, err, rc = ctx.solve(b, x0, rtol=1e-10, itmax=5000)
        ctx.set_options(scatter="owner")
        x2, it2, err2, _ = ctx.solve(b, x0, rtol=1e-10, itmax=5000)
    finally:
        ctx.set_method("bicg")
        ctx.set_options(scatter="owner", zchunk=0, tile=0)
    assert rc == 0 and err <= 1e-10
    # the same owner launches: bitwise the same iterates
    assert it == it2 and err == err2 and np.array_equal(x, x2)


def test_bicgstab_fp32(case, pkg):
    name, p, b, x0, xo, its, ctx = case
    log(f"bicgstab fp32 {name}")
    ctx.set_method("bicgstab")
    try:
        x, it, err, rc = ctx.solve(b, x0, rtol=1e-5, itmax=5000, precision=pkg.FP32)
    finally:
        ctx.set_method("bicg")
    log(f"  iters {it} err {err:.2e} true {ctx.true_residual(b, x):.2e}")
    assert rc == 0 and err <= 1e-5
    # FP32 vectors and operator: the true residual follows to ~1e-4
    assert ctx.true_residual(b, x) <= 1e-3


def test_bicg_unaffected_by_bicgstab(case):
    name, p, b, x0, xo, its, ctx = case
    log(f"bicg after bicgstab {name}")
    xa, ita, erra, _ = ctx.solve(b, x0, rtol=1e-10, itmax=5000)
    ctx.set_method("bicgstab")
    try:
        ctx.solve(b, x0, rtol=1e-10, itmax=5000)
    finally:
        ctx.set_method("bicg")
    xb, itb, errb, _ = ctx.solve(b, x0, rtol=1e-10, itmax=5000)
    assert ita == itb and erra == errb and np.array_equal(xa, xb)


def test_bicgstab_itmax_and_warm_start(case):
    name, p, b, x0, xo, its, ctx = case
    log(f"bicgstab itmax / warm start {name}")
    ctx.set_method("bicgstab")
    try:
        x, it, err, rc = ctx.solve(b, x0, rtol=1e-10, itmax=3, raise_on_fail=False)
        assert it == 3 and rc == 5  # VGPU_NOT_CONVERGED
        # warm start from the oracle's solution: converged at the first test
        x2, it2, err2, rc2 = ctx.solve(b, xo, rtol=1e-8, itmax=5000)
        assert rc2 == 0 and it2 == 0 and err2 <= 1e-8
    finally:
        ctx.set_method("bicg")


def test_mixed_precision_bicgstab_inner(pkg, gpu):
    """c1-sized system: FP32 BiCGSTAB inner solves and FP64 refinement reach
    the FP64 solution; the inner solves do the work (no FP64 fallback: the
    iteration total stays within 3x one FP64 BiCGSTAB solve)."""
    p = pkg.problem.config_problem("c1")
    log("mixed bicgstab c1")
    with pkg.Context.from_problem(p) as ctx:
        b = ctx.assemble_rhs(p.q_in)
        x64, it64, e64, _ = ctx.solve(b, rtol=1e-12, itmax=20000)
        ctx.set_method("bicgstab")
        xs, its, es, _ = ctx.solve(b, rtol=1e-8, itmax=20000)
        ctx.set_method("bicg")
        xm, itm, em, rc = ctx.solve(b, rtol=1e-8, itmax=20000, precision=pkg.MIXED)
        log(f"  fp64 bicg {it64}, fp64 bicgstab {its}, mixed {itm}, true rel {em:.2e}")
        assert rc == 0 and em <= 1e-8
        assert ctx.true_residual(b, xm) <= 1e-8
        assert rel(xm, x64) <= 1e-6
        assert itm <= 3 * its


@pytest.mark.parametrize("method", ["bicg", "bicgstab"])
def test_device_solves_on_rotating_buffers(pkg, gpu, method):
    """Solves of one system into different device buffers (as a time stepper
    rotates its states), timed and untimed, give bitwise the result of the
    first solve (the captured graphs do not depend on the caller's x)."""
    p = pkg.problem.unit_box(30, 22, 14, vel_seed=3, lam=2e-3)
    log(f"rotating buffers {method}")
    with pkg.Context.from_problem(p) as ctx:
        ctx.set_options(scatter="owner", zchunk=5, tile=1)
        ctx.set_method(method)
        b = ctx.assemble_rhs(p.q_in)
        bd = ctx.dev_alloc(8 * p.N)
        ctx.h2d(bd, b)
        xs = [ctx.dev_alloc(8 * p.N) for _ in range(3)]
        res = []
        for rep in range(2):
            for xd in xs:
                ctx.h2d(xd, np.zeros(p.N))
                it, err, _, rc = ctx.solve_dev(bd, xd, rtol=1e-10, itmax=5000, timed=(rep == 0))
                assert rc == 0 and err <= 1e-10
                res.append((it, err, ctx.d2h(xd, p.N)))
        ctx.set_method("bicg")
        for it, err, x in res[1:]:
            assert it == res[0][0] and err == res[0][1] and np.array_equal(x, res[0][2])
        assert ctx.true_residual(b, res[0][2]) <= 1e-8


def test_three_term_combination(pkg, gpu):
    """vgpu_axpbypcz_dev: out = a x + b y + c z on free nodes, 0 on the mask,
    also in place (out = z); to rounding (the device may fuse the products
    into multiply-adds)."""
    p = pkg.problem.unit_box(9, 8, 7)
    rng = np.random.default_rng(5)
    with pkg.Context.from_problem(p) as ctx:
        vs = [rng.standard_normal(p.N) for _ in range(3)]
        ds = [ctx.dev_alloc(8 * p.N) for _ in range(4)]
        for d_, v in zip(ds, vs):
            ctx.h2d(d_, v)
        free = p.mask == 0
        for (a, b, c), out in (((3.0, -3.0, 1.0), 3), ((0.5, 0.25, 2.0), 2)):
            ctx.axpbypcz_dev(a, ds[0], b, ds[1], c, ds[2], ds[out])
            ref = a * vs[0] + b * vs[1] + c * vs[2]
            scale = np.abs(a * vs[0]) + np.abs(b * vs[1]) + np.abs(c * vs[2])
            got = ctx.d2h(ds[out], p.N)
            assert np.all(got[~free] == 0.0)
            assert np.all(np.abs(got[free] - ref[free]) <= 4e-16 * scale[free])
