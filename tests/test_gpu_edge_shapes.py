"""Degenerate and ragged meshes on the GPU against the CPU oracle: one
element in one, two or three directions, long thin rows along x or y far
wider than a tile, a single element layer, anisotropic element sizes, and
every scheme.  On such meshes most nodes are Dirichlet nodes and a tile or
z chunk is mostly outside the grid, which the interior cases do not reach.

Bars (as for the interior cases): FP64 operator <= 1e-13 relative to the
oracle's matvec_0 / mttvec_0 for every scatter, owner tile shape and z chunk
(the owner results bitwise equal across them), zero at Dirichlet nodes;
solves to rtol 1e-10 within 1e-8 rel-L2 of the oracle's FP64 solve to 1e-12
for BiCG (every scatter) and within 1e-9 for BiCGSTAB and mixed precision;
FP32 operator <= 2e-6 and FP32 solves to rtol 1e-5 within 1e-4 (plain BiCG
not on the thin rows, where FP32 rounding at its residual peaks stalls it);
batched solves of all of them bitwise equal to the single solves."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = {
    "1x1x1": dict(dims=(1, 1, 1)),
    "1x1x7_theta": dict(dims=(1, 1, 7), scheme=0, dt=0.05, P=0.3),
    "7x1x1": dict(dims=(7, 1, 1)),
    "1x9x1_v3": dict(dims=(1, 9, 1), variant=3),
    "2x2x2_euler": dict(dims=(2, 2, 2), scheme=2, dt=0.01),
    "3x1x5_aniso": dict(dims=(3, 1, 5), extent=(3.0, 0.5, 0.1)),
    "70x1x3": dict(dims=(70, 1, 3)),
    "1x70x3": dict(dims=(1, 70, 3)),
    "33x17x1_lamfield": dict(dims=(33, 17, 1), lam_range=(1e-4, 1e-2)),
}


def log(msg):
    print(msg, flush=True)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def make(pkg, name):
    c = dict(SHAPES[name])
    dims = c.pop("dims")
    lam = {} if "lam_range" in c else dict(lam=3e-3)
    return pkg.problem.unit_box(*dims, vel_seed=5, **lam, **c)


@pytest.fixture(scope="module", params=list(SHAPES))
def shape(request, pkg, po, gpu):
    name = request.param
    p = make(pkg, name)
    bb, b, d = po.assemble(p)
    ctx = pkg.Context.from_problem(p)
    yield name, p, bb, b, d, ctx
    ctx.close()


def rhs(po, p, b, d):
    b, d, x0 = po.impose_mask(p.mask, b, d, np.zeros(p.N))
    if np.linalg.norm(b) == 0.0:
        b = np.random.default_rng(3).standard_normal(p.N) * (1 - p.mask)
    return b, d, x0


def true_res(po, p, bb, b, x):
    r = b - po.matvec_0(x, p.iss, bb, p.nax, p.nay, p.naz, p.mask, False)
    return np.linalg.norm(r) / np.linalg.norm(b)


def oracle_solution(po, p, bb, b, d, x0):
    """The oracle's FP64 BiCG to 1e-12; where its recursive residual drifts
    from the true one (the 70x1x3 row: true residual 4.5e-10), its BiCGSTAB
    to 1e-13 when that is closer to b."""
    x, it, _, rc = po.dbcg_simple(b, x0, 1e-12, 5000, bb, p.iss, d, p.nax, p.nay, p.naz,
                                  p.mask)
    if rc != 0 or true_res(po, p, bb, b, x) > 1e-11:
        x2, it2, _, _ = po.bicgstab_jacobi(b, x0, 1e-13, 5000, bb, p.iss, d, p.nax, p.nay,
                                           p.naz, p.mask)
        if rc != 0 or true_res(po, p, bb, b, x2) < true_res(po, p, bb, b, x):
            x, it = x2, it2
    return x, it


@pytest.mark.parametrize("scatter", ["owner", "colored", "atomic"])
def test_apply_matches_oracle(shape, po, scatter):
    name, p, bb, _, _, ctx = shape
    log(f"apply {name} {scatter}")
    x = np.random.default_rng(21).standard_normal(p.N)
    for tr in (False, True):
        ref = po.matvec_0(x, p.iss, bb, p.nax, p.nay, p.naz, p.mask, tr)
        y = ctx.apply(x, scatter=scatter, transpose=tr)
        assert rel(y, ref) <= 1e-13, (tr, rel(y, ref))
        assert np.all(y[p.mask == 1] == 0.0)


@pytest.mark.parametrize("scatter", ["owner", "colored", "atomic"])
def test_apply_fp32_matches_oracle(shape, po, pkg, scatter):
    name, p, bb, _, _, ctx = shape
    x = np.random.default_rng(28).standard_normal(p.N)
    ref = po.matvec_0(x, p.iss, bb, p.nax, p.nay, p.naz, p.mask)
    y = ctx.apply(x, scatter=scatter, precision=pkg.FP32)
    log(f"apply fp32 {name} {scatter}: {rel(y, ref):.2e}")
    assert rel(y, ref) <= 2e-6
    assert np.all(y[p.mask == 1] == 0.0)


@pytest.mark.parametrize("method", ["bicg", "bicgstab"])
def test_fp32_solves(shape, po, pkg, method):
    """FP32 solves (FP32 vectors and operator, FP64 scalars) to rtol 1e-5:
    within 1e-4 rel-L2 of the oracle's FP64 solution."""
    name, p, bb, b, d, ctx = shape
    if name in ("70x1x3", "1x70x3") and method == "bicg":
        # plain BiCG's residual peaks far above |b| on the thin rows (FP64:
        # 150-270 iterations for 210 unknowns), and FP32 rounding at those
        # peaks leaves it above 1e-5 at itmax (measured: rel-L2 5e-2 and 3e2)
        pytest.skip("plain BiCG in FP32 does not reach 1e-5 on the thin rows; the FP32 "
                    "BiCGSTAB solve covers them")
    b, d, x0 = rhs(po, p, b, d)
    xo, _ = oracle_solution(po, p, bb, b, d, x0)
    ctx.set_method(method)
    try:
        x, it, err, rc = ctx.solve(b, x0, rtol=1e-5, itmax=5000, precision=pkg.FP32,
                                   raise_on_fail=False)
    finally:
        ctx.set_method("bicg")
    log(f"fp32 {method} {name}: rc {rc}, iterations {it}, rel-L2 vs oracle {rel(x, xo):.2e}")
    assert rc == 0 and err <= 1e-5
    assert rel(x, xo) <= 1e-4


def test_owner_tiles_and_chunks(shape, po):
    """Every owner tile shape with z chunks of 1 and 2 node planes, more
    planes than the mesh has, and the balanced launch: the same operator,
    bitwise."""
    name, p, bb, _, _, ctx = shape
    x = np.random.default_rng(22).standard_normal(p.N)
    ref = [po.matvec_0(x, p.iss, bb, p.nax, p.nay, p.naz, p.mask, tr) for tr in (False, True)]
    base = None
    try:
        for tile in (0, 1, 2):
            for zc in (0, 1, 2, 64):
                log(f"owner {name} tile {tile} zchunk {zc}")
                ctx.set_options("owner", zc, tile)
                y = [ctx.apply(x, scatter="owner", transpose=tr) for tr in (False, True)]
                for tr in (0, 1):
                    assert rel(y[tr], ref[tr]) <= 1e-13, (tile, zc, tr, rel(y[tr], ref[tr]))
                if base is None:
                    base = y
                assert all(np.array_equal(a, c) for a, c in zip(y, base)), (tile, zc)
    finally:
        ctx.set_options("owner", 0, 0)


@pytest.mark.parametrize("scatter", ["owner", "colored", "atomic"])
def test_bicg_matches_oracle(shape, po, scatter):
    name, p, bb, b, d, ctx = shape
    b, d, x0 = rhs(po, p, b, d)
    xo, ito = oracle_solution(po, p, bb, b, d, x0)
    ctx.set_options(scatter=scatter)
    try:
        x, it, err, rc = ctx.solve(b, x0, rtol=1e-10, itmax=5000)
    finally:
        ctx.set_options(scatter="owner", zchunk=0, tile=0)
    log(f"bicg {name} {scatter}: iterations {it} (oracle to 1e-12: {ito}), err {err:.2e}, "
        f"rel-L2 vs oracle {rel(x, xo):.2e}")
    assert rc == 0 and err <= 1e-10
    # 1e-8 (the solution bar against the FP64 CPU solve): on the thin 70x1x3
    # row BiCG's recursive residual drifts from the true one (the oracle's
    # own BiCG: true residual 4.5e-10 when the recursive one is 1e-12), and
    # each scatter's summation order lands elsewhere (over 40 atomic solves of
    # it: rel-L2 up to 2.6e-9, true residual up to 8.5e-9, 197-266
    # iterations; owner and colored 1.2e-10 / 6e-10); BiCG stops on its
    # recursive residual, so its true residual is held to 5e-8, the drift
    # bound stated for plain BiCG
    assert rel(x, xo) <= 1e-8
    assert ctx.true_residual(b, x) <= 5e-8
    assert np.all(x[p.mask == 1] == 0.0)


def test_bicgstab_and_mixed_match_oracle(shape, po, pkg):
    name, p, bb, b, d, ctx = shape
    b, d, x0 = rhs(po, p, b, d)
    xo, _ = oracle_solution(po, p, bb, b, d, x0)
    ctx.set_method("bicgstab")
    try:
        x, it, err, rc = ctx.solve(b, x0, rtol=1e-10, itmax=5000)
        log(f"bicgstab {name}: iterations {it}, rel-L2 vs oracle {rel(x, xo):.2e}")
        assert rc == 0 and rel(x, xo) <= 1e-9
    finally:
        ctx.set_method("bicg")
    x, it, err, rc = ctx.solve(b, x0, rtol=1e-10, itmax=5000, precision=pkg.MIXED)
    log(f"mixed {name}: inner iterations {it}, rel-L2 vs oracle {rel(x, xo):.2e}")
    assert rc == 0 and rel(x, xo) <= 1e-9
    assert ctx.true_residual(b, x) <= 1e-9


def test_assembly_and_diagnostics(shape, po):
    """Diagonal, right-hand side with an old state, mass integral, diagonal
    dominance factor and flag, and damped Jacobi sweeps on the device
    against the oracle."""
    name, p, bb, b, d, ctx = shape
    log(f"assembly and diagnostics {name}")
    _, dref, _ = po.impose_mask(p.mask, b, d, np.zeros(p.N))
    assert rel(ctx.diag(), dref) <= 1e-14
    t_old = np.random.default_rng(25).standard_normal(p.N) * (1 - p.mask)
    _, bref, dref2 = po.assemble(p, t_old=t_old)
    bref, _, _ = po.impose_mask(p.mask, bref, dref2, t_old)
    assert rel(ctx.assemble_rhs(p.q_in, t_old), bref) <= 1e-13
    c = np.random.default_rng(26).random(p.N)
    assert ctx.mass(c) == pytest.approx(po.mass(p.iss, p.xx, p.yy, p.zz, c, p.E), rel=1e-13)
    dd, dom = ctx.ddf()
    assert dd == pytest.approx(po.ddf(p.E, p.N, p.iss, bb), rel=1e-12)
    assert dom == po.is_dd(p.E, p.N, p.iss, bb)
    bm, dm, _ = rhs(po, p, b, d)
    x0 = np.random.default_rng(27).standard_normal(p.N) * (1 - p.mask)
    xr = x0.copy()
    for _ in range(2):
        ax = po.matvec_0(xr, p.iss, bb, p.nax, p.nay, p.naz, p.mask)
        xr = np.where(p.mask == 1, 0.0, xr + 0.8 * (bm - ax) / dm)
    bd, xd = ctx.dev_alloc(8 * p.N), ctx.dev_alloc(8 * p.N)
    try:
        ctx.h2d(bd, bm)
        ctx.h2d(xd, x0)
        ctx.jacobi_dev(bd, xd, 2, 0.8)
        assert rel(ctx.d2h(xd, p.N), xr) <= 1e-13
    finally:
        ctx.dev_free(bd)
        ctx.dev_free(xd)


@pytest.mark.parametrize("scatter", ["owner", "colored", "atomic"])
def test_autotuners(shape, po, pkg, tmp_path, scatter):
    """The owner tuner (FP64 and FP32) and the brick tuners on these meshes,
    where most tile shapes and z chunks cover little but the outside of the
    grid: a valid choice, and the operator with it still the oracle's."""
    name, p, bb, _, _, ctx = shape
    x = np.random.default_rng(24).standard_normal(p.N)
    cache = str(tmp_path / "tune.json")
    try:
        for prec in ((pkg.FP64, pkg.FP32, pkg.MIXED) if scatter == "owner" else (pkg.FP64,)):
            kw = {} if scatter == "owner" else dict(scatter=scatter)
            r = pkg.autotune.tune(ctx, prec, cache_path=cache, reps=2, **kw)
            log(f"autotune {name} {scatter} precision {prec}: tile {r['tile']} zchunk "
                f"{r['zchunk']} {r['ms'] * 1e3:.1f} us")
            assert r["tile"] in (0, 1, 2) and r["zchunk"] >= 0 and r["ms"] > 0
            for tr in (False, True):
                ref = po.matvec_0(x, p.iss, bb, p.nax, p.nay, p.naz, p.mask, tr)
                y = ctx.apply(x, scatter=scatter, transpose=tr)
                assert rel(y, ref) <= 1e-13, (tr, rel(y, ref))
    finally:
        ctx.set_options("owner", 0, 0)


@pytest.mark.parametrize("scatter", ["colored", "atomic"])
def test_legacy_drop_in(shape, po, pkg, scatter):
    """kernel_dbcg_simple_ on the oracle's element matrices (the host's
    role), where most of the eight colours hold no element: the solution
    against the oracle's, and (colored) the device's colour lists against
    the oracle's colouring bit for bit."""
    name, p, bb, b, d, _ = shape
    b, d, t = rhs(po, p, b, d)
    xo, _ = oracle_solution(po, p, bb, b, d, t)
    x, it, err = pkg.kernel_dbcg_simple(p.N, b, t, 1e-10, 5000, bb, p.iss, d, p.nax, p.nay,
                                        p.naz, p.mask, scatter=scatter)
    log(f"legacy {name} {scatter}: iterations {it}, err {err:.2e}, rel-L2 vs oracle "
        f"{rel(x, xo):.2e}")
    assert err <= 1e-10
    assert rel(x, xo) <= 1e-8
    if scatter == "colored":
        el, off, cnt = pkg._lib.legacy_device_color_lists(p.E)
        _, oel, ooff, ocnt = po.colors(p.nax, p.nay, p.naz)
        assert np.array_equal(el, oel) and np.array_equal(off, ooff)
        assert np.array_equal(cnt, ocnt)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_of_edge_shapes_bitwise(pkg, gpu, precision):
    """All the shapes above in one batch (z-chunk launches, 16x16 tiles):
    each member bitwise equal to its single solve on its own context."""
    prec = pkg.FP64 if precision == "fp64" else pkg.MIXED
    probs = [make(pkg, n) for n in SHAPES]
    ctxs = [pkg.Context.from_problem(q) for q in probs]
    try:
        for c in ctxs:
            c.set_options("owner", 3, 1)
        bs = [c.assemble_rhs(q.q_in) for c, q in zip(ctxs, probs)]
        singles = [c.solve(b, rtol=1e-9, itmax=5000, precision=prec, raise_on_fail=False)
                   for c, b in zip(ctxs, bs)]
        log(f"batch of {len(probs)} edge shapes, {precision}")
        with pkg.Batch(ctxs) as bt:
            xs, its, errs, st, rc = bt.solve(bs, rtol=1e-9, itmax=5000, precision=prec,
                                             raise_on_fail=False)
        log(f"  iterations {list(its)}, status {list(st)}")
        assert rc == pkg.OK and (st == 1).all(), st
        for m, (x1, it1, e1, rc1) in enumerate(singles):
            assert rc1 == pkg.OK, m
            assert its[m] == it1 and errs[m] == e1, (m, its[m], it1)
            assert np.array_equal(xs[m], x1), m
    finally:
        for c in ctxs:
            c.close()
