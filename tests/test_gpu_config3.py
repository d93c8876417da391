"""256x256x64 elements with per-element diffusivity log-uniform in
[1e-4, 1e-2] (4.29M nodes): the FP64 solve reaches rtol 1e-8 with a true
residual at the solver's attainable accuracy, the operator agrees between
scatter modes and precisions, and the FP64 BiCG, FP64 BiCGSTAB and mixed
precision (FP32 inner BiCGSTAB at rtol 1e-4, FP64 refinement) solutions are
within 1e-8 rel-L2 of the CPU oracle's FP64 BiCG run to 1e-12 (estimated from
the committed sketch of its solution, tests/golden/oracle_sketch.json)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3(pkg, gpu):
    p = pkg.problem.config_problem("c3")
    ctx = pkg.Context.from_problem(p)
    yield p, ctx
    ctx.close()


def test_c3_operator_modes_agree(c3):
    p, ctx = c3
    x = np.random.default_rng(31).standard_normal(p.N)
    ref = ctx.apply(x)
    for sc in ("atomic", "colored"):
        y = ctx.apply(x, scatter=sc)
        assert np.linalg.norm(y - ref) <= 1e-13 * np.linalg.norm(ref), sc
    yf = ctx.apply(x, precision=1)
    assert np.linalg.norm(yf - ref) <= 2e-6 * np.linalg.norm(ref)


def test_c3_fp64_and_mixed_solves(c3, pkg, po):
    p, ctx = c3
    b = ctx.assemble_rhs(p.q_in)
    x64, it64, err64, rc = ctx.solve(b, rtol=1e-8, itmax=50000, raise_on_fail=False)
    print(f"  c3 fp64: rc {rc} iterations {it64} err {err64:.2e}", flush=True)
    assert rc == pkg.OK and err64 <= 1e-8
    r64 = ctx.true_residual(b, x64)
    assert r64 <= 5e-8
    xm, itm, errm, rcm = ctx.solve(b, rtol=1e-8, itmax=50000, precision=pkg.MIXED,
                                   inner_rtol=1e-4, raise_on_fail=False)
    rm = ctx.true_residual(b, xm)
    d = np.linalg.norm(xm - x64) / np.linalg.norm(x64)
    print(f"  c3 mixed: rc {rcm} inner iterations {itm} true residual {rm:.2e} "
          f"rel-L2 vs fp64 {d:.2e}", flush=True)
    assert rcm == pkg.OK and rm <= 1e-8
    # both against the oracle's FP64 BiCG run to 1e-12 (tests/golden/
    # oracle_sketch.json: 32 Gaussian projections of its solution)
    sk = po.load_sketches()["c3"]
    e64, em = po.sketch_rel_l2(x64, sk), po.sketch_rel_l2(xm, sk)
    print(f"  c3 rel-L2 vs oracle (sketch): fp64 {e64:.2e}, mixed {em:.2e}", flush=True)
    assert e64 <= 1e-8 and em <= 1e-8
    assert d <= 2e-8
    ctx.set_method("bicgstab")
    try:
        xs, its, errs, rcs = ctx.solve(b, rtol=1e-8, itmax=50000, raise_on_fail=False)
    finally:
        ctx.set_method("bicg")
    es = po.sketch_rel_l2(xs, sk)
    print(f"  c3 bicgstab: rc {rcs} iterations {its} rel-L2 vs oracle (sketch) {es:.2e}",
          flush=True)
    assert rcs == pkg.OK and es <= 1e-8


def test_c3_mixed_loose_inner_tolerance(c3, pkg, po):
    """The mixed solve with inner rtol 1e-2 at the FP32 iteration launch's
    autotuned shape (the setting the benchmark reports beside the specified
    1e-4): true residual <= 1e-8 and within 1e-8 rel-L2 of the oracle's FP64
    solve (sketch)."""
    p, ctx = c3
    b = ctx.assemble_rhs(p.q_in)
    t32 = ctx.autotune(pkg.MIXED, reps=3)
    try:
        x, it, err, rc = ctx.solve(b, rtol=1e-8, itmax=50000, precision=pkg.MIXED,
                                   inner_rtol=1e-2, raise_on_fail=False)
    finally:
        ctx.set_options("owner", 0, 0)
    tr = ctx.true_residual(b, x)
    e = po.sketch_rel_l2(x, po.load_sketches()["c3"])
    print(f"  c3 mixed, inner 1e-2, tile {t32['tile']} z chunk {t32['zchunk']}: rc {rc} "
          f"inner iterations {it} true residual {tr:.2e} rel-L2 vs oracle (sketch) {e:.2e}",
          flush=True)
    assert rc == pkg.OK and err <= 1e-8 and tr <= 1e-8
    assert e <= 1e-8
