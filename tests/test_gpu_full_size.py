"""Parity at the benchmark size (128x128x64 elements, 1,081,665 nodes): the
operator against the oracle's matvec_0/mttvec_0 on oracle-assembled element
matrices, and the solutions against the oracle's own FP64 solves: its BiCG to
1e-8 (summary in tests/golden/oracle_c2.json, made by
tests/golden/make_oracle_c2.py) and its BiCG to 1e-12, through 32 Gaussian
projections (tests/golden/oracle_sketch.json, made by
tests/golden/make_oracle_sketch.py): every solve to rtol 1e-8 lands within
1e-8 rel-L2 of it."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def c2(pkg, po, gpu):
    p = pkg.problem.config_problem("c2")
    bb, b, d = po.assemble(p)
    ctx = pkg.Context.from_problem(p)
    yield p, bb, ctx
    ctx.close()


@pytest.mark.parametrize("scatter,tile", [("owner", 0), ("owner", 1), ("owner", 2),
                                          ("atomic", 0), ("colored", 0)])
def test_apply_full_size(c2, po, scatter, tile):
    p, bb, ctx = c2
    print(f"full-size apply {scatter} tile {tile}", flush=True)
    ctx.set_options(scatter, 0, tile)
    x = np.random.default_rng(21).standard_normal(p.N)
    for tr in (False, True):
        ref = po.matvec_0(x, p.iss, bb, p.nax, p.nay, p.naz, p.mask, tr)
        y = ctx.apply(x, scatter=scatter, transpose=tr)
        r = np.linalg.norm(y - ref) / np.linalg.norm(ref)
        assert r <= 1e-13, (tr, r)
    ctx.set_options("owner", 0, 0)


@pytest.mark.parametrize("scatter", ["owner", "atomic", "colored"])
def test_solve_full_size_matches_oracle(c2, pkg, po, scatter):
    p, bb, ctx = c2
    g = json.load(open(os.path.join(HERE, "golden", "oracle_c2.json")))
    print(f"full-size solve {scatter}", flush=True)
    ctx.set_options(scatter, 0, 0)
    b = ctx.assemble_rhs(p.q_in)
    assert np.linalg.norm(b) == pytest.approx(g["b_norm2"], rel=1e-12)
    x, it, err, rc = ctx.solve(b, rtol=1e-8, itmax=5000)
    ctx.set_options("owner", 0, 0)
    print(f"  iterations gpu {it} oracle {g['iterations']}", flush=True)
    assert err <= 1e-8
    # same algorithm, different summation order: iteration counts within 20 %
    # of the oracle's 242 for the fixed orders (owner 227-240, colored 251);
    # the atomic order changes from run to run and so does the count (over 20
    # solves: 216-320), held within 50 %
    frac = 0.5 if scatter == "atomic" else 0.2
    assert abs(it - g["iterations"]) <= frac * g["iterations"]
    # Both stop on the recursive residual at 1e-8.  The true residual of a
    # BiCG iterate differs from the recursive one by up to ~eps * iterations *
    # max_k |r_k| / |b|; this solve peaks at |r_k| ~ 1e6 |b|, so the
    # attainable true residual is a few 1e-8 and depends on summation order.
    assert ctx.true_residual(b, x) <= 5e-8
    assert np.linalg.norm(x) == pytest.approx(g["x_norm2"], rel=1e-8)
    assert ctx.mass(x) == pytest.approx(g["mass"], rel=1e-8)
    assert x.sum() == pytest.approx(g["x_sum"], rel=1e-8)
    est = po.sketch_rel_l2(x, po.load_sketches()["c2"])
    print(f"  rel-L2 vs the oracle's 1e-12 solve (sketch estimate) {est:.2e}", flush=True)
    assert est <= 1e-8


def test_mixed_full_size(c2, pkg, po):
    p, bb, ctx = c2
    g = json.load(open(os.path.join(HERE, "golden", "oracle_c2.json")))
    b = ctx.assemble_rhs(p.q_in)
    x, it, err, rc = ctx.solve(b, rtol=1e-8, itmax=20000, precision=pkg.MIXED)
    est = po.sketch_rel_l2(x, po.load_sketches()["c2"])
    print(f"  mixed inner iterations {it}, true rel residual {err:.2e}, rel-L2 vs oracle "
          f"(sketch) {est:.2e}", flush=True)
    assert err <= 1e-8 and ctx.true_residual(b, x) <= 1e-8
    assert np.linalg.norm(x) == pytest.approx(g["x_norm2"], rel=1e-8)
    assert est <= 1e-8


def test_bicgstab_full_size(c2, pkg, po):
    """FP64 BiCGSTAB (same preconditioner and stopping rule) reaches the
    oracle's BiCG solution at the benchmark size."""
    p, bb, ctx = c2
    g = json.load(open(os.path.join(HERE, "golden", "oracle_c2.json")))
    b = ctx.assemble_rhs(p.q_in)
    ctx.set_options("owner", 0, 0)
    ctx.set_method("bicgstab")
    try:
        x, it, err, rc = ctx.solve(b, rtol=1e-8, itmax=5000)
    finally:
        ctx.set_method("bicg")
    print(f"  bicgstab iterations {it}, recursive {err:.2e}", flush=True)
    assert rc == 0 and err <= 1e-8
    assert ctx.true_residual(b, x) <= 1e-8
    assert np.linalg.norm(x) == pytest.approx(g["x_norm2"], rel=1e-8)
    assert ctx.mass(x) == pytest.approx(g["mass"], rel=1e-8)
    assert x.sum() == pytest.approx(g["x_sum"], rel=1e-8)
    assert po.sketch_rel_l2(x, po.load_sketches()["c2"]) <= 1e-8


def test_colored_then_atomic_solves_one_context(c2, pkg):
    """A COLORED FP64 BiCG solve followed by an ATOMIC one on the same
    context (the drop and re-capture of the solver graph on a scatter change),
    then the owner scatter again: each reaches rtol."""
    p, bb, ctx = c2
    b = ctx.assemble_rhs(p.q_in)
    res = {}
    try:
        for sc in ("colored", "atomic", "owner"):
            print(f"full-size solve sequence: {sc}", flush=True)
            ctx.set_options(sc, 0, 0)
            x, it, err, rc = ctx.solve(b, rtol=1e-8, itmax=5000)
            res[sc] = x
            assert rc == 0 and err <= 1e-8, (sc, rc, err)
    finally:
        ctx.set_options("owner", 0, 0)
    for sc in ("colored", "atomic"):
        assert np.linalg.norm(res[sc] - res["owner"]) <= 1e-8 * np.linalg.norm(res["owner"])
