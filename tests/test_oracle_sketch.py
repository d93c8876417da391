"""The committed sketches of the oracle's FP64 solutions of the large
workloads (tests/golden/oracle_sketch.json, made by
tests/golden/make_oracle_sketch.py) and the rel-L2 estimator the GPU tests and
the bench read them with.  CPU only."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("k,tol", [(32, 0.27), (16, 0.41)])
def test_estimator_on_known_differences(po, k, tol):
    """sqrt(mean (g . v)^2) estimates |v| for Gaussian g: on random, smooth
    and single-spike differences of known norm, at scales from 1e-4 down to
    1e-12 of the reference, the estimate is within the stated band."""
    rng = np.random.default_rng(5)
    n = 40000
    xo = np.cumsum(rng.standard_normal(n)) / 50.0 + 1.0
    ent = {"sketch": list(po.sketch_projections(xo, k, 910000)), "sketch_seed0": 910000,
           "x_norm2": float(np.linalg.norm(xo))}
    shapes = [rng.standard_normal(n), np.sin(np.linspace(0, 7, n)), np.eye(1, n, 1234)[0]]
    for v in shapes:
        v = v / np.linalg.norm(v)
        for scale in (1e-4, 1e-8, 1e-12):
            x = xo + scale * np.linalg.norm(xo) * v
            true = np.linalg.norm(x - xo) / np.linalg.norm(xo)
            est = po.sketch_rel_l2(x, ent)
            assert abs(est / true - 1.0) <= tol, (scale, est, true)
    assert po.sketch_rel_l2(xo, ent) == 0.0


def test_sketch_fixture_entries_are_consistent(po):
    """Every entry: the oracle reached its 1e-12 reference (its BiCG, the
    reference's method, or its BiCGSTAB where BiCG stalled); where both
    converged they agree to 1e-9 (BiCG's true residual can stop at a few 1e-9
    on the high-Peclet batch members while its recursive one reaches 1e-12,
    and its solution is still within 6e-10 of BiCGSTAB's at 1e-13), so every
    reference is good to well below the 1e-8 the GPU solutions are held to;
    the 128x128x64 entry's norm and mass equal those of the oracle's 1e-8
    solve in oracle_c2.json to 1e-8."""
    sk = po.load_sketches()
    for name in ("c2", "c3", "c4a", "c4b"):
        assert name in sk, name
    entries = [sk["c2"], sk["c3"]] + sk["c4a"]["members"] + sk["c4b"]["members"]
    assert len(sk["c4a"]["members"]) == 32 and len(sk["c4b"]["members"]) == 16
    for e in entries:
        assert e["true_rel_residual"] <= 1e-8, e["true_rel_residual"]
        if e["bicg_converged"]:
            assert e["bicg_vs_bicgstab_rel_l2"] <= 1e-9
        else:
            assert e["reference_solver"] == "bicgstab" and e["bicgstab_status"] == 1
        assert len(e["sketch"]) in (16, 32) and e["x_norm2"] > 0
    g = json.load(open(os.path.join(HERE, "golden", "oracle_c2.json")))
    assert sk["c2"]["x_norm2"] == pytest.approx(g["x_norm2"], rel=1e-8)
    assert sk["c2"]["mass"] == pytest.approx(g["mass"], rel=1e-8)


def test_sketch_matches_a_live_oracle_solve(pkg, po):
    """The 32x32x16 entry against the oracle run here: its own BiCG to 1e-12
    has a sketch estimate of ~0 (the same solution), and its 1e-8 solve's
    estimated rel-L2 matches the exact one within the estimator's band."""
    sk = po.load_sketches()
    if "c1" not in sk:
        pytest.skip("no 32x32x16 entry in the fixture")
    p = pkg.problem.config_problem("c1")
    bb, b, d = po.assemble(p)
    b, d, x0 = po.impose_mask(p.mask, b, d, np.zeros(p.N))
    args = (bb, p.iss, d, p.nax, p.nay, p.naz, p.mask)
    x12, _, _, _ = po.dbcg_simple_omp(b, x0, 1e-12, 20000, *args, 8)  # as the fixture
    x8, _, _, _ = po.dbcg_simple(b, x0, 1e-8, 20000, *args)
    assert po.sketch_rel_l2(x12, sk["c1"]) <= 1e-11
    true = np.linalg.norm(x8 - x12) / np.linalg.norm(x12)
    assert abs(po.sketch_rel_l2(x8, sk["c1"]) / true - 1.0) <= 0.27
