"""The oracle against the reference's own run transcript (the only golden
values the reference ships): src_v2026_2_1/vanadis_log.txt, default case of
VANADIS.FOR (7x20x20 nodes, vanadis.cfg:6, theta scheme dt = 100 s, V1 DR,
Robin on face 3, atomic CUDA BiCG to 1e-8)."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def run20(pkg, po):
    p = pkg.problem.reference_default_case()
    t = np.zeros(p.N)
    out = []
    for step in range(20):
        bb, b, d = po.assemble(p, t_old=t)
        ddf = po.ddf(p.E, p.N, p.iss, bb) if step == 0 else None
        isdd = po.is_dd(p.E, p.N, p.iss, bb) if step == 0 else None
        b, d, t = po.impose_mask(p.mask, b, d, t)
        t, it, err, rc = po.dbcg_simple(b, t, 1e-8, 1000, bb, p.iss, d, p.nax, p.nay, p.naz,
                                        p.mask)
        m = po.mass(p.iss, p.xx, p.yy, p.zz, t, p.E)
        out.append(dict(iters=it, err=err, mass=m, ddf=ddf, isdd=isdd, rc=rc))
    return p, out


def test_sizes(run20, golden_log):
    p, _ = run20
    s0 = golden_log["steps"][0]
    assert (p.E, p.N) == (s0["elements"], s0["nodes"])


def test_ddf_bitwise(run20, golden_log):
    """DDF pins every assembled element matrix (2H + 0.03 C, Robin, DR tau)."""
    _, out = run20
    assert out[0]["ddf"] == float(golden_log["steps"][0]["ddf_text"])
    assert out[0]["isdd"] is golden_log["steps"][0]["diag_dominant"]


def test_iterations_every_step(run20, golden_log):
    _, out = run20
    assert [o["iters"] for o in out] == [s["iters"] for s in golden_log["steps"]]


def test_residual_per_step(run20, golden_log):
    """The transcript prints the final recursive residual to 4 digits (e10.4);
    its run summed with FP64 atomics in a nondeterministic order, so the
    residual at the stopping iteration differs by up to 0.14 % (step 2)."""
    _, out = run20
    for o, s in zip(out, golden_log["steps"]):
        assert o["err"] == pytest.approx(s["residual"], rel=2e-3)
        assert o["err"] <= 1e-8


def test_mass_every_step(run20, golden_log):
    """Masses to 1e-10 relative (the transcript ran atomic CUDA BiCG, so the
    last bits of the iterates differ; all 20 steps agree to <3e-12 here)."""
    _, out = run20
    for o, s in zip(out, golden_log["steps"]):
        assert o["mass"] == pytest.approx(float(s["mass_text"]), rel=1e-10)


def _mass_after_step1(pkg, po, **kw):
    p = pkg.problem.reference_default_case(**kw)
    bb, b, d = po.assemble(p)
    b, d, t = po.impose_mask(p.mask, b, d, np.zeros(p.N))
    t, it, err, rc = po.dbcg_simple(b, t, 1e-8, 1000, bb, p.iss, d, p.nax, p.nay, p.naz, p.mask)
    return po.mass(p.iss, p.xx, p.yy, p.zz, t, p.E)


@pytest.mark.parametrize("kw", [dict(alpha=0.0), dict(P=0.0), dict(variant=3),
                                dict(facemap=3), dict(K=20.0)])
def test_planted_mistake_is_caught(pkg, po, golden_log, kw):
    """Each plausible mistake (a dropped Robin or decay term, the V3 tau, the
    V3 face table, a wrong diffusivity) moves the step-1 mass by far more than
    the 1e-10 tolerance; the V3 tau also moves the DDF."""
    m = _mass_after_step1(pkg, po, **kw)
    assert abs(m / float(golden_log["steps"][0]["mass_text"]) - 1) > 1e-6
    if "variant" in kw:
        p = pkg.problem.reference_default_case(**kw)
        bb, _, _ = po.assemble(p)
        assert po.ddf(p.E, p.N, p.iss, bb) != float(golden_log["steps"][0]["ddf_text"])
