"""Pins of the oracle that do not use the reference: exact Kronecker closed
forms of box elements (2x2x2 Gauss is exact for them), invariants, scipy
sparse algebra and brute force on tiny meshes."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

# local node -> (ax, ay, az), VANADIS.FOR:255-262 / par_3d 332-357
LOC = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def kron3(Fx, Fy, Fz):
    A = np.zeros((8, 8))
    for a, (ax, ay, az) in enumerate(LOC):
        for b, (bx, by, bz) in enumerate(LOC):
            A[a, b] = Fx[ax, bx] * Fy[ay, by] * Fz[az, bz]
    return A


def box_closed_form(h, lam, v, tau, P, alpha=0.0, robin=False, q=0.0):
    """Exact H, C (= M_w) and F of a box element from 1-D integrals."""
    M = [hk / 6.0 * np.array([[2.0, 1.0], [1.0, 2.0]]) for hk in h]
    K = [1.0 / hk * np.array([[1.0, -1.0], [-1.0, 1.0]]) for hk in h]
    Cn = 0.5 * np.array([[-1.0, 1.0], [-1.0, 1.0]])  # int phi_a phi_b'
    Ct = Cn.T                                        # int phi_a' phi_b

    def pat(sub):  # sub: dict axis -> matrix, M elsewhere
        F = [sub.get(k, M[k]) for k in range(3)]
        return kron3(*F)

    D = lam * (pat({0: K[0]}) + pat({1: K[1]}) + pat({2: K[2]}))
    G = sum(v[k] * pat({k: Cn}) for k in range(3))
    S = np.zeros((8, 8))
    for k in range(3):
        for l in range(3):
            T = pat({k: K[k]}) if k == l else pat({k: Ct, l: Cn})
            S += v[k] * v[l] * T
    Mw = pat({}) + tau * sum(v[k] * pat({k: Ct}) for k in range(3))
    H = D + G + tau * S + P * Mw
    if robin:
        E0 = np.array([[1.0, 0.0], [0.0, 0.0]])
        E0p = np.array([[-1.0 / h[0], 0.0], [1.0 / h[0], 0.0]])
        R = kron3(E0, M[1], M[2]) + tau * (v[0] * kron3(E0p, M[1], M[2]) +
                                           v[1] * kron3(E0, Ct, M[2]) +
                                           v[2] * kron3(E0, M[1], Ct))
        H = H + alpha * R
    s = np.array([[2 * c - 1 for c in loc] for loc in LOC], dtype=float)
    F = q * (0.125 + tau * (s @ (np.asarray(v) / (4.0 * np.asarray(h)))))
    return H, Mw, F


def box_nodes(x0, h):
    xe = np.array([x0[0] + h[0] * c[0] for c in LOC])
    ye = np.array([x0[1] + h[1] * c[1] for c in LOC])
    ze = np.array([x0[2] + h[2] * c[2] for c in LOC])
    return xe, ye, ze


def tau_of(po, variant, xe, ye, ze, lam, v):
    import ctypes as C
    n = np.zeros(8); w = np.zeros(8); dx = np.zeros(8); dy = np.zeros(8); dz = np.zeros(8)
    dj = C.c_double(); t = C.c_double()
    po.lib().orc_gauss_point(variant, 0.1, -0.2, 0.3, xe, ye, ze, np.full(3, lam),
                             np.asarray(v, float), C.byref(dj), n, w, dx, dy, dz, C.byref(t))
    return t.value


@pytest.mark.parametrize("variant", [1, 2, 3])
@pytest.mark.parametrize("seed", range(6))
def test_el3d_equals_kronecker_closed_form(po, variant, seed):
    rng = np.random.default_rng(100 * variant + seed)
    h = rng.uniform(0.05, 3.0, 3)
    x0 = rng.uniform(-50, 50, 3)
    v = rng.uniform(-3, 3, 3)
    # V2/V3 evaluate coth(Pe) with exp(Pe), which overflows to NaN for
    # Pe > ~709 (tested separately); keep Pe moderate here
    lam = 10 ** rng.uniform(-3 if variant == 1 else -1, 0)
    P = rng.uniform(0, 0.1)
    alpha = rng.uniform(0, 0.5)
    robin = seed % 2 == 0
    q = rng.uniform(0, 2)
    xe, ye, ze = box_nodes(x0, h)
    H, Cm, F = po.el3d(variant, xe, ye, ze, np.full(3, lam), P, v, q, int(robin), alpha, 0.0)
    tau = tau_of(po, variant, xe, ye, ze, lam, v)
    Hc, Cc, Fc = box_closed_form(h, lam, v, tau, P, alpha, robin, q)
    scale = np.abs(Hc).max()
    assert np.abs(H - Hc).max() <= 1e-13 * scale
    assert np.abs(Cm - Cc).max() <= 1e-13 * np.abs(Cc).max()
    assert np.abs(F - Fc).max() <= 1e-13 * max(np.abs(Fc).max(), 1e-300)


def test_el3d_invariants(po):
    """Diffusion and advection annihilate constants; int N_a N_b sums to the
    volume; with v = 0 the element matrix is symmetric."""
    h = np.array([0.3, 0.7, 1.1])
    xe, ye, ze = box_nodes([1.0, 2.0, 3.0], h)
    H, Cm, F = po.el3d(1, xe, ye, ze, np.full(3, 0.2), 0.0, np.array([0.4, -0.3, 0.2]), 1.0,
                       0, 0.0, 0.0)
    assert np.abs(H @ np.ones(8)).max() < 1e-14 * np.abs(H).max()
    vol = h.prod()
    assert Cm.sum() == pytest.approx(vol, rel=1e-14)
    assert F.sum() == pytest.approx(1.0, rel=1e-14)  # q_in = 1 distributed
    H0, C0, _ = po.el3d(1, xe, ye, ze, np.full(3, 0.2), 0.05, np.zeros(3), 0.0, 0, 0.0, 0.0)
    assert np.allclose(H0, H0.T, rtol=0, atol=1e-15 * np.abs(H0).max())
    assert np.allclose(C0, C0.T, rtol=0, atol=1e-16)


def test_tau_limits_and_variants(po):
    xe, ye, ze = box_nodes([0, 0, 0], [0.1, 0.1, 0.1])
    # tau -> 0 as v -> 0
    assert tau_of(po, 1, xe, ye, ze, 1.0, [0, 0, 0]) == 0.0
    # large Pe: tau -> h/(2|v|) (beta -> 1)
    t = tau_of(po, 1, xe, ye, ze, 1e-9, [1.0, 0, 0])
    assert t == pytest.approx(0.05, rel=1e-6)
    # small Pe: beta ~ Pe/3 -> tau ~ h^2/(12 lam)
    t = tau_of(po, 1, xe, ye, ze, 1e3, [1e-4, 0, 0])
    assert t == pytest.approx(0.1 ** 2 / (12 * 1e3), rel=1e-6)
    # V1 and V2 agree for isotropic lambda away from the branch points
    for pe_v in (0.3, 2.0, 7.0):
        t1 = tau_of(po, 1, xe, ye, ze, 0.01, [pe_v, 0.2, -0.1])
        t2 = tau_of(po, 2, xe, ye, ze, 0.01, [pe_v, 0.2, -0.1])
        assert t1 == pytest.approx(t2, rel=1e-12)
    # V3 differs when v is oblique (axis-sum streamline length)
    t1 = tau_of(po, 1, xe, ye, ze, 0.01, [1.0, 0.3, 0.0])
    t3 = tau_of(po, 3, xe, ye, ze, 0.01, [1.0, 0.3, 0.0])
    assert abs(t1 - t3) > 1e-3 * t1


def test_tau_overflow_behaviour(po):
    """V1 switches to 1 - 1/Pe above Pe = 50 (VANADIS.FOR:2000-2001); V2/V3
    form cosh/sinh from exp(Pe) and give NaN once exp overflows
    (omp_timing.for:1682-1686).  The oracle keeps both behaviours."""
    xe, ye, ze = box_nodes([0, 0, 0], [1.0, 1.0, 1.0])
    t1 = tau_of(po, 1, xe, ye, ze, 1e-4, [1.0, 0, 0])  # Pe = 5000
    assert t1 == pytest.approx(0.5 * (1 - 1 / 5000.0), rel=1e-14)
    t3 = tau_of(po, 3, xe, ye, ze, 1e-4, [1.0, 0, 0])
    assert np.isnan(t3)


def csr_from_bb(p, bb):
    E = p.E
    iss = p.iss.reshape(8, E) - 1
    rows = np.repeat(iss.T, 8, axis=1).ravel()
    cols = np.tile(iss.T, (1, 8)).ravel()
    return sp.csr_matrix((bb, (rows, cols)), shape=(p.N, p.N))


def test_matvec0_equals_masked_csr(pkg, po):
    p = pkg.problem.unit_box(5, 4, 3, vel_seed=2, lam=3e-2)
    bb, b, d = po.assemble(p)
    A = csr_from_bb(p, bb)
    D = sp.diags(1.0 - p.mask)
    At = D @ A @ D
    x = np.random.default_rng(1).standard_normal(p.N)
    assert np.allclose(po.matvec_0(x, p.iss, bb, p.nax, p.nay, p.naz, p.mask), At @ x,
                       rtol=0, atol=1e-14 * np.abs(At @ x).max())
    assert np.allclose(po.matvec_0(x, p.iss, bb, p.nax, p.nay, p.naz, p.mask, True),
                       At.T @ x, rtol=0, atol=1e-14 * np.abs(At.T @ x).max())
    # diag is the diagonal of the assembled matrix
    assert np.allclose(d, A.diagonal(), rtol=1e-15, atol=0)


@pytest.mark.parametrize("dims", [(5, 4, 3), (12, 9, 7)])
def test_colored_matvec0_equals_masked_csr(pkg, po, dims):
    """The 8-colour OpenMP matvec_0 (the CPU baseline's parallel operator)
    is the masked CSR product, bitwise the same for any thread count, and
    within rounding of the serial matvec_0 (only a node's addition order
    differs)."""
    p = pkg.problem.unit_box(*dims, vel_seed=4, lam=1e-2)
    bb, b, d = po.assemble(p)
    A = csr_from_bb(p, bb)
    D = sp.diags(1.0 - p.mask)
    ref = D @ A @ D @ (x := np.random.default_rng(2).standard_normal(p.N))
    y1 = po.matvec_0_colored(x, p.iss, bb, p.nax, p.nay, p.naz, p.mask, threads=1)
    y4 = po.matvec_0_colored(x, p.iss, bb, p.nax, p.nay, p.naz, p.mask, threads=4)
    ys = po.matvec_0(x, p.iss, bb, p.nax, p.nay, p.naz, p.mask)
    assert np.array_equal(y1, y4)
    assert np.allclose(y1, ref, rtol=0, atol=1e-14 * np.abs(ref).max())
    assert np.abs(y1 - ys).max() <= 1e-15 * np.abs(ys).max() * 8
    assert np.all(y1[p.mask == 1] == 0.0)


def test_bicg_converges_to_direct_solution(pkg, po):
    p = pkg.problem.unit_box(6, 5, 4, vel_seed=5, lam=2e-2)
    bb, b, d = po.assemble(p)
    b, d, x0 = po.impose_mask(p.mask, b, d, np.zeros(p.N))
    x, it, err, rc = po.dbcg_simple(b, x0, 1e-12, 2000, bb, p.iss, d, p.nax, p.nay, p.naz,
                                    p.mask)
    assert rc == 0 and err <= 1e-12
    A = csr_from_bb(p, bb).tolil()
    m = p.mask == 1
    Ad = A.tocsr()
    Dm = sp.diags(1.0 - p.mask)
    Asys = (Dm @ Ad @ Dm + sp.diags(p.mask.astype(float))).tocsc()
    xd = spla.spsolve(Asys, b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9
    assert np.all(x[m] == 0.0)


@pytest.mark.parametrize("dims,seed,lam", [((6, 5, 4), 5, 2e-2), ((9, 7, 5), 3, 3e-3)])
def test_bicgstab_converges_to_direct_solution(pkg, po, dims, seed, lam):
    """The oracle's Jacobi-BiCGSTAB reaches the direct solution of the masked
    system, and its first iterates are those of scipy's preconditioned
    BiCGSTAB (the same algorithm, Templates sec. 2.3.8) to rounding (BiCGSTAB
    amplifies rounding differences from iteration to iteration; scipy also
    stops half an iteration early, on |s|, so the counts are not compared)."""
    p = pkg.problem.unit_box(*dims, vel_seed=seed, lam=lam)
    bb, b, d = po.assemble(p)
    b, d, x0 = po.impose_mask(p.mask, b, d, np.zeros(p.N))
    x, it, err, st = po.bicgstab_jacobi(b, x0, 1e-12, 2000, bb, p.iss, d, p.nax, p.nay,
                                        p.naz, p.mask)
    assert st == 1 and err <= 1e-12
    Dm = sp.diags(1.0 - p.mask)
    Asys = (Dm @ csr_from_bb(p, bb) @ Dm + sp.diags(p.mask.astype(float))).tocsc()
    xd = spla.spsolve(Asys, b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9
    assert np.all(x[p.mask == 1] == 0.0)
    # scipy on the same system, M = D
    M = spla.LinearOperator(Asys.shape, matvec=lambda v: v / d)
    hist = []
    spla.bicgstab(Asys, b, rtol=1e-30, atol=0.0, maxiter=5, M=M,
                  callback=lambda xk: hist.append(xk.copy()))
    assert len(hist) == 5
    for k in (1, 2, 5):
        xk, itk, _, _ = po.bicgstab_jacobi(b, x0, 1e-30, k, bb, p.iss, d, p.nax, p.nay,
                                           p.naz, p.mask)
        assert itk == k
        assert np.linalg.norm(xk - hist[k - 1]) <= 1e-11 * np.linalg.norm(hist[k - 1])


def test_bicgstab_exact_step_converges(pkg, po):
    """A mesh of one element with seven Dirichlet nodes leaves a 1x1 system:
    BiCGSTAB's first half step solves it exactly (s = 0, so t = 0), which
    must end the solve converged after one iteration with x = b / A (the
    Templates' stop on small s), not report a breakdown on t . t = 0."""
    p = pkg.problem.unit_box(1, 1, 1, vel_seed=5, lam=3e-3)
    bb, b, d = po.assemble(p)
    b, d, x0 = po.impose_mask(p.mask, b, d, np.zeros(p.N))
    free = np.flatnonzero(p.mask == 0)
    assert len(free) == 1
    x, it, err, st = po.bicgstab_jacobi(b, x0, 1e-12, 50, bb, p.iss, d, p.nax, p.nay,
                                        p.naz, p.mask)
    Asys = csr_from_bb(p, bb)
    f = free[0]
    assert st == 1 and it == 1 and err == 0.0
    assert x[f] == pytest.approx(b[f] / Asys[f, f], rel=1e-14)
    assert np.all(x[p.mask == 1] == 0.0)


def test_theta_scheme_assembly(pkg, po):
    """BB = 2H + (3/dt)C and b = 3F - (H - (3/dt)C) T_old per element
    (VANADIS.FOR:723-748), checked against el3d outputs."""
    p = pkg.problem.reference_default_case()
    rng = np.random.default_rng(3)
    t_old = rng.standard_normal(p.N)
    bb, b, d = po.assemble(p, t_old=t_old)
    E = p.E
    iss = p.iss.reshape(8, E) - 1
    bref = np.zeros(p.N)
    kod = p.kod.reshape(6, E)
    for e in range(E):
        nd = iss[:, e]
        H, Cm, F = po.el3d(1, p.xx[nd], p.yy[nd], p.zz[nd], np.full(3, p.lam[e]), p.P,
                           p.vel[[e, E + e, 2 * E + e]], p.q_in[e], int(kod[2, e] == 3),
                           p.alpha, 0.0)
        A = 2 * H + 3.0 / p.dt * Cm
        assert np.array_equal(bb[64 * e:64 * e + 64].reshape(8, 8), A)
        np.add.at(bref, nd, 3 * F - (H - 3.0 / p.dt * Cm) @ t_old[nd])
    assert np.allclose(b, bref, rtol=1e-13, atol=1e-13 * np.abs(bref).max())


def test_mass_integral_exact_for_trilinear(pkg, po):
    p = pkg.problem.reference_default_case()
    vol = (420.0 - 0.0) * 8000.0 * 8000.0
    assert po.mass(p.iss, p.xx, p.yy, p.zz, np.ones(p.N), p.E) == pytest.approx(vol, rel=1e-14)
    c = 2.0 + 0.001 * p.xx - 3e-5 * p.yy
    exact = 2.0 * vol + 0.001 * (420.0 ** 2 / 2) * 8000 * 8000
    assert po.mass(p.iss, p.xx, p.yy, p.zz, c, p.E) == pytest.approx(exact, rel=1e-12)


def test_ddf_brute_force(pkg, po):
    p = pkg.problem.unit_box(3, 3, 2, vel_seed=1, lam=1e-2)
    bb, _, _ = po.assemble(p)
    A = csr_from_bb(p, bb)
    # reference definition: |entries| summed per element contribution
    iss = p.iss.reshape(8, p.E) - 1
    dg = np.zeros(p.N)
    sr = np.zeros(p.N)
    for e in range(p.E):
        B = np.abs(bb[64 * e:64 * e + 64].reshape(8, 8))
        for i, j in itertools.product(range(8), range(8)):
            (dg if i == j else sr)[iss[i, e]] += B[i, j]
    expect = max(sr[k] / dg[k] for k in range(p.N) if dg[k] > 0)
    assert po.ddf(p.E, p.N, p.iss, bb) == pytest.approx(expect, rel=1e-14)
    assert A.shape == (p.N, p.N)
