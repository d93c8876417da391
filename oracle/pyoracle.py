"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) and of the
flang-built reference (oracle/_ref/*.so, when present).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the reference the GPU results are compared with or the timed CPU baseline.  The
product path never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int)


def build(force: bool = False) -> None:
    """Compile liboracle.so (and the reference build when its sources exist)."""
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-B" if force else "-s", "-C", HERE, "liboracle.so"],
                       check=True, stdout=subprocess.DEVNULL)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_el3d.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_double, _dp, C.c_double,
                               C.c_int, C.c_double, C.c_double, _dp, _dp, _dp]
        L.orc_gauss_point.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp,
                                      _dp, _dp, _dp, _pd, _dp, _dp, _dp, _dp, _dp, _pd]
        L.orc_assemble.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                   _dp, _dp, _dp, _ip, _ip, _dp, C.c_double, _dp, _dp,
                                   C.c_double, C.c_double, C.c_void_p, _dp, _dp, _dp]
        L.orc_assemble.restype = C.c_int
        for f in (L.orc_matvec_0, L.orc_mttvec_0):
            f.argtypes = [_dp, _dp, _ip, _dp, C.c_int, C.c_int, C.c_int, _ip]
        L.orc_matvec_0_colored.argtypes = [_dp, _dp, _ip, _dp, C.c_int, C.c_int, C.c_int, _ip,
                                           C.c_int]
        L.orc_matvec_0_colored.restype = C.c_int
        L.orc_dbcg_simple.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_int, _pi, _pd, _dp,
                                      _ip, _dp, C.c_int, C.c_int, C.c_int, _ip]
        L.orc_dbcg_simple.restype = C.c_int
        L.orc_mttvec_0_colored.argtypes = [_dp, _dp, _ip, _dp, C.c_int, C.c_int, C.c_int, _ip,
                                           C.c_int]
        L.orc_mttvec_0_colored.restype = C.c_int
        L.orc_dbcg_simple_omp.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_int, _pi, _pd,
                                          _dp, _ip, _dp, C.c_int, C.c_int, C.c_int, _ip,
                                          C.c_int]
        L.orc_dbcg_simple_omp.restype = C.c_int
        L.orc_assemble_omp.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                       C.c_int, _dp, _dp, _dp, _ip, _ip, _dp, C.c_double, _dp,
                                       _dp, C.c_double, C.c_double, C.c_void_p, _dp, _dp, _dp,
                                       C.c_int]
        L.orc_assemble_omp.restype = C.c_int
        L.orc_integrate_mass_hex8.argtypes = [C.c_int, _ip, _dp, _dp, _dp, _dp]
        L.orc_integrate_mass_hex8.restype = C.c_double
        L.orc_diagonal_dominance_factor.argtypes = [C.c_int, C.c_int, _ip, _dp]
        L.orc_diagonal_dominance_factor.restype = C.c_double
        L.orc_is_diagonally_dominant.argtypes = [C.c_int, C.c_int, _ip, _dp]
        L.orc_is_diagonally_dominant.restype = C.c_int
        L.orc_w_el.argtypes = [C.c_int, C.c_int, C.c_int, _ip]
        L.orc_kod_el_dr.argtypes = [C.c_int, C.c_int, C.c_int, _ip]
        L.orc_build_bc_nodes.argtypes = [C.c_int, _ip, _ip, C.c_int, C.c_int, C.c_int, _ip]
        L.orc_siatka.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_generate_hex8_colors.argtypes = [C.c_int, C.c_int, C.c_int, _ip]
        L.orc_build_color_element_lists.argtypes = [C.c_int, _ip, _ip, _ip, _ip]
        _fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        L.orc_net_grads.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                                    _dp, _ip, _fp, C.c_double, C.c_double, _fp, _fp, C.c_int,
                                    _pd, _dp,
                                    _dp, _dp]
        L.orc_net_grads.restype = C.c_int
        _lib = L
    return _lib


# ---- thin numpy-facing helpers --------------------------------------------
def w_el(nax, nay, naz):
    E = (nax - 1) * (nay - 1) * (naz - 1)
    iss = np.zeros(8 * E, np.int32)
    lib().orc_w_el(nax, nay, naz, iss)
    return iss


def kod_el_dr(nax, nay, naz):
    E = (nax - 1) * (nay - 1) * (naz - 1)
    kod = np.zeros(6 * E, np.int32)
    lib().orc_kod_el_dr(nax, nay, naz, kod)
    return kod


def build_bc_nodes(facemap, iss, kod, nax, nay, naz):
    m = np.zeros(nax * nay * naz, np.int32)
    lib().orc_build_bc_nodes(facemap, iss, kod, nax, nay, naz, m)
    return m


def el3d(variant, xe, ye, ze, lam3, P, v3, q_in=0.0, robin=0, alpha=0.0, qs=0.0):
    H = np.zeros(64)
    Cm = np.zeros(64)
    F = np.zeros(8)
    lib().orc_el3d(variant, np.ascontiguousarray(xe, np.float64),
                   np.ascontiguousarray(ye, np.float64), np.ascontiguousarray(ze, np.float64),
                   np.ascontiguousarray(lam3, np.float64), P,
                   np.ascontiguousarray(v3, np.float64), q_in, robin, alpha, qs, H, Cm, F)
    return H.reshape(8, 8), Cm.reshape(8, 8), F


def assemble(p, t_old=None, scheme=None, dt=None, threads=0):
    """BB [64E], b [N], diag [N] of problem p (no mask imposition).
    threads > 0: element systems in parallel (orc_assemble_omp, bitwise the
    serial result)."""
    N, E = p.N, p.E
    bb = np.empty(64 * E)
    b = np.zeros(N)
    d = np.zeros(N)
    t = None if t_old is None else np.ascontiguousarray(t_old, np.float64)
    args = (p.variant, p.scheme if scheme is None else scheme,
            p.dt if dt is None else dt, p.nax, p.nay, p.naz,
            p.xx, p.yy, p.zz, p.iss, p.kod, np.ascontiguousarray(p.lam, np.float64),
            p.P, np.ascontiguousarray(p.vel, np.float64),
            np.ascontiguousarray(p.q_in, np.float64), p.alpha, p.qs,
            None if t is None else t.ctypes.data_as(C.c_void_p), bb, b, d)
    if threads > 0:
        if lib().orc_assemble_omp(*args, int(threads)) != 0:
            raise MemoryError("orc_assemble_omp")
    else:
        lib().orc_assemble(*args)
    return bb, b, d


def impose_mask(mask, b, diag, x):
    """VANADIS.FOR:799-805: b = 0, diag = 1, x = 0 on masked nodes."""
    m = mask == 1
    b = b.copy()
    diag = diag.copy()
    x = x.copy()
    b[m] = 0.0
    diag[m] = 1.0
    x[m] = 0.0
    return b, diag, x


def matvec_0(x, iss, bb, nax, nay, naz, mask, transpose=False):
    xin = np.array(x, np.float64, copy=True)
    y = np.zeros_like(xin)
    f = lib().orc_mttvec_0 if transpose else lib().orc_matvec_0
    f(xin, y, iss, bb, nax, nay, naz, mask)
    return y


def matvec_0_colored(x, iss, bb, nax, nay, naz, mask, threads=1):
    """matvec_0 over 8 parity colours of elements, `threads` OpenMP threads
    within a colour (oracle.c orc_matvec_0_colored)."""
    xin = np.array(x, np.float64, copy=True)
    y = np.zeros_like(xin)
    if lib().orc_matvec_0_colored(xin, y, iss, bb, nax, nay, naz, mask, int(threads)) != 0:
        raise MemoryError("orc_matvec_0_colored")
    return y


def mttvec_0_colored(x, iss, bb, nax, nay, naz, mask, threads=1):
    """mttvec_0 over 8 parity colours (ompbase.c orc_mttvec_0_colored)."""
    xin = np.array(x, np.float64, copy=True)
    y = np.zeros_like(xin)
    if lib().orc_mttvec_0_colored(xin, y, iss, bb, nax, nay, naz, mask, int(threads)) != 0:
        raise MemoryError("orc_mttvec_0_colored")
    return y


def dbcg_simple_omp(b, x0, tol, itmax, bb, iss, diag, nax, nay, naz, mask, threads):
    """dbcg_simple with colour-parallel products and fixed-order OpenMP
    reductions (ompbase.c); returns (x, iterations, err, rc)."""
    x = np.array(x0, np.float64, copy=True)
    it = C.c_int(0)
    err = C.c_double(0.0)
    rc = lib().orc_dbcg_simple_omp(len(b), np.ascontiguousarray(b, np.float64), x, tol,
                                   itmax, C.byref(it), C.byref(err), bb, iss,
                                   np.ascontiguousarray(diag, np.float64), nax, nay, naz,
                                   mask, int(threads))
    if rc < 0:
        raise MemoryError("orc_dbcg_simple_omp")
    return x, it.value, err.value, rc


def dbcg_simple(b, x0, tol, itmax, bb, iss, diag, nax, nay, naz, mask):
    x = np.array(x0, np.float64, copy=True)
    it = C.c_int(0)
    err = C.c_double(0.0)
    rc = lib().orc_dbcg_simple(len(b), np.ascontiguousarray(b, np.float64), x, tol, itmax,
                               C.byref(it), C.byref(err), bb, iss,
                               np.ascontiguousarray(diag, np.float64), nax, nay, naz, mask)
    return x, it.value, err.value, rc


def bicgstab_jacobi(b, x0, tol, itmax, bb, iss, diag, nax, nay, naz, mask, threads=0):
    """Jacobi-BiCGSTAB, right preconditioned (van der Vorst 1992; the
    preconditioned BiCGSTAB of Barrett et al., Templates, 1994, sec. 2.3.8 with
    M = D), with the stopping rule of dbcg_simple (recursive ||r||/||b|| <= tol,
    VANADIS.FOR:977-1055, cudatest.cu:491-553) tested before each iteration.
    Not in the reference: the reference solution for the library's BiCGSTAB method.
    Shadow residual r~ = r0.  Returns (x, iterations, err, status) with
    status 1 converged, 2 itmax, 3 breakdown."""
    if threads > 0:
        A = lambda v: matvec_0_colored(v, iss, bb, nax, nay, naz, mask, threads)  # noqa: E731
    else:
        A = lambda v: matvec_0(v, iss, bb, nax, nay, naz, mask)  # noqa: E731
    d = np.asarray(diag, np.float64)
    x = np.array(x0, np.float64, copy=True)
    x[np.asarray(mask) == 1] = 0.0
    r = np.asarray(b, np.float64) - A(x)
    rt = r.copy()
    bn = np.linalg.norm(b)
    rho_old = alpha = omega = 1.0
    p = np.zeros_like(r)
    v = np.zeros_like(r)
    k = 0
    while True:
        err = np.linalg.norm(r) / bn
        if err <= tol:
            return x, k, err, 1
        if k >= itmax:
            return x, k, err, 2
        if not np.isfinite(err):
            return x, k, err, 3
        rho = rt @ r
        if k == 0:
            p = r.copy()
        else:
            beta = (rho / rho_old) * (alpha / omega)
            if not np.isfinite(beta) or rho == 0.0 or omega == 0.0:
                return x, k, err, 3
            p = r + beta * (p - omega * v)
        ph = p / d
        v = A(ph)
        gam = rt @ v
        if gam == 0.0 or not np.isfinite(rho / gam):
            return x, k, err, 3
        alpha = rho / gam
        s_ = r - alpha * v
        sh = s_ / d
        t = A(sh)
        tt = t @ t
        if tt == 0.0:
            # s = 0 (t = A D^-1 s vanishes only with s for a regular A): the
            # step solved the system; Templates' "s small: x += alpha p, stop"
            # at s = 0, taken as omega = 0 so that r = s and the test above
            # ends the loop (a singular A breaks down at beta)
            omega = 0.0
        elif not np.isfinite((t @ s_) / tt):
            return x, k, err, 3
        else:
            omega = (t @ s_) / tt
        x = x + alpha * ph + omega * sh
        r = s_ - omega * t
        rho_old = rho
        k += 1


def mass(iss, xx, yy, zz, c, nelem):
    return lib().orc_integrate_mass_hex8(nelem, iss, xx, yy, zz,
                                         np.ascontiguousarray(c, np.float64))


def ddf(nelem, nnode, iss, bb):
    return lib().orc_diagonal_dominance_factor(nelem, nnode, iss, bb)


def is_dd(nelem, nnode, iss, bb):
    return bool(lib().orc_is_diagonally_dominant(nelem, nnode, iss, bb))


def colors(nax, nay, naz):
    E = (nax - 1) * (nay - 1) * (naz - 1)
    c = np.zeros(E, np.int32)
    lib().orc_generate_hex8_colors(nax, nay, naz, c)
    el = np.zeros(E, np.int32)
    off = np.zeros(9, np.int32)
    cnt = np.zeros(8, np.int32)
    lib().orc_build_color_element_lists(E, c, el, off, cnt)
    return c, el, off, cnt


# ---- the flang-built reference (oracle/_ref) ------------------------------
class RefLib:
    """Direct calls into the reference's own subroutines (VANADIS.FOR for V1,
    vanadis_full_DR_OpenMP_timing.for for V3), built by `make -C oracle ref`.
    COMMON blocks are read and written as the globals the compiler emits."""

    def __init__(self, which: str):
        path = os.path.join(REF_DIR, f"libvanadis_ref_{which}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        # lazy binding: the V1 build leaves VANADIS.FOR's call of the CUDA
        # solver (kernel_dbcg_simple_, from przest3d) unresolved, and no
        # harness calls przest3d
        self.L = C.CDLL(path, mode=os.RTLD_LAZY)
        self.which = which
        self.L.par_3d_()

    @staticmethod
    def available(which: str) -> bool:
        return os.path.exists(os.path.join(REF_DIR, f"libvanadis_ref_{which}.so"))

    def _common(self, name, n, ctype=C.c_double):
        return (ctype * n).in_dll(self.L, name)

    def set_cfg(self, Q, K, P, V, alpha):
        cfg = self._common("cfg_", 7)
        for i, v in enumerate((Q, K, P, V[0], V[1], V[2], alpha)):
            cfg[i] = v

    def set_sizes(self, nelem, nnode):
        cal = self._common("cal_el_", 10, C.c_int)
        cal[0] = nelem
        cal[1] = nnode

    def set_dt(self, idt: int):
        self._common("d_time_", 1, C.c_int)[0] = idt

    def el3d(self, kel, xe, ye, ze, kod, nax, nay, naz):
        """Run the reference el3d for element kel (1-based) with the given
        node coordinates; returns (H, C, F) row-major like the oracle."""
        wsp = self._common("wsp_lok_", 24)
        for j in range(8):
            wsp[j] = xe[j]
            wsp[8 + j] = ye[j]
            wsp[16 + j] = ze[j]
        t = (C.c_double * 6)()
        ic = C.c_int(100)
        k = C.c_int(kel)
        nx, ny, nz = C.c_int(nax), C.c_int(nay), C.c_int(naz)
        kod_c = np.ascontiguousarray(kod, np.int32)
        self.L.el3d_(C.byref(ic), C.byref(k), t, kod_c.ctypes.data_as(_pi), C.byref(nx),
                     C.byref(ny), C.byref(nz))
        skl = np.ctypeslib.as_array(self._common("skl_el_", 64 * 11 + 8))
        hma = skl[0:64].reshape(8, 8, order="F").copy()
        fmv = skl[128:136].copy()
        cma_n = skl[136 + 64 * 8: 136 + 64 * 9].reshape(8, 8, order="F").copy()
        return hma, cma_n, fmv

    def build_bc_nodes(self, iss, kod, nax, nay, naz):
        """The reference's Dirichlet mask (build_bc_nodes, VANADIS.FOR:2026-2117;
        V3 omp_timing.for:1719-1811); set_sizes first."""
        E = (nax - 1) * (nay - 1) * (naz - 1)
        m = np.full(nax * nay * naz, -7, np.int32)
        iss_c = np.ascontiguousarray(iss, np.int32)
        kod_c = np.ascontiguousarray(kod, np.int32)
        assert iss_c.size == 8 * E and kod_c.size == 6 * E
        nx, ny, nz = C.c_int(nax), C.c_int(nay), C.c_int(naz)
        self.L.build_bc_nodes_(iss_c.ctypes.data_as(_pi), kod_c.ctypes.data_as(_pi),
                               m.ctypes.data_as(_pi), C.byref(nx), C.byref(ny), C.byref(nz))
        return m

    def matvec_0(self, x, iss, bb, nax, nay, naz, mask, transpose=False):
        xin = np.array(x, np.float64, copy=True)
        y = np.zeros_like(xin)
        nx, ny, nz = C.c_int(nax), C.c_int(nay), C.c_int(naz)
        f = self.L.mttvec_0_ if transpose else self.L.matvec_0_
        f(xin.ctypes.data_as(_pd), y.ctypes.data_as(_pd), iss.ctypes.data_as(_pi),
          bb.ctypes.data_as(_pd), C.byref(nx), C.byref(ny), C.byref(nz),
          mask.ctypes.data_as(_pi))
        return y

    def dbcg_simple(self, b, x0, tol, itmax, bb, iss, diag, nax, nay, naz, mask):
        """The reference CPU BiCG; run on a thread with a large stack because
        its work vectors are automatic arrays."""
        out = {}

        def run():
            x = np.array(x0, np.float64, copy=True)
            n = C.c_int(len(b))
            t = C.c_double(tol)
            im = C.c_int(itmax)
            it = C.c_int(0)
            err = C.c_double(0.0)
            nx, ny, nz = C.c_int(nax), C.c_int(nay), C.c_int(naz)
            bc = np.ascontiguousarray(b, np.float64)
            dg = np.ascontiguousarray(diag, np.float64)
            self.L.dbcg_simple_(C.byref(n), bc.ctypes.data_as(_pd), x.ctypes.data_as(_pd),
                                C.byref(t), C.byref(im), C.byref(it), C.byref(err),
                                bb.ctypes.data_as(_pd), iss.ctypes.data_as(_pi),
                                dg.ctypes.data_as(_pd), C.byref(nx), C.byref(ny),
                                C.byref(nz), mask.ctypes.data_as(_pi))
            out["r"] = (x, it.value, err.value)

        old = threading.stack_size()
        need = max(64 << 20, 6 * 8 * len(b) + (16 << 20))
        threading.stack_size(min(need, 1 << 30))
        try:
            th = threading.Thread(target=run)
            th.start()
            th.join()
        finally:
            threading.stack_size(old)
        return out["r"]

    def mass(self, iss, xx, yy, zz, c, nelem):
        m = C.c_double(0)
        ne = C.c_int(nelem)
        cc = np.ascontiguousarray(c, np.float64)
        self.L.integrate_mass_hex8_(C.byref(ne), iss.ctypes.data_as(_pi),
                                    xx.ctypes.data_as(_pd), yy.ctypes.data_as(_pd),
                                    zz.ctypes.data_as(_pd), cc.ctypes.data_as(_pd), C.byref(m))
        return m.value

    def ddf(self, nelem, nnode, iss, bb):
        f = self.L.diagonal_dominance_factor_
        f.restype = C.c_double
        ne, nn = C.c_int(nelem), C.c_int(nnode)
        return f(C.byref(ne), C.byref(nn), iss.ctypes.data_as(_pi), bb.ctypes.data_as(_pd))


def net_grads(nx, ny, nz, x1, x2, src, vx, vy, vz, target, mask, in_scale, out_scale, w, b,
              quant=True, extrap=0.0):
    """Loss, weight/bias gradients and output channel 0 of the learned-guess
    network (netoracle.c; quant mirrors the GPU's BF16 storage points)."""
    N = nx * ny * nz
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
    loss = C.c_double(0)
    dw = np.zeros(5 * 27 * 256)
    db = np.zeros(5 * 16)
    out = np.zeros(N)
    rc = lib().orc_net_grads(nx, ny, nz, f(x1), f(x2), f(src), f(vx), f(vy), f(vz), f(target),
                             np.ascontiguousarray(mask, dtype=np.int32),
                             np.ascontiguousarray(in_scale, dtype=np.float32), float(out_scale),
                             float(extrap),
                             np.ascontiguousarray(w, dtype=np.float32),
                             np.ascontiguousarray(b, dtype=np.float32), int(bool(quant)),
                             C.byref(loss), dw, db, out)
    if rc:
        raise MemoryError("orc_net_grads")
    return loss.value, dw, db, out



# ---- solution sketches of the large workloads (tests/golden/oracle_sketch.json)
def sketch_projections(x, k, seed0):
    """g_i . x for the seeded Gaussian vectors g_i = default_rng(seed0 + i)
    .standard_normal(len(x)), i < k (tests/golden/make_oracle_sketch.py)."""
    x = np.ascontiguousarray(x, np.float64)
    return np.array([np.dot(np.random.default_rng(seed0 + i).standard_normal(x.size), x)
                     for i in range(k)])


def sketch_rel_l2(x, entry):
    """Estimate of |x - x_o| / |x_o| for the oracle solution x_o that `entry`
    sketches: sqrt(mean_i (g_i . x - g_i . x_o)^2) / |x_o|.  E (g . v)^2 =
    |v|^2, so with k projections the squared estimate is |v|^2 chi2_k / k
    (k = 32: within -27 % / +27 % of the true value with 99 % probability;
    k = 16: -40 % / +41 %)."""
    s = np.asarray(entry["sketch"], np.float64)
    d = sketch_projections(x, len(s), entry["sketch_seed0"]) - s
    return float(np.sqrt(np.mean(d * d)) / entry["x_norm2"])


def load_sketches(path=None):
    import json
    if path is None:
        path = os.path.join(os.path.dirname(HERE), "tests", "golden", "oracle_sketch.json")
    with open(path) as f:
        return json.load(f)
