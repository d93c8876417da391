"""ctypes binding of libvanadis_gpu.so (C ABI declared in include/vanadis_gpu.h).

The library is built in place next to this file by the package Makefile.  If it
is missing or fails to load, every entry point raises: there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvanadis_gpu.so")

# status codes (vanadis_gpu.h)
OK, ERR_ARG, ERR_HIP, ERR_MESH, ERR_STATE, NOT_CONVERGED, BREAKDOWN, ERR_NOMEM = range(8)
SCATTER_OWNER, SCATTER_COLORED, SCATTER_ATOMIC = 0, 1, 2
DR_V1, DR_V2, DR_V3 = 1, 2, 3
SCHEME_THETA, SCHEME_STEADY, SCHEME_EULER = 0, 1, 2
FP64, FP32, MIXED = 0, 1, 2
METHOD_BICG, METHOD_BICGSTAB = 0, 1
_METHOD_NAMES = {"bicg": METHOD_BICG, "bicgstab": METHOD_BICGSTAB}

_SCATTER_NAMES = {"owner": SCATTER_OWNER, "colored": SCATTER_COLORED, "atomic": SCATTER_ATOMIC}

# every symbol include/vanadis_gpu.h declares
EXPORTED = (
    "vgpu_last_error", "vgpu_version", "vgpu_device_count", "kernel_dbcg_simple_",
    "vgpu_legacy_set_scatter", "vgpu_legacy_release", "vgpu_create", "vgpu_destroy",
    "vgpu_get_sizes", "vgpu_set_coeffs", "vgpu_set_scheme", "vgpu_set_options", "vgpu_apply",
    "vgpu_apply_dev", "vgpu_diag", "vgpu_assemble_rhs", "vgpu_solve", "vgpu_solve_dev",
    "vgpu_true_residual", "vgpu_mass", "vgpu_dev_alloc", "vgpu_dev_free", "vgpu_memcpy_h2d",
    "vgpu_memcpy_d2h", "vgpu_time_kernel", "vgpu_autotune", "vgpu_batch_create",
    "vgpu_batch_destroy", "vgpu_batch_solve", "vgpu_batch_layout", "vgpu_batch_solve_dev",
    "vgpu_assemble_rhs_dev", "vgpu_gaussian_dev", "vgpu_net_create", "vgpu_net_destroy",
    "vgpu_net_set_params", "vgpu_net_get_params", "vgpu_net_get_grads", "vgpu_net_set_static",
    "vgpu_net_predict_dev", "vgpu_net_train_step_dev", "vgpu_net_reset_optimizer",
    "vgpu_copy_dev", "vgpu_axpby_dev", "vgpu_residual_norm_dev", "vgpu_net_set_loss",
    "vgpu_jacobi_dev", "vgpu_net_set_smoothing", "vgpu_ddf", "vgpu_write_tfile",
    "vgpu_set_method", "vgpu_axpbypcz_dev", "vgpu_net_time_layers",
    "vgpu_net_set_zchunk", "vgpu_batch_set_fallback", "vgpu_legacy_colors",
    "vgpu_legacy_get_color_lists", "vgpu_legacy_last_times", "kernel_dbcg_simple_noalloc",
    "vgpu_noalloc_max_nodes", "vgpu_autotune_bricks",
)


class VgpuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"vgpu error {code}: {msg}")
        self.code = code


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int)
_vp = C.c_void_p

_lib = None


def scatter_code(s) -> int:
    return _SCATTER_NAMES[s] if isinstance(s, str) else int(s)


def load():
    """Load the HIP library (raises if it is absent: no fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not built: run __graft_entry__.build() "
                          "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    L.vgpu_last_error.restype = C.c_char_p
    L.vgpu_version.restype = C.c_char_p
    L.vgpu_device_count.argtypes = [_pi]
    L.kernel_dbcg_simple_.argtypes = [_pi, _dp, _dp, _pd, _pi, _pi, _pd, _dp, _ip, _dp, _pi,
                                      _pi, _pi, _ip]
    L.kernel_dbcg_simple_.restype = None
    L.vgpu_legacy_set_scatter.argtypes = [C.c_int]
    L.vgpu_legacy_colors.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _ip]
    L.vgpu_legacy_get_color_lists.argtypes = [C.c_int64, _ip, _ip, _ip]
    L.vgpu_legacy_last_times.argtypes = [_pd, _pd, _pd]
    L.kernel_dbcg_simple_noalloc.argtypes = [C.c_int, _vp, _vp, C.c_double, C.c_int, _pi, _pd,
                                             _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp,
                                             C.c_int, _vp]
    L.vgpu_noalloc_max_nodes.argtypes = [C.POINTER(C.c_int64)]
    L.vgpu_create.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _vp, _ip,
                              C.POINTER(_vp)]
    L.vgpu_destroy.argtypes = [_vp]
    L.vgpu_get_sizes.argtypes = [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.vgpu_set_coeffs.argtypes = [_vp, _dp, _dp, C.c_double, C.c_double, C.c_int, C.c_int,
                                  C.c_double]
    L.vgpu_set_scheme.argtypes = [_vp, C.c_int, C.c_double]
    L.vgpu_set_options.argtypes = [_vp, C.c_int, C.c_int, C.c_int]
    L.vgpu_set_method.argtypes = [_vp, C.c_int]
    L.vgpu_apply.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp]
    L.vgpu_apply_dev.argtypes = [_vp, C.c_int, C.c_int, _vp, _vp]
    L.vgpu_diag.argtypes = [_vp, _dp]
    L.vgpu_assemble_rhs.argtypes = [_vp, _dp, _vp, _dp]
    L.vgpu_solve.argtypes = [_vp, C.c_int, _dp, _dp, C.c_double, C.c_int, C.c_double, _pi, _pd]
    L.vgpu_solve_dev.argtypes = [_vp, C.c_int, _vp, _vp, C.c_double, C.c_int, C.c_double,
                                 _pi, _pd, _pd]
    L.vgpu_true_residual.argtypes = [_vp, _dp, _dp, _pd]
    L.vgpu_mass.argtypes = [_vp, _dp, _pd]
    L.vgpu_dev_alloc.argtypes = [_vp, C.c_int64, C.POINTER(_vp)]
    L.vgpu_dev_free.argtypes = [_vp, _vp]
    L.vgpu_memcpy_h2d.argtypes = [_vp, _vp, _vp, C.c_int64]
    L.vgpu_memcpy_d2h.argtypes = [_vp, _vp, _vp, C.c_int64]
    L.vgpu_time_kernel.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _pd]
    L.vgpu_autotune.argtypes = [_vp, C.c_int, C.c_int, _pi, _pi, _pi, _pd]
    L.vgpu_autotune_bricks.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _pi, _pi, _pd]
    L.vgpu_batch_create.argtypes = [C.c_int, C.POINTER(_vp), C.POINTER(_vp)]
    L.vgpu_batch_destroy.argtypes = [_vp]
    L.vgpu_batch_solve.argtypes = [_vp, C.c_int, C.POINTER(_vp), C.POINTER(_vp), C.c_double,
                                   C.c_int, _pi, _pd, _pi]
    L.vgpu_batch_layout.argtypes = [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.vgpu_batch_solve_dev.argtypes = [_vp, C.c_int, _vp, _vp, C.c_double, C.c_int, _pi, _pd,
                                       _pi, _pd]
    L.vgpu_assemble_rhs_dev.argtypes = [_vp, _vp, _vp, _vp]
    L.vgpu_gaussian_dev.argtypes = [_vp, C.c_double, C.c_double, C.c_double, C.c_double, _vp,
                                    _vp]
    _fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    L.vgpu_net_create.argtypes = [_vp, C.POINTER(_vp)]
    L.vgpu_net_destroy.argtypes = [_vp]
    L.vgpu_net_set_params.argtypes = [_vp, _fp, _fp]
    L.vgpu_net_get_params.argtypes = [_vp, _fp, _fp]
    L.vgpu_net_get_grads.argtypes = [_vp, _fp, _fp]
    L.vgpu_net_set_static.argtypes = [_vp, _dp, _dp, _dp, _fp, C.c_double, C.c_double]
    L.vgpu_copy_dev.argtypes = [_vp, _vp, _vp, C.c_int64]
    L.vgpu_axpby_dev.argtypes = [_vp, C.c_double, _vp, C.c_double, _vp, _vp]
    L.vgpu_axpbypcz_dev.argtypes = [_vp, C.c_double, _vp, C.c_double, _vp, C.c_double, _vp,
                                    _vp]
    L.vgpu_residual_norm_dev.argtypes = [_vp, _vp, _vp, _pd]
    L.vgpu_jacobi_dev.argtypes = [_vp, _vp, _vp, C.c_int, C.c_double]
    L.vgpu_ddf.argtypes = [_vp, _pd, _pi]
    L.vgpu_write_tfile.argtypes = [C.c_char_p, C.c_int64, _dp, _dp, _dp, _dp]
    L.vgpu_net_predict_dev.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.vgpu_net_train_step_dev.argtypes = [_vp, _vp, _vp, _vp, _vp, C.c_float, _pd]
    L.vgpu_net_reset_optimizer.argtypes = [_vp]
    L.vgpu_net_set_loss.argtypes = [_vp, C.c_int]
    L.vgpu_net_set_smoothing.argtypes = [_vp, C.c_int]
    L.vgpu_net_time_layers.argtypes = [_vp, C.c_int, C.c_int, _pd]
    L.vgpu_net_set_zchunk.argtypes = [_vp, C.c_int]
    L.vgpu_batch_set_fallback.argtypes = [_vp, C.c_int, C.c_int]
    _lib = L
    return L


def check(rc: int, what: str = "") -> int:
    if rc not in (OK,):
        msg = load().vgpu_last_error().decode(errors="replace")
        raise VgpuError(rc, f"{what}: {msg}" if what else msg)
    return rc


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# --------------------------------------------------------------------------
# reference entry point: CALL kernel_dbcg_simple(NON, FMV3, TNW2, TOL, IL_I,
# ITER, ERR, BB, ISS, DIAG, NAX, NAY, NAZ, bc_dirichlet)  (VANADIS.FOR:809-810)
# --------------------------------------------------------------------------
def kernel_dbcg_simple(non, fmv3, tnw2, tol, il_i, bb, iss, diag, nax, nay, naz,
                       bc_dirichlet, scatter="colored"):
    """Legacy drop-in: returns (x, iter, err).  tnw2 is not modified (the
    Fortran host passes it in/out; here the updated copy is returned)."""
    L = load()
    check(L.vgpu_legacy_set_scatter(scatter_code(scatter)), "legacy scatter")
    x = np.array(tnw2, dtype=np.float64, copy=True)
    n = C.c_int(int(non))
    t = C.c_double(float(tol))
    im = C.c_int(int(il_i))
    it = C.c_int(0)
    err = C.c_double(0.0)
    nx, ny, nz = C.c_int(nax), C.c_int(nay), C.c_int(naz)
    L.kernel_dbcg_simple_(C.byref(n), _f64(fmv3), x, C.byref(t), C.byref(im), C.byref(it),
                          C.byref(err), _f64(bb), _i32(iss), _f64(diag), C.byref(nx),
                          C.byref(ny), C.byref(nz), _i32(bc_dirichlet))
    return x, it.value, err.value


def legacy_colors(nax, nay, naz):
    """The legacy path's colouring (color[E], elems[E], offsets[9],
    counts[8]), computed by the library on the host."""
    E = (nax - 1) * (nay - 1) * (naz - 1)
    color = np.zeros(E, np.int32)
    elems = np.zeros(E, np.int32)
    off = np.zeros(9, np.int32)
    cnt = np.zeros(8, np.int32)
    check(load().vgpu_legacy_colors(nax, nay, naz, color, elems, off, cnt), "legacy colors")
    return color, elems, off, cnt


def legacy_last_times():
    """(upload_s, solve_s, download_s) of the last kernel_dbcg_simple call."""
    u, s_, d = C.c_double(), C.c_double(), C.c_double()
    check(load().vgpu_legacy_last_times(C.byref(u), C.byref(s_), C.byref(d)), "legacy times")
    return u.value, s_.value, d.value


def legacy_device_color_lists(E):
    """The element lists the legacy path holds on the device (read back after
    a kernel_dbcg_simple call with the colour scatter), offsets, counts."""
    elems = np.zeros(E, np.int32)
    off = np.zeros(9, np.int32)
    cnt = np.zeros(8, np.int32)
    check(load().vgpu_legacy_get_color_lists(E, elems, off, cnt), "legacy color lists")
    return elems, off, cnt


class Context:
    """Device-resident matrix-free operator and solver for one structured mesh
    (the vgpu_* context API)."""

    def __init__(self, nax, nay, naz, xx, yy, zz, iss, bc):
        L = load()
        h = _vp()
        iss_arr = None if iss is None else _i32(iss)
        check(L.vgpu_create(nax, nay, naz, _f64(xx), _f64(yy), _f64(zz),
                            None if iss_arr is None else iss_arr.ctypes.data_as(_vp),
                            _i32(bc), C.byref(h)), "vgpu_create")
        self._keep = iss_arr
        self._children = weakref.WeakSet()  # nets and batches built on this context
        self.h = h
        self.nax, self.nay, self.naz = nax, nay, naz
        self.N = nax * nay * naz
        self.E = (nax - 1) * (nay - 1) * (naz - 1)

    @classmethod
    def from_problem(cls, p, coeffs=True):
        c = cls(p.nax, p.nay, p.naz, p.xx, p.yy, p.zz, p.iss, p.mask)
        if coeffs:
            c.set_coeffs(p.lam, p.vel, p.P, p.alpha, p.variant, p.scheme, p.dt)
        return c

    def close(self):
        # networks and batches built on this context go first (a collector
        # may finalise the context before them)
        for child in list(getattr(self, "_children", ())):
            child.close()
        if getattr(self, "h", None):
            load().vgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_coeffs(self, lam, vel, P=0.0, alpha=0.0, variant=1, scheme=1, dt=1.0):
        check(load().vgpu_set_coeffs(self.h, _f64(lam), _f64(vel), float(P), float(alpha),
                                     int(variant), int(scheme), float(dt)), "set_coeffs")

    def set_scheme(self, scheme, dt):
        check(load().vgpu_set_scheme(self.h, int(scheme), float(dt)), "set_scheme")

    def set_options(self, scatter="owner", zchunk=-1, tile=-1):
        """Scatter mode; z chunk (-1 keep, 0 balanced over resident workgroups,
        n > 0 one workgroup per (chunk of n node planes, tile) group, chunk by
        chunk); owner tile shape (-1 keep)."""
        check(load().vgpu_set_options(self.h, scatter_code(scatter), int(zchunk), int(tile)),
              "set_options")

    def set_method(self, method="bicg"):
        """Krylov method of the FP64 / FP32 solves: "bicg" (the reference's
        dbcg_simple) or "bicgstab" (Jacobi-BiCGSTAB on the owner launches)."""
        code = _METHOD_NAMES[method] if isinstance(method, str) else int(method)
        check(load().vgpu_set_method(self.h, code), "set_method")

    def apply(self, x, scatter="owner", precision=FP64, transpose=False):
        y = np.zeros(self.N)
        check(load().vgpu_apply(self.h, scatter_code(scatter), int(precision), int(transpose),
                                _f64(x), y), "apply")
        return y

    def diag(self):
        d = np.zeros(self.N)
        check(load().vgpu_diag(self.h, d), "diag")
        return d

    def assemble_rhs(self, q_in, t_old=None):
        b = np.zeros(self.N)
        t = None if t_old is None else _f64(t_old)
        check(load().vgpu_assemble_rhs(self.h, _f64(q_in),
                                       None if t is None else t.ctypes.data_as(_vp), b),
              "assemble_rhs")
        return b

    def solve(self, b, x0=None, rtol=1e-8, itmax=20000, precision=FP64, inner_rtol=0.0,
              raise_on_fail=True):
        """Returns (x, iters, err, status code)."""
        x = np.zeros(self.N) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
        it = C.c_int(0)
        err = C.c_double(0.0)
        rc = load().vgpu_solve(self.h, int(precision), _f64(b), x, float(rtol), int(itmax),
                               float(inner_rtol), C.byref(it), C.byref(err))
        if raise_on_fail and rc != OK:
            check(rc, "solve")
        return x, it.value, err.value, rc

    def true_residual(self, b, x):
        r = C.c_double(0)
        check(load().vgpu_true_residual(self.h, _f64(b), _f64(x), C.byref(r)), "true_residual")
        return r.value

    def ddf(self):
        """(diagonal dominance factor, all rows dominant) of the element matrices."""
        d = C.c_double(0)
        dom = C.c_int(0)
        check(load().vgpu_ddf(self.h, C.byref(d), C.byref(dom)), "ddf")
        return d.value, bool(dom.value)

    def mass(self, c):
        m = C.c_double(0)
        check(load().vgpu_mass(self.h, _f64(c), C.byref(m)), "mass")
        return m.value

    # ---- device-resident helpers (bench) ----
    def dev_alloc(self, nbytes):
        p = _vp()
        check(load().vgpu_dev_alloc(self.h, int(nbytes), C.byref(p)), "dev_alloc")
        return p

    def dev_free(self, ptr):
        """Free a buffer of dev_alloc (also freed with the context)."""
        check(load().vgpu_dev_free(self.h, ptr), "dev_free")

    def h2d(self, dst, arr):
        a = np.ascontiguousarray(arr)
        check(load().vgpu_memcpy_h2d(self.h, dst, a.ctypes.data_as(_vp), a.nbytes), "h2d")

    def d2h(self, src, n, dtype=np.float64):
        a = np.zeros(n, dtype=dtype)
        check(load().vgpu_memcpy_d2h(self.h, a.ctypes.data_as(_vp), src, a.nbytes), "d2h")
        return a

    def solve_dev(self, b_dev, x_dev, rtol=1e-8, itmax=20000, precision=FP64,
                  inner_rtol=0.0, timed=True):
        """Returns (iterations, err, seconds, status code); timed=False skips
        the solve's own timer (and the stream synchronisation before it):
        seconds is then 0."""
        it = C.c_int(0)
        err = C.c_double(0.0)
        sec = C.c_double(0.0)
        rc = load().vgpu_solve_dev(self.h, int(precision), b_dev, x_dev, float(rtol),
                                   int(itmax), float(inner_rtol), C.byref(it), C.byref(err),
                                   C.byref(sec) if timed else None)
        return it.value, err.value, sec.value, rc

    def time_kernel(self, what=0, scatter="owner", precision=FP64, reps=20):
        ms = C.c_double(0)
        check(load().vgpu_time_kernel(self.h, int(what), scatter_code(scatter), int(precision),
                                      int(reps), C.byref(ms)), "time_kernel")
        return ms.value

    def assemble_rhs_dev(self, q_dev, t_old_dev, b_dev):
        check(load().vgpu_assemble_rhs_dev(self.h, q_dev, t_old_dev, b_dev), "assemble_rhs_dev")

    def copy_dev(self, dst, src, nbytes):
        check(load().vgpu_copy_dev(self.h, dst, src, int(nbytes)), "copy_dev")

    def axpby_dev(self, a, x, b, y, out):
        check(load().vgpu_axpby_dev(self.h, float(a), x, float(b), y, out), "axpby_dev")

    def axpbypcz_dev(self, a, x, b, y, c, z, out):
        check(load().vgpu_axpbypcz_dev(self.h, float(a), x, float(b), y, float(c), z, out),
              "axpbypcz_dev")

    def residual_norm_dev(self, b, x):
        r = C.c_double(0)
        check(load().vgpu_residual_norm_dev(self.h, b, x, C.byref(r)), "residual_norm_dev")
        return r.value

    def jacobi_dev(self, b, x, sweeps=1, omega=0.8):
        check(load().vgpu_jacobi_dev(self.h, b, x, int(sweeps), float(omega)), "jacobi_dev")

    def gaussian_dev(self, centre, sigma, q_elem_dev=None, src_node_dev=None):
        check(load().vgpu_gaussian_dev(self.h, float(centre[0]), float(centre[1]),
                                       float(centre[2]), float(sigma), q_elem_dev, src_node_dev),
              "gaussian_dev")

    def autotune(self, precision=FP64, reps=20):
        """Time every owner tile shape / z chunk, keep the fastest; returns
        {"scatter", "tile", "zchunk", "ms"}."""
        sc, t, z = C.c_int(0), C.c_int(0), C.c_int(0)
        ms = C.c_double(0)
        check(load().vgpu_autotune(self.h, int(precision), int(reps), C.byref(sc), C.byref(t),
                                   C.byref(z), C.byref(ms)), "autotune")
        return {"scatter": sc.value, "tile": t.value, "zchunk": z.value, "ms": ms.value}


    def autotune_bricks(self, scatter="colored", precision=FP64, reps=20):
        """Time every brick shape / depth of the colored or atomic scatter's
        dual apply, keep the fastest (the context then solves with that
        scatter); returns {"scatter", "tile", "zchunk", "ms"}."""
        t, z = C.c_int(0), C.c_int(0)
        ms = C.c_double(0)
        check(load().vgpu_autotune_bricks(self.h, scatter_code(scatter), int(precision),
                                          int(reps), C.byref(t), C.byref(z), C.byref(ms)),
              "autotune_bricks")
        return {"scatter": scatter_code(scatter), "tile": t.value, "zchunk": z.value,
                "ms": ms.value}


class Batch:
    """Several contexts solved together (vgpu_batch_*): one persistent owner
    launch per BiCG iteration over all systems, one hipGraph per batch."""

    def __init__(self, contexts):
        self.ctxs = list(contexts)
        arr = (_vp * len(self.ctxs))(*[c.h for c in self.ctxs])
        h = _vp()
        check(load().vgpu_batch_create(len(self.ctxs), arr, C.byref(h)), "batch_create")
        self.h = h
        for c in self.ctxs:
            c._children.add(self)

    def close(self):
        if getattr(self, "h", None):
            load().vgpu_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def layout(self):
        off = (C.c_int64 * len(self.ctxs))()
        tot = C.c_int64(0)
        check(load().vgpu_batch_layout(self.h, off, C.byref(tot)), "batch_layout")
        return list(off), tot.value

    def set_fallback(self, method=None, itmax=20000):
        """FP64 members left at itmax or in breakdown by the batched BiCG, or
        stopped on the recursive residual with a true residual above rtol, are
        finished by BiCGSTAB on their own contexts ("bicgstab"; None: off)."""
        code = METHOD_BICG if method is None else _METHOD_NAMES[method]
        check(load().vgpu_batch_set_fallback(self.h, int(code), int(itmax)), "batch_set_fallback")

    def solve(self, bs, x0s=None, rtol=1e-8, itmax=20000, precision=FP64,
              raise_on_fail=True):
        """Returns (xs, iters, errs, status, rc); xs[m] is system m's solution."""
        nm = len(self.ctxs)
        bs = [_f64(b) for b in bs]
        xs = [np.zeros(c.N) if x0s is None else np.array(x0s[i], dtype=np.float64, copy=True)
              for i, c in enumerate(self.ctxs)]
        bp = (_vp * nm)(*[b.ctypes.data_as(_vp) for b in bs])
        xp = (_vp * nm)(*[x.ctypes.data_as(_vp) for x in xs])
        it = np.zeros(nm, dtype=np.int32)
        er = np.zeros(nm)
        st = np.zeros(nm, dtype=np.int32)
        rc = load().vgpu_batch_solve(self.h, int(precision), bp, xp, float(rtol), int(itmax),
                                     it.ctypes.data_as(_pi), er.ctypes.data_as(_pd),
                                     st.ctypes.data_as(_pi))
        if raise_on_fail and rc != OK:
            check(rc, "batch_solve")
        return xs, it, er, st, rc

    def solve_dev(self, b_dev, x_dev, rtol=1e-8, itmax=20000, precision=FP64):
        nm = len(self.ctxs)
        it = np.zeros(nm, dtype=np.int32)
        er = np.zeros(nm)
        st = np.zeros(nm, dtype=np.int32)
        sec = C.c_double(0)
        rc = load().vgpu_batch_solve_dev(self.h, int(precision), b_dev, x_dev, float(rtol),
                                         int(itmax), it.ctypes.data_as(_pi),
                                         er.ctypes.data_as(_pd), st.ctypes.data_as(_pi),
                                         C.byref(sec))
        return it, er, st, sec.value, rc


def device_count() -> int:
    n = C.c_int(0)
    load().vgpu_device_count(C.byref(n))
    return n.value


NET_LAYERS, NET_TAPS, NET_CH = 5, 27, 16
NET_W = NET_LAYERS * NET_TAPS * NET_CH * NET_CH
NET_B = NET_LAYERS * NET_CH


class Net:
    """The learned-guess network on a context's node grid (vgpu_net_*)."""

    def __init__(self, ctx):
        h = _vp()
        check(load().vgpu_net_create(ctx.h, C.byref(h)), "net_create")
        self.h = h
        self.ctx = ctx
        ctx._children.add(self)

    def close(self):
        if getattr(self, "h", None):
            load().vgpu_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, w, b):
        check(load().vgpu_net_set_params(self.h, np.ascontiguousarray(w, np.float32).ravel(),
                                         np.ascontiguousarray(b, np.float32).ravel()),
              "net_set_params")

    def get_params(self):
        w = np.zeros(NET_W, np.float32)
        b = np.zeros(NET_B, np.float32)
        check(load().vgpu_net_get_params(self.h, w, b), "net_get_params")
        return w, b

    def get_grads(self):
        w = np.zeros(NET_W, np.float32)
        b = np.zeros(NET_B, np.float32)
        check(load().vgpu_net_get_grads(self.h, w, b), "net_get_grads")
        return w, b

    def set_static(self, vx, vy, vz, in_scale, out_scale, extrap=0.0):
        check(load().vgpu_net_set_static(self.h, _f64(vx), _f64(vy), _f64(vz),
                                         np.ascontiguousarray(in_scale, np.float32),
                                         float(out_scale), float(extrap)), "net_set_static")

    def predict_dev(self, x1, x2, src, guess):
        check(load().vgpu_net_predict_dev(self.h, x1, x2, src, guess), "net_predict")

    def train_step_dev(self, x1, x2, src, target, lr):
        loss = C.c_double(0)
        check(load().vgpu_net_train_step_dev(self.h, x1, x2, src, target, float(lr),
                                             C.byref(loss)), "net_train_step")
        return loss.value

    def reset_optimizer(self):
        check(load().vgpu_net_reset_optimizer(self.h), "net_reset_optimizer")

    def set_smoothing(self, passes):
        """Passes of the [1 4 1]^3 / 216 low-pass filter on the output."""
        check(load().vgpu_net_set_smoothing(self.h, int(passes)), "net_set_smoothing")

    def set_loss(self, kind):
        """0: increment L2 loss; 1: initial-residual loss through the operator."""
        check(load().vgpu_net_set_loss(self.h, int(kind)), "net_set_loss")

    def set_zchunk(self, zchunk=0):
        """Node planes per wave of the layer kernels (0: automatic)."""
        check(load().vgpu_net_set_zchunk(self.h, int(zchunk)), "net_set_zchunk")

    def time_layers(self, what=0, reps=20):
        """Average ms of 0 a hidden forward layer, 1 a backward-data layer,
        2 a layer's weight gradients, 3 the whole prediction network."""
        ms = C.c_double(0)
        check(load().vgpu_net_time_layers(self.h, int(what), int(reps), C.byref(ms)),
              "net_time_layers")
        return ms.value


def write_tfile(path, xx, yy, zz, c):
    """The host's T-file: x, y, z, c*1e6 per node in Fortran e12.5."""
    xx, yy, zz, c = _f64(xx), _f64(yy), _f64(zz), _f64(c)
    check(load().vgpu_write_tfile(os.fsencode(path), len(c), xx, yy, zz, c), "write_tfile")

