"""The device element arithmetic (vgpu_elem.hpp, compiled for the host with
g++) against the oracle's el3d on random box elements: A = cH H + cC C with
mu = cH P + cC and Robin weight cH alpha, forward and transpose."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    out = tmp_path_factory.mktemp("elem") / "libelem.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                    "-o", str(out), os.path.join(HERE, "host", "elem_shim.cpp")], check=True)
    L = C.CDLL(str(out))
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    for fn in (L.elem_matrix, L.elem_matrix_faces):
        fn.argtypes = [dp, C.c_double, dp, C.c_double, C.c_double, C.c_double,
                       C.c_double, C.c_int, C.c_int, dp]
    return L


LOC = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


@pytest.mark.parametrize("seed", range(12))
@pytest.mark.parametrize("scheme", ["theta", "steady", "euler"])
def test_device_element_matrix_matches_oracle(shim, po, seed, scheme):
    rng = np.random.default_rng(seed)
    h = rng.uniform(0.01, 2.0, 3)
    v = rng.uniform(-3, 3, 3)
    lam = 10 ** rng.uniform(-3, 0)
    P = rng.uniform(0, 0.2) if seed % 3 else 0.0
    alpha = rng.uniform(0, 0.5)
    robin = seed % 2
    dt = rng.uniform(0.01, 2.0)
    cH, cC = {"theta": (2.0, 3.0 / dt), "steady": (1.0, 0.0), "euler": (1.0, 1.0 / dt)}[scheme]
    x0 = rng.uniform(-10, 10, 3)
    xe = np.array([x0[0] + h[0] * c[0] for c in LOC])
    ye = np.array([x0[1] + h[1] * c[1] for c in LOC])
    ze = np.array([x0[2] + h[2] * c[2] for c in LOC])
    H, Cm, _ = po.el3d(1, xe, ye, ze, np.full(3, lam), P, v, 0.0, robin, alpha, 0.0)
    Aref = cH * H + cC * Cm
    # tau exactly as the oracle evaluates it at a Gauss point
    import ctypes
    n = np.zeros(8); w = np.zeros(8); d1 = np.zeros(8); d2 = np.zeros(8); d3 = np.zeros(8)
    dj = ctypes.c_double(); tau = ctypes.c_double()
    po.lib().orc_gauss_point(1, 0.5, 0.5, 0.5, xe, ye, ze, np.full(3, lam), v,
                             ctypes.byref(dj), n, w, d1, d2, d3, ctypes.byref(tau))
    # the per-element path of the brick kernels and the factored face path of
    # the marching owner kernel
    for fn in (shim.elem_matrix, shim.elem_matrix_faces):
        for tr in (0, 1):
            A = np.zeros(64)
            fn(h, lam, v, tau.value, cH, cH * P + cC, cH * alpha, robin, tr, A)
            A = A.reshape(8, 8)
            ref = Aref.T if tr else Aref
            err = np.abs(A - ref).max()
            assert err <= 1e-13 * np.abs(ref).max(), (fn.__name__, tr, err)
