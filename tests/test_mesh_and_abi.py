"""Host logic (mesh, connectivity, masks, colouring, seeded fields) and the C
ABI of libvanadis_gpu.so (loads, exports every declared symbol).  CPU only."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("dims", [(2, 2, 2), (7, 20, 20), (33, 33, 17), (5, 9, 4)])
def test_connectivity_bit_exact(pkg, po, dims):
    nax, nay, naz = dims
    iss = pkg.problem.w_el(nax, nay, naz)
    assert np.array_equal(iss, po.w_el(nax, nay, naz))
    # closed form 1 + i + NAX j + NAX NAY k (+1, +NAX, +NAX NAY offsets)
    E = (nax - 1) * (nay - 1) * (naz - 1)
    e = np.arange(E)
    i, j, k = e % (nax - 1), (e // (nax - 1)) % (nay - 1), e // ((nax - 1) * (nay - 1))
    assert np.array_equal(iss[:E], 1 + i + nax * j + nax * nay * k)
    assert np.array_equal(iss[6 * E:7 * E], iss[:E] + 1 + nax + nax * nay)


@pytest.mark.parametrize("dims", [(7, 20, 20), (4, 6, 5)])
@pytest.mark.parametrize("facemap", [1, 3])
def test_face_codes_and_masks(pkg, po, dims, facemap):
    nax, nay, naz = dims
    kod = pkg.problem.kod_el_dr(nax, nay, naz)
    assert np.array_equal(kod, po.kod_el_dr(nax, nay, naz))
    iss = po.w_el(nax, nay, naz)
    m = pkg.problem.build_bc_nodes(iss, kod, nax, nay, naz, facemap)
    assert np.array_equal(m, po.build_bc_nodes(facemap, iss, kod, nax, nay, naz))
    n = np.arange(nax * nay * naz)
    gi, gj, gk = n % nax, (n // nax) % nay, n // (nax * nay)
    if facemap == 1:  # every face but the ground x = 0
        expect = (gk == 0) | (gk == naz - 1) | (gi == nax - 1) | (gj == 0) | (gj == nay - 1)
        assert np.array_equal(m.astype(bool), expect)
    else:  # V3 masks the planes one element inside y-min / y-max instead
        assert m.sum() > 0


@pytest.mark.parametrize("dims", [(7, 20, 20), (33, 33, 17), (6, 5, 4)])
def test_coloring_bit_exact_and_disjoint(pkg, po, dims):
    nax, nay, naz = dims
    col, el, off, cnt = po.colors(nax, nay, naz)
    E = (nax - 1) * (nay - 1) * (naz - 1)
    e = np.arange(E)
    i, j, k = e % (nax - 1), (e // (nax - 1)) % (nay - 1), e // ((nax - 1) * (nay - 1))
    assert np.array_equal(col, (i & 1) + 2 * (j & 1) + 4 * (k & 1))
    # stable counting sort
    assert np.array_equal(el, np.argsort(col, kind="stable"))
    assert np.array_equal(off[1:] - off[:-1], cnt)
    iss = po.w_el(nax, nay, naz).reshape(8, E)
    for c in range(8):
        nodes = iss[:, el[off[c]:off[c + 1]]].ravel()
        assert len(np.unique(nodes)) == len(nodes)  # same colour shares no node
    if dims == (7, 20, 20):  # closed form of the colour counts for the shipped mesh
        assert list(cnt) == [300, 300, 270, 270, 270, 270, 243, 243]


@pytest.mark.parametrize("dims", [(7, 20, 20), (33, 33, 17), (6, 5, 4), (2, 2, 2)])
def test_library_colour_lists_bit_exact(pkg, po, dims):
    """The colouring and per-colour element lists the library's legacy colour
    path builds (vgpu_legacy_colors; host code, no device) equal the oracle's
    generate_hex8_colors / build_color_element_lists bit for bit."""
    nax, nay, naz = dims
    col, el, off, cnt = pkg._lib.legacy_colors(nax, nay, naz)
    ocol, oel, ooff, ocnt = po.colors(nax, nay, naz)
    assert np.array_equal(col, ocol) and np.array_equal(el, oel)
    assert np.array_equal(off, ooff) and np.array_equal(cnt, ocnt)


def test_splitmix_determinism(pkg):
    a = pkg.problem.splitmix64(42, 10)
    b = pkg.problem.splitmix64(42, 20)[:10]
    assert np.array_equal(a, b)
    u = pkg.problem.uniform01(1, 100000)
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01


def test_unit_box_fields(pkg):
    p = pkg.problem.config_problem("c1")
    assert (p.nax, p.nay, p.naz) == (33, 33, 17) and p.N == 18513 and p.E == 16384
    vx, vy, vz = p.vel.reshape(3, -1)
    assert np.abs(vx - 1.0).max() <= 0.2 + 1e-12
    assert abs(vx.mean() - 1.0) < 0.05 and abs(vy.mean() - 0.5) < 0.05
    # mesh Peclet |v| h / (2 lam) ~ 1.1/32/2e-3 ~ 17 (per-axis h = 1/32, 1/16)
    h = 1.0 / 32
    assert 10 < np.sqrt(vx ** 2 + vy ** 2 + vz ** 2).mean() * h / (2 * 1e-3) < 40
    assert np.isclose(p.q_in.sum(), (2 * np.pi) ** 1.5 * 0.1 ** 3, rtol=0.05)


def _declared_functions():
    txt = open(os.path.join(ROOT, "include", "vanadis_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b((?:vgpu_|kernel_)\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(pkg):
    import ctypes
    lib = pkg.load()
    names = _declared_functions()
    assert "kernel_dbcg_simple_" in names and len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(pkg.EXPORTED)
    assert b"gfx950" in lib.vgpu_version()
    n = ctypes.c_int(-1)
    assert lib.vgpu_device_count(ctypes.byref(n)) == 0 and n.value >= 0


def test_library_is_gfx950_code_object():
    so = os.path.join(ROOT, "vanadis_solver-directional_residual_amd", "libvanadis_gpu.so")
    data = open(so, "rb").read()
    assert b"gfx950" in data


def test_fortran_interface_module_compiles(tmp_path):
    """The ISO_C_BINDING module a Fortran host uses to call the library."""
    import shutil
    import subprocess
    flang = shutil.which("flang") or shutil.which("amdflang")
    if flang is None:
        pytest.skip("no Fortran compiler")
    src = os.path.join(ROOT, "vanadis_solver-directional_residual_amd", "fortran",
                       "vanadis_gpu_iface.f90")
    r = subprocess.run([flang, "-c", src, "-o", str(tmp_path / "m.o"),
                        "-module-dir", str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
