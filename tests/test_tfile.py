"""The T-file writer (vgpu_write_tfile) against a Fortran compiler's own e12.5
output of the reference's format 122 (VANADIS.FOR:866), on random values over
many magnitudes and edge cases (zero, signed zero, 3-digit exponents, rounding
carries).  Pure host code: runs without a GPU."""
import os
import shutil
import subprocess

import numpy as np
import pytest

FORTRAN = """
program tfile
  implicit none
  integer :: n, i
  double precision, allocatable :: x(:), y(:), z(:), c(:)
  character(len=54) :: line
  open(10, file='in.txt', status='old')
  read(10, *) n
  allocate(x(n), y(n), z(n), c(n))
  do i = 1, n
    read(10, *) x(i), y(i), z(i), c(i)
  end do
  close(10)
  open(11, file='out.txt', status='unknown')
  do i = 1, n
    write(line, 122) x(i), y(i), z(i), c(i) * 1d6
    write(11, '(a54)') line
  end do
  close(11)
122 format (e12.5, 2x, e12.5, 2x, e12.5, 2x, e12.5)
end program
"""


def _values():
    rng = np.random.default_rng(9)
    v = rng.standard_normal(400) * 10.0 ** rng.integers(-130, 130, 400)
    edge = [0.0, -0.0, 1.0, 9.999996, 9.999994, 0.999995, 123.456, -0.001234567, 1e-120,
            1e100, 0.5, 2.5e5, 7.77777e-7, 1e-99, 9.99996e99]
    return np.concatenate([v, edge])


def test_tfile_matches_fortran_e12_5(pkg, tmp_path):
    flang = shutil.which("flang") or shutil.which("gfortran")
    if flang is None:
        pytest.skip("no Fortran compiler on this machine")
    v = _values()
    n = len(v)
    x, y, z, c = v, v[::-1].copy(), np.roll(v, 7), np.roll(v, 13) / 1e6
    (tmp_path / "t.f90").write_text(FORTRAN)
    subprocess.run([flang, "-o", str(tmp_path / "t"), str(tmp_path / "t.f90")], check=True,
                   capture_output=True)
    with open(tmp_path / "in.txt", "w") as f:
        f.write(f"{n}\n")
        for i in range(n):
            f.write("%.17e %.17e %.17e %.17e\n" % (x[i], y[i], z[i], c[i]))
    subprocess.run([str(tmp_path / "t")], cwd=tmp_path, check=True, capture_output=True)
    ref = (tmp_path / "out.txt").read_text().splitlines()
    pkg._lib.write_tfile(str(tmp_path / "mine.txt"), x, y, z, c)
    mine = (tmp_path / "mine.txt").read_text().splitlines()
    assert len(mine) == n
    assert mine == ref
    assert all(len(line) == 54 for line in mine)


def test_tfile_fixed_examples(pkg, tmp_path):
    v = np.array([1.0, -0.001234567, 9.999996, 1e-120, 0.0])
    pkg._lib.write_tfile(str(tmp_path / "t.txt"), v, v, v, v / 1e6)
    lines = (tmp_path / "t.txt").read_text().splitlines()
    assert [ln[:12] for ln in lines] == [" 0.10000E+01", "-0.12346E-02", " 0.10000E+02",
                                        " 0.10000-119", " 0.00000E+00"]


def test_tfile_special_values(pkg, tmp_path):
    """NaN and infinities as gfortran writes them in e12.5 (the reference's
    build compiles VANADIS.FOR with gfortran, makefile:296: "NaN",
    "Infinity", "-Infinity", right-justified; flang would write "Inf");
    subnormals and the largest double with 3-digit exponents."""
    v = np.array([np.nan, np.inf, -np.inf, 5e-324, 1.7976931348623157e308, -1e-310])
    pkg._lib.write_tfile(str(tmp_path / "t.txt"), v, v, v, np.zeros(6))
    lines = (tmp_path / "t.txt").read_text().splitlines()
    assert [ln[:12] for ln in lines] == ["         NaN", "    Infinity", "   -Infinity",
                                        " 0.49407-323", " 0.17977+309", "-0.10000-309"]
