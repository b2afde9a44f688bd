"""Text bundle dumps (LoadParticleBundleDat / DumpParticleBundleDat,
R/src/pic/particle/ParticleBundle.f90:151-174, 230-262), CPU checks:

* a dump written by a Fortran program (compiled here with flang) with the
  same list-directed WRITE sequence reads back bit for bit, entries past
  NPar included;
* a dump written by ipm2d.restart reads back bit for bit in a Fortran
  program using the same list-directed READ sequence (bit patterns printed
  in hexadecimal);
* the forms other compilers emit -- D exponents, r*c repeat counts, commas,
  three-digit exponents without the letter, .TRUE. -- parse as Fortran
  defines them;
* to_grid_units rescales an unscaled dump as PositionRescaleParticleBundle
  and VelocityRescaleParticleBundle do (multiply by 1/XFactor, 1/VFactor).
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from ipm2d import restart

FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
needs_flang = pytest.mark.skipif(not os.path.exists(FLANG) and shutil.which("flang") is None,
                                 reason="no Fortran compiler")

WRITER = """
program w
  implicit none
  integer(4) :: npar, sz, i
  real(8) :: charge, mass, weight, xf, vf
  logical :: lx, lv
  real(8), allocatable :: a(:, :)
  npar = 5; sz = 7
  charge = -1.6022d-19; mass = 9.1095d-31; weight = 123456.789d0
  xf = 7.8125d-4; vf = 1.5625d7; lx = .true.; lv = .false.
  allocate(a(sz, 8))
  a = 0.d0
  do i = 1, npar
    a(i, 1) = dble(i) + 1.d0 / 3.d0
    a(i, 2) = dble(i) * 1.d-300
    a(i, 3) = -dble(i) * 2.5d7
    a(i, 4) = 1.d0 / dble(i + 6)
    a(i, 5) = dble(i) * 1.d300
    a(i, 6) = -4.9d-324 * dble(i)
    a(i, 7) = dble(i) * 0.1d0
    a(i, 8) = 1.d5 + dble(i)
  end do
  a(6, 1) = 42.d0
  open(10, file='bundle.dat')
  write(10, *) npar, sz
  write(10, *) charge, mass, weight
  write(10, *) xf, vf
  write(10, *) lx, lv
  do i = 1, 8
    write(10, *) a(:, i)
  end do
  close(10)
end program
"""

READER = """
program r
  implicit none
  integer(4) :: npar, sz, i, k
  real(8) :: charge, mass, weight, xf, vf
  logical :: lx, lv
  real(8), allocatable :: a(:, :)
  open(10, file='ours.dat')
  read(10, *) npar, sz
  read(10, *) charge, mass, weight
  read(10, *) xf, vf
  read(10, *) lx, lv
  allocate(a(sz, 8))
  do i = 1, 8
    read(10, *) a(:, i)
  end do
  close(10)
  write(*, '(2I12)') npar, sz
  write(*, '(5Z17.16)') charge, mass, weight, xf, vf
  write(*, '(2L2)') lx, lv
  do i = 1, 8
    do k = 1, sz
      write(*, '(Z17.16)') a(k, i)
    end do
  end do
end program
"""


def _compile(tmp_path, name, src):
    f = tmp_path / f"{name}.f90"
    f.write_text(src)
    exe = tmp_path / name
    cc = FLANG if os.path.exists(FLANG) else shutil.which("flang")
    subprocess.run([cc, "-O0", str(f), "-o", str(exe)], check=True, cwd=tmp_path, capture_output=True)
    return exe


def _bits(x):
    return struct.unpack("<Q", struct.pack("<d", float(x)))[0]


@needs_flang
def test_reads_a_fortran_written_dump_bit_for_bit(tmp_path):
    exe = _compile(tmp_path, "w", WRITER)
    subprocess.run([str(exe)], check=True, cwd=tmp_path)
    d = restart.read_dat(tmp_path / "bundle.dat")
    assert (d.npar, d.size) == (5, 7)
    assert (d.lxscaled, d.lvscaled) == (True, False)
    assert d.charge == -1.6022e-19 and d.mass == 9.1095e-31 and d.weight == 123456.789
    assert d.xfactor == 7.8125e-4 and d.vfactor == 1.5625e7
    i = np.arange(1, 6, dtype=np.float64)
    want = np.zeros((8, 7))
    want[0, :5] = i + 1.0 / 3.0
    want[1, :5] = i * 1e-300
    want[2, :5] = -i * 2.5e7
    want[3, :5] = 1.0 / (i + 6)
    want[4, :5] = i * 1e300
    want[5, :5] = -4.9e-324 * i
    want[6, :5] = i * 0.1
    want[7, :5] = 1e5 + i
    want[0, 5] = 42.0
    assert [_bits(x) for x in d.po.ravel()] == [_bits(x) for x in want.ravel()]
    assert d.particles().shape == (8, 5)


@needs_flang
def test_fortran_reads_our_dump_bit_for_bit(tmp_path, rng):
    P = rng.standard_normal((8, 9)) * np.logspace(-200, 200, 9)
    P[5, 3] = 5e-324
    P[6, 0] = -0.0
    d = restart.BundleDump(6, 9, 1.6022e-19, 2.2e-25, 3.5e9, 1e-3, 2e7, True, True, P)
    restart.write_dat(tmp_path / "ours.dat", d, per_line=3)
    exe = _compile(tmp_path, "r", READER)
    out = subprocess.run([str(exe)], check=True, cwd=tmp_path, capture_output=True, text=True).stdout.split()
    assert out[:2] == ["6", "9"]
    head = [int(v, 16) for v in out[2:7]]
    assert head == [_bits(x) for x in (1.6022e-19, 2.2e-25, 3.5e9, 1e-3, 2e7)]
    assert out[7:9] == ["T", "T"]
    vals = [int(v, 16) for v in out[9:]]
    assert vals == [_bits(x) for x in P.ravel()]


def test_other_compilers_forms(tmp_path):
    text = """ 3 4
 -1.6022D-19, 9.1095E-31 , 1.0Q+05
 1.0d-3 2.0D+07 extra-ignored
 .TRUE. .false.
 1.5 2*2.5D0
 7.0
 4*0.
 1.0-300 2.0+300 3.0 4.0
 4*1
 1 2 3 4
 4*-1.
 0.5, 0.25, 0.125, 0.0625
 9 8 7 6
"""
    p = tmp_path / "x.dat"
    p.write_text(text)
    d = restart.read_dat(p)
    assert (d.npar, d.size) == (3, 4)
    assert d.charge == -1.6022e-19 and d.mass == 9.1095e-31 and d.weight == 1e5
    assert (d.xfactor, d.vfactor) == (1e-3, 2e7)
    assert (d.lxscaled, d.lvscaled) == (True, False)
    np.testing.assert_array_equal(d.po[0], [1.5, 2.5, 2.5, 7.0])
    np.testing.assert_array_equal(d.po[1], [0.0] * 4)
    np.testing.assert_array_equal(d.po[2], [1e-300, 2e300, 3.0, 4.0])
    np.testing.assert_array_equal(d.po[3], [1.0] * 4)
    np.testing.assert_array_equal(d.po[5], [-1.0] * 4)
    np.testing.assert_array_equal(d.po[6], [0.5, 0.25, 0.125, 0.0625])
    np.testing.assert_array_equal(d.po[7], [9.0, 8.0, 7.0, 6.0])


def test_round_trip_and_rescale(tmp_path, rng):
    P = rng.standard_normal((8, 50))
    d = restart.BundleDump(40, 50, -1.0, 2.0, 3.0, 0.25, 8.0, False, False, P)
    restart.write_dat(tmp_path / "a.dat", d)
    e = restart.read_dat(tmp_path / "a.dat")
    assert [_bits(x) for x in e.po.ravel()] == [_bits(x) for x in P.ravel()]
    g = e.to_grid_units()
    assert g.lxscaled and g.lvscaled
    np.testing.assert_array_equal(g.po[0:2, :40], P[0:2, :40] * (1.0 / 0.25))
    np.testing.assert_array_equal(g.po[2:5, :40], P[2:5, :40] * (1.0 / 8.0))
    np.testing.assert_array_equal(g.po[5:8], P[5:8])            # Az, Ar, WQ untouched
    np.testing.assert_array_equal(g.po[:, 40:], P[:, 40:])       # past NPar untouched


def test_malformed_dumps_rejected(tmp_path):
    p = tmp_path / "bad.dat"
    p.write_text(" 5 3\n 1. 2. 3.\n 1. 2.\n T T\n")
    with pytest.raises(ValueError):
        restart.read_dat(p)
    p.write_text(" 1 2\n 1. 2. 3.\n 1. 2.\n T T\n 1. 2.\n")
    with pytest.raises(ValueError):
        restart.read_dat(p)
