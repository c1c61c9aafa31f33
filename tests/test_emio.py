"""On-disk formats (EM.dat, magnetic_<tag>.dat) against the Fortran compiler's
own output of the reference's format statements (CPU; needs amdflang)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from matjes_amd import emio

FC = shutil.which("amdflang") or (os.path.exists("/opt/rocm/llvm/bin/amdflang") and "/opt/rocm/llvm/bin/amdflang")

# the reference's write statements (spindynamics.f90:136-139,358-360; io-utils.f90:235-236)
PROG = """
program fmt
  implicit none
  real(8) :: v(18), m(3*%(n)d)
  integer :: io, j
  character(len=30) :: rw_format
  open(newunit=io, file='in.bin', access='stream', form='unformatted', action='read')
  read(io) v, m
  close(io)
  open(7, file='EM.dat', action='write', status='replace', form='formatted')
  Write(7,'(19(a,2x))') '# 1:step','2:real_time (fs)','3:E_av (eV)','4:M', &
   &  '5:Mx','6:My','7:Mz','8:vorticity','9:vx', &
   &  '10:vy','11:vz','12:qeuler','13:q+','14:q-','15:T= (K)', &
   &  '16:Tfin=','17:Hx (T)','18:H (T)y=','19:H (T)z='
  j = 123456
  Write(7,'(I12,17(E20.12E3,2x),E20.12E3)') j, v
  close(7)
  open(8, file='cfg.dat', action='write', status='replace', form='formatted')
  write(rw_format,'( "(", I4, "E24.16,2x)" )') 3
  write(8, rw_format) m
  close(8)
  ! print_Beff (print_Beff.f90:18-25) and the generator's print_base (random_base.f90:163-166)
  open(9, file='beff.dat', action='write', status='replace', form='formatted')
  do j = 1, %(n)d
    write(9, '(3(E20.12E3,2x))') m(3*j-2:3*j)
  enddo
  close(9)
  open(10, file='rnd.dat', action='write', status='replace', form='formatted')
  do j = 1, 3*%(n)d
    write(10, '(E20.12E3)') m(j)
  enddo
  close(10)
end program
"""


@pytest.mark.skipif(not FC, reason="no Fortran compiler")
def test_formats_match_fortran(tmp_path):
    rng = np.random.default_rng(3)
    n = 40
    m = rng.normal(size=3 * n) * 10.0 ** rng.integers(-12, 12, size=3 * n)
    m[:6] = [0.0, -0.0, 1.0, -1.0, 0.5, 1e-300]
    m[6:9] = [1e100, -2.5e-120, 0.1]
    v = rng.normal(size=18) * 10.0 ** rng.integers(-5, 5, size=18)
    v[3] = 0.0
    (tmp_path / "fmt.f90").write_text(PROG % {"n": n})
    np.concatenate([v, m]).astype("<f8").tofile(tmp_path / "in.bin")
    subprocess.check_call([FC, "fmt.f90", "-o", "fmt"], cwd=tmp_path)
    subprocess.check_call(["./fmt"], cwd=tmp_path)
    ref_em = (tmp_path / "EM.dat").read_text().splitlines()
    ref_cfg = (tmp_path / "cfg.dat").read_text().splitlines()
    # header line and one data row
    emio.write_em_dat(tmp_path / "ours_em.dat", [])
    ours_head = (tmp_path / "ours_em.dat").read_text().splitlines()[0]
    assert ours_head.rstrip() == ref_em[0].rstrip()
    row = "%12d" % 123456 + "".join(emio.fortran_e(x, 20, 12, 3) + ("  " if k < 17 else "") for k, x in enumerate(v))
    assert row.rstrip() == ref_em[1].rstrip()
    emio.write_config(tmp_path / "ours_cfg.dat", m)
    ours = (tmp_path / "ours_cfg.dat").read_text().splitlines()
    assert [s.rstrip() for s in ours] == [s.rstrip() for s in ref_cfg]
    emio.write_beff(tmp_path / "ours_beff.dat", m)
    assert [x.rstrip() for x in (tmp_path / "ours_beff.dat").read_text().splitlines()] == \
        [x.rstrip() for x in (tmp_path / "beff.dat").read_text().splitlines()]
    emio.write_rnd_num(tmp_path / "ours_rnd.dat", m)
    assert (tmp_path / "ours_rnd.dat").read_text().splitlines() == (tmp_path / "rnd.dat").read_text().splitlines()
    # E24.16 keeps 16 significant digits: the round trip is exact to 1e-15 relative, not bitwise
    np.testing.assert_allclose(emio.read_config(tmp_path / "ours_cfg.dat").reshape(-1), m, rtol=1e-15, atol=0)


def test_em_row_columns():
    r = emio.em_row(7, 6.0, -1.5e-3, [0.0, 0.6, 0.8], T=10.0, H=(0.0, 0.0, 1.0))
    f = r.split()
    assert f[0] == "7" and len(f) == 19
    assert float(f[3]) == 1.0 and float(f[14]) == 10.0 and float(f[18]) == 1.0
