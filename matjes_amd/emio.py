"""The reference's on-disk formats for the dynamics outputs, written from
device snapshots.

  EM.dat                 header (spindynamics.f90:135-139), one row per Efreq
                         steps with '(I12,17(E20.12E3,2x),E20.12E3)'
                         (spindynamics.f90:358-360)
  magnetic_<tag>.dat     write_config -> dump_config(io, dim_mode, modes)
                         with '(   NE24.16,2x)', N = dim_mode values per line,
                         x fastest (IO-outputs/write_config.f90:62-82,
                         IO-utils/io-utils.f90:227-237)
  Beff_<tag>.dat         print_Beff: the effective field of the step's last
                         stage, '(N(E20.12E3,2x))' per site (IO-outputs/
                         print_Beff.f90:10-29, spindynamics.f90:400)
  rnd_num_<tag>          random_numbers%print_base: the step's thermal draws
                         (sigma-scaled, 3N), one '(E20.12E3)' per line
                         (random-numbers/random_base.f90:153-170,
                         spindynamics.f90:424)

Numbers are formatted as Fortran's Ew.d[Ee] edit descriptor does (mantissa
0.d1d2..., correctly rounded, exponent sign always written), so files diff
cleanly against the reference's.
"""
import math
import os
import re

import numpy as np

# columns 8-14 when the topological charge is not computed (spindynamics.f90:344-346)
_NO_TOPO = 1.2345789e+98 * math.pi * 4.0
_NO_VORT = 1.2345789e+98 * 3.0 * math.sqrt(3.0)

EM_HEADER = ["# 1:step", "2:real_time (fs)", "3:E_av (eV)", "4:M", "5:Mx", "6:My", "7:Mz", "8:vorticity", "9:vx",
             "10:vy", "11:vz", "12:qeuler", "13:q+", "14:q-", "15:T= (K)", "16:Tfin=", "17:Hx (T)", "18:H (T)y=",
             "19:H (T)z="]


def fortran_e(x, w, d, e=None):
    """x formatted with the Fortran edit descriptor Ew.d (Ee if e is given)."""
    x = float(x)
    if math.isnan(x) or math.isinf(x):
        s = ("-" if x < 0 else "") + ("NaN" if math.isnan(x) else "Infinity")
        return s.rjust(w)
    neg = x < 0 or (x == 0.0 and math.copysign(1.0, x) < 0)
    if x == 0.0:
        digits, exp = "0" * d, 0
    else:
        m, ex = ("%.*e" % (d - 1, abs(x))).split("e")
        digits = m.replace(".", "")
        exp = int(ex) + 1
    if e is None:
        ea = abs(exp)
        es = ("E%+03d" % exp) if ea <= 99 else ("%+04d" % exp)
    else:
        es = "E" + ("+" if exp >= 0 else "-") + str(abs(exp)).rjust(e, "0")
    s = ("-" if neg else "") + "0." + digits + es
    if len(s) > w:
        s = s.replace("-0.", "-.", 1) if neg else s.replace("0.", ".", 1)
    if len(s) > w:
        return "*" * w
    return s.rjust(w)


def em_row(step, real_time, e_av, M, vortex=None, q=None, T=0.0, T_measured=0.0, H=(0.0, 0.0, 0.0)):
    """One EM.dat row: step, time, E/Ncell, |M|, M, |vortex|, vortex, q_euler,
    q+, q-, T, T_measured, H_ext (the reference's placeholders when the
    topology is not computed)."""
    M = np.asarray(M, dtype=np.float64)
    if vortex is None:
        v = np.full(3, _NO_VORT / 3.0 / math.sqrt(3.0))
    else:
        v = np.asarray(vortex, dtype=np.float64)
    if q is None:
        qp = qm = _NO_TOPO / math.pi / 4.0
    else:
        qp, qm = q
    vals = [real_time, e_av, math.sqrt(float(np.sum(M * M))), *M.tolist(), math.sqrt(float(np.sum(v * v))),
            *v.tolist(), qp + qm, qp, qm, T, T_measured, *list(H)]
    out = "%12d" % step
    for k, x in enumerate(vals):
        out += fortran_e(x, 20, 12, 3)
        if k < len(vals) - 1:
            out += "  "
    return out


def write_em_dat(path, rows):
    """rows: dicts or tuples for em_row (step, real_time, e_av, M, ...)."""
    with open(path, "w") as f:
        f.write("".join(h + "  " for h in EM_HEADER).rstrip(" ") + "  \n")
        for r in rows:
            f.write((em_row(**r) if isinstance(r, dict) else em_row(*r)) + "\n")


def write_config(path, m, dim_mode=3):
    """magnetic_<tag>.dat: dim_mode values per line in E24.16, then 2 blanks."""
    m = np.asarray(m, dtype=np.float64).reshape(-1, dim_mode)
    with open(path, "w") as f:
        for row in m:
            f.write("".join(fortran_e(x, 24, 16) for x in row) + "  \n")


def config_name(tag, directory="."):
    return os.path.join(directory, "magnetic_%s.dat" % str(tag).strip())


def write_beff(path, B, dim_mode=3):
    """Beff_<tag>.dat: dim_mode values per line, each E20.12E3 then 2 blanks."""
    B = np.asarray(B, dtype=np.float64).reshape(-1, dim_mode)
    with open(path, "w") as f:
        for row in B:
            f.write("".join(fortran_e(x, 20, 12, 3) + "  " for x in row) + "\n")


def beff_name(tag, directory="."):
    return os.path.join(directory, "Beff_%s.dat" % str(tag).strip())


def write_rnd_num(path, x):
    """rnd_num_<tag>: the 3N thermal draws, one E20.12E3 per line."""
    with open(path, "w") as f:
        for v in np.asarray(x, dtype=np.float64).reshape(-1):
            f.write(fortran_e(v, 20, 12, 3) + "\n")


def rnd_num_name(tag, directory="."):
    return os.path.join(directory, "rnd_num_%s" % str(tag).strip())


_EXP3 = re.compile(r"(?<=\d)([+-]\d{3})\b")


def read_config(path, dim_mode=3):
    """Read a magnetic_<tag>.dat (also the Fortran form 0.1000-299 of 3-digit exponents)."""
    with open(path) as f:
        vals = [float(_EXP3.sub(r"E\1", t)) for line in f for t in line.split()]
    return np.asarray(vals, dtype=np.float64).reshape(-1, dim_mode)
