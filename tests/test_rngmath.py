"""Accuracy of the thermal generator's ln and sincos (matjes_amd/csrc/mj_rngmath.h)
against libm, on the host (the same source compiles for the device)."""
import ctypes
import math
import os
import subprocess

import numpy as np

from conftest import REPO

SRC = r'''
#include <stdint.h>
#include "mj_rngmath.h"
extern "C" void logs(const double *u, double *o, long n) { for (long i = 0; i < n; ++i) o[i] = mj_log01(u[i]); }
extern "C" void sc(const double *u, double *s, double *c, long n) { for (long i = 0; i < n; ++i) mj_sincos2pi(u[i], s + i, c + i); }
extern "C" void logw(const uint32_t *w, double *o, long n) { for (long i = 0; i < n; ++i) o[i] = mj_ln_u32(w[i]); }
extern "C" void scw(const uint32_t *w, double *s, double *c, double *c2, long n)
{ for (long i = 0; i < n; ++i) { mj_sincos2pi_u32(w[i], s + i, c + i); c2[i] = mj_cos2pi_u32(w[i]); } }
'''


def _lib(tmp_path):
    cpp = tmp_path / "w.cpp"
    cpp.write_text(SRC)
    so = tmp_path / "w.so"
    subprocess.check_call(["g++", "-O2", "-fPIC", "-shared", "-I", os.path.join(REPO, "matjes_amd", "csrc"),
                           str(cpp), "-o", str(so)])
    L = ctypes.CDLL(str(so))
    dp = ctypes.POINTER(ctypes.c_double)
    L.logs.argtypes = [dp, dp, ctypes.c_long]
    L.sc.argtypes = [dp, dp, dp, ctypes.c_long]
    up = ctypes.POINTER(ctypes.c_uint32)
    L.logw.argtypes = [up, dp, ctypes.c_long]
    L.scw.argtypes = [up, dp, dp, dp, ctypes.c_long]
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _uniform53(n, seed):
    k = np.random.default_rng(seed).integers(0, 2 ** 53, size=n, dtype=np.int64)
    return k.astype(np.float64) * 2.0 ** -53


def test_log01_accuracy(tmp_path):
    L = _lib(tmp_path)
    u = _uniform53(400000, 1) + 2.0 ** -53          # (0, 1]
    u = np.concatenate([u, [2.0 ** -53, 0.5, 1.0 - 2.0 ** -53, 1.0, 0.70710678118654752440]])
    o = np.empty_like(u)
    L.logs(_p(u), _p(o), len(u))
    ref = np.log(u)
    ulp = np.spacing(np.abs(ref)) + 1e-300
    err = np.abs(o - ref) / ulp
    assert np.max(err[ref != 0]) <= 4.0
    assert o[-2] == 0.0


def test_sincos2pi_accuracy(tmp_path):
    L = _lib(tmp_path)
    u = np.concatenate([_uniform53(400000, 2), np.arange(8) / 8.0, [0.0625, 0.1875, 1.0 - 2.0 ** -53]])
    s = np.empty_like(u)
    c = np.empty_like(u)
    L.sc(_p(u), _p(s), _p(c), len(u))
    # reference in long double via mpmath-free exact path: use numpy longdouble sin/cos of 2*pi*u
    a = 2 * np.pi * u.astype(np.longdouble)
    rs, rc = np.sin(a), np.cos(a)
    assert np.max(np.abs(s - rs)) <= 4e-16
    assert np.max(np.abs(c - rc)) <= 4e-16
    assert np.max(np.abs(s * s + c * c - 1.0)) <= 4e-16


def _words(n, seed):
    w = np.random.default_rng(seed).integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
    edge = np.array([0, 1, 2, 2 ** 29 - 1, 2 ** 29, 2 ** 30, 2 ** 31, 3 * 2 ** 30, 2 ** 32 - 2, 2 ** 32 - 1],
                    dtype=np.uint32)
    return np.ascontiguousarray(np.concatenate([w, edge]))


def test_word_forms_of_the_thermal_normals(tmp_path):
    """ln((w + 1) 2^-32), (sin, cos)(2 pi w 2^-32) and cos alone from the
    32-bit word (the forms the thermal normals use), host build (IEEE quotient
    and square root; the device's hardware-seeded ones are checked through
    the normals against the oracle on the GPU)."""
    L = _lib(tmp_path)
    w = _words(400000, 5)
    up = ctypes.POINTER(ctypes.c_uint32)
    o = np.empty(len(w))
    L.logw(w.ctypes.data_as(up), _p(o), len(w))
    u = (w.astype(np.float64) + 1.0) * 2.0 ** -32
    ref = np.log(u)
    ok = ref != 0
    assert np.max(np.abs(o[ok] - ref[ok]) / np.spacing(np.abs(ref[ok]))) <= 4.0
    assert o[-1] == 0.0
    s, c, c2 = np.empty(len(w)), np.empty(len(w)), np.empty(len(w))
    L.scw(w.ctypes.data_as(up), _p(s), _p(c), _p(c2), len(w))
    a = 2 * np.pi * (w.astype(np.longdouble) * np.longdouble(2.0) ** -32)
    assert np.max(np.abs(s - np.sin(a))) <= 4e-16
    assert np.max(np.abs(c - np.cos(a))) <= 4e-16
    assert np.max(np.abs(c2 - np.cos(a))) <= 4e-16
