"""ctypes access to the FP64 CPU oracle (oracle/liborc.so) for the tests.

Test infrastructure: the oracle is the checker of the HIP engine.
"""
import ctypes
import os
import subprocess
from ctypes import POINTER, byref, c_double, c_int32, c_int64, c_uint32, c_uint64, c_void_p

import numpy as np

from md_brushes_amd.abi import MdParams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC_PATH = os.path.join(ROOT, "oracle", "liborc.so")

_lib = None
_fast = None


def _bind(L):
    P = POINTER(MdParams)
    dp, ip = POINTER(c_double), POINTER(c_int32)
    sig = {
        "orc_derive": ([P], c_int32),
        "orc_build_topology": ([P, ip, ip, ip], c_int32),
        "orc_create": ([P], c_void_p),
        "orc_destroy": ([c_void_p], None),
        "orc_set_state": ([c_void_p, dp, dp, dp], None),
        "orc_get_state": ([c_void_p, dp, dp, dp, dp, dp], None),
        "orc_set_drmax": ([c_void_p, c_double, c_double], None),
        "orc_lj_verlet_list": ([c_void_p], c_int64),
        "orc_real_verlet_list": ([c_void_p], c_int64),
        "orc_get_list": ([c_void_p, c_int32, ip, c_int64], c_int64),
        "orc_lj_force": ([c_void_p, dp], c_int32),
        "orc_fene_force": ([c_void_p, dp], c_int32),
        "orc_real_space": ([c_void_p, dp], None),
        "orc_spme": ([c_void_p, dp], None),
        "orc_standard_ewald": ([c_void_p, c_double, c_int32, c_int32, c_int32, dp], None),
        "orc_error_analysis": ([c_void_p, dp, dp, dp], None),
        "orc_compute_force": ([c_void_p, c_int32, c_int32, dp], c_int32),
        "orc_new_position": ([c_void_p, c_int32, c_int32, dp], c_int32),
        "orc_update_verlet_list": ([c_void_p], c_int32),
        "orc_rescale_velocity": ([c_void_p], None),
        "orc_run": ([c_void_p, c_int32, c_int32, c_int32], c_int32),
        "orc_energies": ([c_void_p, dp], None),
        "orc_set_threads": ([c_int32], None),
        "orc_fft3d": ([c_int32, c_int32, c_int32, dp, c_int32], None),
        "orc_bspln_coeffs": ([c_int32, dp, dp], None),
        "orc_bspln": ([c_double, c_int32], c_double),
        "orc_splcof": ([c_int32, dp, c_int32, dp, dp, dp, dp], None),
        "orc_pme_bc_axis": ([c_int32, c_int32, c_double, c_double, dp, dp, dp], None),
        "orc_philox4x32_10": ([POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32)], None),
        "orc_noise_normals": ([c_uint64, c_int32, c_int32, dp], None),
    }
    for name, (a, r) in sig.items():
        f = getattr(L, name)
        f.argtypes = a
        f.restype = r
    return L


def lib():
    """The checker: oracle/liborc.so (serial, -O2)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORC_PATH):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    _lib = _bind(ctypes.CDLL(ORC_PATH))
    return _lib


def fast_lib():
    """The same oracle built -O3 with OpenMP (oracle/liborc_fast.so): the CPU
    baseline; orc_set_threads picks its thread count."""
    global _fast
    if _fast is None:
        path = os.path.join(ROOT, "oracle", "liborc_fast.so")
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        _fast = _bind(ctypes.CDLL(path))
    return _fast


def dp(a):
    return None if a is None else a.ctypes.data_as(POINTER(c_double))


def ip(a):
    return None if a is None else a.ctypes.data_as(POINTER(c_int32))


def colmajor(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).T).reshape(-1)


def derive(p):
    q = p.copy()
    rc = lib().orc_derive(byref(q))
    assert rc == 0
    return q


class Oracle:
    """One system in the oracle; arrays are (NN,k).  L: the library (the
    checker by default, fast_lib() for the CPU baseline)."""

    def __init__(self, p, L=None):
        self.p = p.copy()
        self.NN = p.NN
        self.L = L or lib()
        self.h = self.L.orc_create(byref(self.p))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def set_state(self, pos, vel=None, acc=None):
        self.L.orc_set_state(self.h, dp(colmajor(pos)), dp(None if vel is None else colmajor(vel)),
                            dp(None if acc is None else colmajor(acc)))

    def state(self):
        N = self.NN
        pos, vel, acc, accc, dr = (np.zeros(4 * N), np.zeros(3 * N), np.zeros(3 * N), np.zeros(3 * N),
                                   np.zeros(2))
        self.L.orc_get_state(self.h, dp(pos), dp(vel), dp(acc), dp(accc), dp(dr))
        return pos.reshape(4, N).T, vel.reshape(3, N).T, acc.reshape(3, N).T, accc.reshape(3, N).T, dr

    def set_drmax(self, d1, d2):
        self.L.orc_set_drmax(self.h, d1, d2)

    def build_lists(self):
        n1 = self.L.orc_lj_verlet_list(self.h)
        assert n1 >= 0
        if self.p.real_verlet and self.p.qq != 0:
            assert self.L.orc_real_verlet_list(self.h) >= 0

    def lists(self, which):
        n = self.L.orc_get_list(self.h, which, None, 0)
        pairs = np.zeros(2 * max(n, 1), np.int32)
        self.L.orc_get_list(self.h, which, ip(pairs), n)
        return pairs[: 2 * n].reshape(2, n).T.copy()

    def _f(self, fn):
        f = np.zeros(3 * self.NN)
        rc = fn(self.h, dp(f))
        return f.reshape(3, self.NN).T.copy(), rc

    def lj_force(self):
        return self._f(self.L.orc_lj_force)[0]

    def fene_force(self):
        f, rc = self._f(self.L.orc_fene_force)
        return f, rc

    def real_space(self):
        return self._f(self.L.orc_real_space)[0]

    def spme(self):
        return self._f(self.L.orc_spme)[0]

    def coulomb(self):
        return self.spme() + self.real_space()

    def standard_ewald(self, alpha, K):
        f = np.zeros(3 * self.NN)
        self.L.orc_standard_ewald(self.h, alpha, K[0], K[1], K[2], dp(f))
        return f.reshape(3, self.NN).T.copy()

    def error_analysis(self):
        """(summary dict, f_se (NN,3), f_c (NN,3)) of error_analysis."""
        out = np.zeros(8)
        fse, fc = np.zeros(3 * self.NN), np.zeros(3 * self.NN)
        self.L.orc_error_analysis(self.h, dp(out), dp(fse), dp(fc))
        keys = ["rmsf", "f_r", "f_k", "alpha", "rcc", "K1", "K2", "K3"]
        return dict(zip(keys, out)), fse.reshape(3, self.NN).T.copy(), fc.reshape(3, self.NN).T.copy()

    def compute_force(self, step, noise=0, eta=None):
        return self.L.orc_compute_force(self.h, step, noise, dp(None if eta is None else colmajor(eta)))

    def new_position(self, step, noise=0, eta=None):
        return self.L.orc_new_position(self.h, step, noise, dp(None if eta is None else colmajor(eta)))

    def update_verlet_list(self):
        return self.L.orc_update_verlet_list(self.h)

    def rescale_velocity(self):
        self.L.orc_rescale_velocity(self.h)

    def run(self, step0, n, preheat=False):
        return self.L.orc_run(self.h, step0, n, int(preheat))

    ENERGIES = ["lj", "wall", "fene", "real", "recip", "slab", "self", "coulomb", "field", "potential",
                "kinetic", "kinetic_per_particle", "virial_lj", "virial_fene", "virial_real"]

    def energies(self):
        """Energies and pair virials at the current state (dict)."""
        out = np.zeros(16)
        self.L.orc_energies(self.h, dp(out))
        return dict(zip(self.ENERGIES, out))


def philox(ctr, key):
    c = (c_uint32 * 4)(*ctr)
    k = (c_uint32 * 2)(*key)
    o = (c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def rel_rms(a, b):
    """sqrt(sum|a-b|^2 / sum|b|^2): the relative force RMS of error_analysis
    (compt_acc.f90:795-803)."""
    a, b = np.asarray(a), np.asarray(b)
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))
