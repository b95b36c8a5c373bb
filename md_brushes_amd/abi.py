"""ctypes binding of the engine's C-ABI (include/md_engine.h).

The parameter struct mirrors ``md_params`` field for field; the Engine class
mirrors the reference's hot-path subroutines (compute_acceleration and
initialize_update modules of MD_Brushes) with explicit arguments.
"""
import ctypes
import os
from ctypes import POINTER, byref, c_char_p, c_double, c_int32, c_int64, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmd_engine.so")

MD_OK, MD_EINVAL, MD_EBOND, MD_ECELL, MD_ENONFINITE, MD_EOVERFLOW, MD_EHIP, MD_ENOMEM, MD_EIO = range(9)
MD_FP64, MD_FP32 = 0, 1
MD_NOISE_NONE, MD_NOISE_ETA, MD_NOISE_PHILOX = 0, 1, 2
MD_MAX_ARM = 16

_D = c_double
_I = c_int32


class MdParams(ctypes.Structure):
    """Field-for-field image of ``md_params``."""

    _fields_ = [
        ("uniform_or_random", _I), ("arm", _I), ("Nma", _I), ("Nga", _I), ("Nml", _I),
        ("Ngl", _I), ("man_l", _I), ("man_s", _I),
        ("StepNum0", _I), ("StepNum", _I), ("DeltaStep1", _I), ("DeltaStep2", _I),
        ("DeltaStep3", _I), ("multistep", _I),
        ("Lz", _D), ("Z_empty", _D), ("sigmag", _D), ("Beta", _D), ("qq", _D), ("qqi", _D),
        ("ion_ratio", _D), ("R_bond", _D), ("dt", _D),
        ("epsilon", _D), ("sigma", _D), ("sigma_s", _D), ("rcl", _D), ("rvl", _D), ("rsk", _D),
        ("R0_2", _D), ("kFENE", _D),
        ("ordr", _I * 3), ("pad0", _I),
        ("lb", _D), ("xi", _D), ("EF", _D), ("tol", _D), ("tau_rf", _D),
        ("charged_arm", _I * MD_MAX_ARM),
        ("ovr_ng", _I * 3), ("pad1", _I),
        ("ovr_alpha", _D), ("ovr_rcc", _D),
        ("Nta", _I), ("Ntl", _I), ("Npe", _I), ("N_anchor", _I), ("Nq_PE", _I),
        ("Nq_salt_ions", _I), ("Nq", _I), ("NN", _I), ("N_bond", _I),
        ("real_verlet", _I), ("ng", _I * 3), ("pad2", _I),
        ("Lx", _D), ("Ly", _D), ("Zfac", _D), ("sigma_sp", _D), ("rcl_s", _D), ("rcl_sp", _D),
        ("rvl_eff", _D), ("alpha", _D), ("rcc", _D), ("rvc", _D), ("gdim", _D * 3),
        ("precision", _I), ("pad3", _I), ("seed", c_uint64),
        ("ovr_skin", _D), ("ovr_cell_scale", _D), ("ovr_tile", _I * 3), ("pad4", _I),
    ]

    def copy(self):
        out = MdParams()
        ctypes.memmove(byref(out), byref(self), ctypes.sizeof(MdParams))
        return out

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if hasattr(v, "__len__") else v
        return d


class MDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"md status {code}: {msg}")
        self.code = code


_lib = None


def load_library(path=None):
    """Loads libmd_engine.so; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with __graft_entry__.build() or make -C md_brushes_amd")
    lib = ctypes.CDLL(p)
    P = POINTER(MdParams)
    dp = POINTER(c_double)
    ip = POINTER(c_int32)
    sig = {
        "md_read_params": ([c_char_p, c_char_p, c_char_p, P], _I),
        "md_derive_params": ([P], _I),
        "md_build_topology": ([P, ip, ip, ip], _I),
        "md_initialize_position": ([P, c_uint64, dp], _I),
        "md_initialize_velocity": ([P, c_uint64, dp], _I),
        "md_create": ([POINTER(c_void_p), P], _I),
        "md_upload_state": ([c_void_p, dp, dp, dp], _I),
        "md_download_state": ([c_void_p, dp, dp, dp, dp], _I),
        "md_download_coulomb": ([c_void_p, dp], _I),
        "md_build_lists": ([c_void_p], _I),
        "md_download_list": ([c_void_p, _I, ip, c_int64, POINTER(c_int64)], _I),
        "md_compute_force": ([c_void_p, _I, _I, dp], _I),
        "md_new_position": ([c_void_p, _I, _I, dp], _I),
        "md_update_verlet_list": ([c_void_p], _I),
        "md_rescale_velocity": ([c_void_p], _I),
        "md_run": ([c_void_p, _I, _I, _I], _I),
        "md_force_parts": ([c_void_p, dp, dp], _I),
        "md_last_error": ([c_void_p, c_char_p, _I], _I),
        "md_time_kernels": ([c_void_p, _I, dp, _I], _I),
        "md_autotune": ([c_void_p, _I, ip], _I),
        "md_error_analysis": ([c_void_p, dp, dp], _I),
        "md_heights": ([c_void_p, dp], _I),
        "md_energies": ([c_void_p, dp], _I),
        "md_batch_energies": ([c_void_p, _I, dp], _I),
        "md_state_size": ([c_void_p, POINTER(c_int64)], _I),
        "md_state_save": ([c_void_p, c_void_p, c_int64], _I),
        "md_state_load": ([c_void_p, c_void_p, c_int64], _I),
        "md_batch_heights": ([c_void_p, _I, dp], _I),
        "md_batch_error_analysis": ([c_void_p, _I, dp, dp], _I),
        "md_batch_error_analysis_noalloc": ([c_void_p, _I, dp, dp], _I),
        "md_destroy": ([c_void_p], None),
        "md_batch_create": ([POINTER(c_void_p), _I, _I], _I),
        "md_batch_add": ([c_void_p, P, dp, dp, ip], _I),
        "md_batch_remove": ([c_void_p, _I], _I),
        "md_batch_run": ([c_void_p, _I], _I),
        "md_batch_status": ([c_void_p, _I, ip, c_char_p, _I], _I),
        "md_batch_download": ([c_void_p, _I, dp, dp, dp], _I),
        "md_batch_info": ([c_void_p, _I, POINTER(c_int64), ip], _I),
        "md_replica_bytes": ([P, dp, POINTER(c_int64)], _I),
        "md_autotune_geometry": ([P, dp, dp, _I, _I, P, dp], _I),
        "md_batch_observe": ([c_void_p, _I, dp, dp, _I], _I),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    if path is None:
        _lib = lib
    return lib


def _ewald_summary(out):
    keys = ["rmsf", "f_r", "f_k", "alpha", "rcc", "K1", "K2", "K3"]
    return {k: (int(v) if k[0] == "K" else float(v)) for k, v in zip(keys, out)}


def _dp(a):
    return None if a is None else a.ctypes.data_as(POINTER(c_double))


def _ip(a):
    return None if a is None else a.ctypes.data_as(POINTER(c_int32))


def derive(p):
    """md_derive_params on a copy of p."""
    q = p.copy()
    rc = load_library().md_derive_params(byref(q))
    if rc:
        raise MDError(rc, "md_derive_params")
    return q


def build_topology(p):
    """(charge, fene(N_bond,2), anchor), 1-based as in the reference."""
    lib = load_library()
    charge = np.zeros(max(p.Nq, 1), np.int32)
    fene = np.zeros(2 * max(p.N_bond, 1), np.int32)
    anchor = np.zeros(max(p.N_anchor, 1), np.int32)
    rc = lib.md_build_topology(byref(p), _ip(charge) if p.qq != 0 else None, _ip(fene), _ip(anchor))
    if rc:
        raise MDError(rc, "md_build_topology")
    nb = p.N_bond
    return charge[: p.Nq], fene[: 2 * nb].reshape(2, nb).T.copy(), anchor[: p.N_anchor]


def replica_bytes(p, pos):
    """Device bytes md_batch_add would allocate for this replica (no device call)."""
    b = c_int64()
    rc = load_library().md_replica_bytes(byref(p), _dp(_colmajor(pos)), byref(b))
    if rc:
        raise MDError(rc, "md_replica_bytes")
    return b.value


MD_TUNE_LISTS = 1


def autotune_geometry(p, pos, vel, cycles=50, lists=False):
    """Engine geometry chosen by running the system (md_autotune_geometry):
    spreading tile edges (results unchanged) and, with lists=True, the LJ
    Verlet skin and the real-space sweep's cell scale (forces change by
    rounding).  Returns (derived params with the chosen overrides, steps/s)."""
    best = MdParams()
    sps = c_double()
    rc = load_library().md_autotune_geometry(byref(p), _dp(_colmajor(pos)), _dp(_colmajor(vel)), cycles,
                                             MD_TUNE_LISTS if lists else 0, byref(best), byref(sps))
    if rc:
        raise MDError(rc, "md_autotune_geometry")
    return best, sps.value


def initialize_position(p, seed):
    """pos(NN,4) as a Fortran-order array (x,y,z,q columns)."""
    pos = np.zeros(4 * p.NN)
    rc = load_library().md_initialize_position(byref(p), seed, _dp(pos))
    if rc:
        raise MDError(rc, "md_initialize_position")
    return pos.reshape(4, p.NN).T.copy()


def initialize_velocity(p, seed):
    vel = np.zeros(3 * p.NN)
    rc = load_library().md_initialize_velocity(byref(p), seed, _dp(vel))
    if rc:
        raise MDError(rc, "md_initialize_velocity")
    return vel.reshape(3, p.NN).T.copy()


def _colmajor(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).T).reshape(-1)


class Engine:
    """One system on the GPU; methods follow the reference's subroutines.

    Arrays in and out are (NN,k) numpy arrays (pos: x,y,z,q columns).
    """

    def __init__(self, params):
        self.lib = load_library()
        self.p = params.copy()
        self.NN = params.NN
        self.h = c_void_p()
        rc = self.lib.md_create(byref(self.h), byref(self.p))
        self._check(rc, "md_create")

    def _check(self, rc, what):
        if rc:
            buf = ctypes.create_string_buffer(512)
            if self.h:
                self.lib.md_last_error(self.h, buf, 512)
            raise MDError(rc, f"{what}: {buf.value.decode(errors='replace')}")

    def close(self):
        if self.h:
            self.lib.md_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, pos, vel, acc=None):
        a = None if acc is None else _colmajor(acc)
        self._check(self.lib.md_upload_state(self.h, _dp(_colmajor(pos)), _dp(_colmajor(vel)), _dp(a)),
                    "md_upload_state")

    def download(self):
        N = self.NN
        pos, vel, acc, dr = np.zeros(4 * N), np.zeros(3 * N), np.zeros(3 * N), np.zeros(2)
        self._check(self.lib.md_download_state(self.h, _dp(pos), _dp(vel), _dp(acc), _dp(dr)),
                    "md_download_state")
        return pos.reshape(4, N).T, vel.reshape(3, N).T, acc.reshape(3, N).T, dr

    def coulomb(self):
        c = np.zeros(3 * self.NN)
        self._check(self.lib.md_download_coulomb(self.h, _dp(c)), "md_download_coulomb")
        return c.reshape(3, self.NN).T

    def build_lists(self):
        """lj_verlet_list + real_verlet_list."""
        self._check(self.lib.md_build_lists(self.h), "md_build_lists")

    def lists(self, which):
        """Sorted (npair,2) array of 1-based (i,j) pairs."""
        n = c_int64()
        self._check(self.lib.md_download_list(self.h, which, None, 0, byref(n)), "md_download_list")
        cap = n.value
        pairs = np.zeros(2 * max(cap, 1), np.int32)
        self._check(self.lib.md_download_list(self.h, which, _ip(pairs), cap, byref(n)), "md_download_list")
        return pairs[: 2 * cap].reshape(2, cap).T.copy()

    def compute_force(self, step, noise=MD_NOISE_NONE, eta=None):
        e = None if eta is None else _colmajor(eta)
        self._check(self.lib.md_compute_force(self.h, step, noise, _dp(e)), "md_compute_force")

    def new_position(self, step, noise=MD_NOISE_NONE, eta=None):
        e = None if eta is None else _colmajor(eta)
        self._check(self.lib.md_new_position(self.h, step, noise, _dp(e)), "md_new_position")

    def update_verlet_list(self):
        self._check(self.lib.md_update_verlet_list(self.h), "md_update_verlet_list")

    def rescale_velocity(self):
        self._check(self.lib.md_rescale_velocity(self.h), "md_rescale_velocity")

    def run(self, step0, nsteps, preheat=False):
        self._check(self.lib.md_run(self.h, step0, nsteps, int(preheat)), "md_run")

    def autotune(self, iters=5):
        """Times the launch forms (fused FFT planes, z-pass lines per
        workgroup, real-space sweep form) and keeps the fastest; results do
        not depend on them.  Returns the choice."""
        ch = (c_int32 * 9)()
        self._check(self.lib.md_autotune(self.h, iters, ch), "md_autotune")
        return {"fused_planes": ch[0], "z_lines": ch[1], "real_workgroup": ch[2], "real_shared_launch": ch[3],
                "real_grid": ch[4], "plane_spreading": ch[5], "ride_parts": ch[6], "ride_lead": ch[7],
                "ride_lead_gather": ch[8]}

    KERNELS = ["kick_drift", "short_range", "bin_q", "sort_q", "spread", "fft_x_fwd", "fft_y_fwd",
               "fft_z_conv", "fft_y_inv", "fft_x_inv", "gather", "real", "bin_lj", "sort_lj", "build_lj"]
    COULOMB = KERNELS[2:12]
    LISTS = KERNELS[12:]
    RUN_START = KERNELS[:1]  # once per run: replayed steps drift in short_range

    def time_kernels(self, iters=20):
        """Mean ms per launch of each kernel (HIP events); positions, velocities
        and forces are restored, lists and the held Coulomb force recomputed."""
        ms = np.zeros(15)
        self._check(self.lib.md_time_kernels(self.h, iters, _dp(ms), 15), "md_time_kernels")
        return dict(zip(self.KERNELS, ms))

    def force_parts(self):
        N = self.NN
        fs, fc = np.zeros(3 * N), np.zeros(3 * N)
        self._check(self.lib.md_force_parts(self.h, _dp(fs), _dp(fc)), "md_force_parts")
        return fs.reshape(3, N).T, fc.reshape(3, N).T

    def save_state(self):
        """The engine's complete state (bytes) for an exact resume."""
        n = c_int64()
        self._check(self.lib.md_state_size(self.h, byref(n)), "md_state_size")
        buf = np.zeros(n.value, dtype=np.uint8)
        self._check(self.lib.md_state_save(self.h, buf.ctypes.data_as(c_void_p), n.value), "md_state_save")
        return buf.tobytes()

    def load_state(self, state):
        """Restores save_state's bytes: the run continues bit for bit."""
        buf = np.frombuffer(state, dtype=np.uint8)
        self._check(self.lib.md_state_load(self.h, buf.ctypes.data_as(c_void_p), buf.size), "md_state_load")

    HEIGHTS = ["h_avg", "hl_max", "hl_end", "hl_avg", "hs_max", "hs_end", "hs_branch", "hs_avg"]

    def heights(self):
        """The height analysis of write_height (in_output.f90:1017-1079)."""
        out = np.zeros(8)
        self._check(self.lib.md_heights(self.h, _dp(out)), "md_heights")
        return dict(zip(self.HEIGHTS, out))

    ENERGIES = ["lj", "wall", "fene", "real", "recip", "slab", "self", "coulomb", "field", "potential",
                "kinetic", "kinetic_per_particle", "virial_lj", "virial_fene", "virial_real"]

    def energies(self):
        """Energies and pair virials at the current state (md_energies), FP64
        sums: WCA, wall, FENE, real space, reciprocal, slab, Ewald self,
        Coulomb, field, potential, kinetic, kinetic per mobile particle (the
        reference's kenetic_energy, in_output.f90:1307-1311), pair virials."""
        out = np.zeros(16)
        self._check(self.lib.md_energies(self.h, _dp(out)), "md_energies")
        return dict(zip(self.ENERGIES, out))

    def error_analysis(self):
        """error_analysis (compt_acc.f90:729-812) on the device: (summary,
        standard-Ewald reference Coulomb force (NN,3))."""
        out = np.zeros(8)
        f = np.zeros(3 * self.NN)
        self._check(self.lib.md_error_analysis(self.h, _dp(out), _dp(f)), "md_error_analysis")
        return _ewald_summary(out), f.reshape(3, self.NN).T.copy()


class Batch:
    """Many replicas of uneven size in one engine (md_batch_* entry points)."""

    HIST_BINS = 500

    def __init__(self, precision, multistep=10):
        self.lib = load_library()
        self.h = c_void_p()
        self.sizes = {}
        self.multistep = multistep
        rc = self.lib.md_batch_create(byref(self.h), precision, multistep)
        self._check(rc, "md_batch_create")

    _check = Engine._check

    def close(self):
        if self.h:
            self.lib.md_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self, iters=5):
        """As Engine.autotune, over all replicas of the batch."""
        ch = (c_int32 * 9)()
        self._check(self.lib.md_autotune(self.h, iters, ch), "md_autotune")
        return {"fused_planes": ch[0], "z_lines": ch[1], "real_workgroup": ch[2], "real_shared_launch": ch[3],
                "real_grid": ch[4], "plane_spreading": ch[5], "ride_parts": ch[6], "ride_lead": ch[7],
                "ride_lead_gather": ch[8]}

    def time_kernels(self, iters=20):
        """Mean ms per launch of each kernel over all replicas (HIP events);
        every replica's positions, velocities and forces are restored."""
        ms = np.zeros(15)
        self._check(self.lib.md_time_kernels(self.h, iters, _dp(ms), 15), "md_time_kernels")
        return dict(zip(Engine.KERNELS, ms))

    def add(self, params, pos, vel):
        slot = c_int32()
        self._check(self.lib.md_batch_add(self.h, byref(params), _dp(_colmajor(pos)), _dp(_colmajor(vel)),
                                          byref(slot)), "md_batch_add")
        self.sizes[slot.value] = params.NN
        return slot.value

    def remove(self, slot):
        self._check(self.lib.md_batch_remove(self.h, slot), "md_batch_remove")
        self.sizes.pop(slot, None)

    def run(self, ncycles):
        self._check(self.lib.md_batch_run(self.h, ncycles), "md_batch_run")

    def status(self, slot):
        """(status code, next step, message)."""
        step = c_int32()
        buf = ctypes.create_string_buffer(256)
        rc = self.lib.md_batch_status(self.h, slot, byref(step), buf, 256)
        return rc, step.value, buf.value.decode(errors="replace")

    def download(self, slot):
        N = self.sizes[slot]
        pos, vel, acc = np.zeros(4 * N), np.zeros(3 * N), np.zeros(3 * N)
        self._check(self.lib.md_batch_download(self.h, slot, _dp(pos), _dp(vel), _dp(acc)), "md_batch_download")
        return pos.reshape(4, N).T, vel.reshape(3, N).T, acc.reshape(3, N).T

    def info(self, slot=-1):
        b = c_int64()
        n = c_int32()
        self._check(self.lib.md_batch_info(self.h, slot, byref(b), byref(n)), "md_batch_info")
        return b.value, n.value

    def heights(self, slot):
        """Engine.heights of one replica."""
        out = np.zeros(8)
        self._check(self.lib.md_batch_heights(self.h, slot, _dp(out)), "md_batch_heights")
        return dict(zip(Engine.HEIGHTS, out))

    def error_analysis(self, slot):
        """Accuracy spot check of one replica: (summary, reference force)."""
        out = np.zeros(8)
        f = np.zeros(3 * self.sizes[slot])
        self._check(self.lib.md_batch_error_analysis(self.h, slot, _dp(out), _dp(f)),
                    "md_batch_error_analysis")
        return _ewald_summary(out), f.reshape(3, -1).T.copy()

    def energies(self, slot):
        """Engine.energies of one replica."""
        out = np.zeros(16)
        self._check(self.lib.md_batch_energies(self.h, slot, _dp(out)), "md_batch_energies")
        return dict(zip(Engine.ENERGIES, out))

    def observe(self, slot):
        """(brush height, z-density histogram of the polymer beads)."""
        h = c_double()
        prof = np.zeros(self.HIST_BINS)
        self._check(self.lib.md_batch_observe(self.h, slot, byref(h), _dp(prof), self.HIST_BINS),
                    "md_batch_observe")
        return h.value, prof
