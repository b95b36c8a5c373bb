"""Host-side checks that need no GPU: the C-ABI library loads and exports
every entry point of include/md_engine.h, the ctypes struct matches the C
layout, and the engine's set-up code (parameter files, derivation, topology,
initial configuration) agrees with the oracle and with the reference's rules.
"""
import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import orc
from md_brushes_amd import MdParams, build_topology, config, derive, initialize_position, load_library
from md_brushes_amd.configs import _apply, _base, reference_defaults, write_param_files

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "md_engine.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void)\s+(md_\w+)\s*\(", text, re.M)))


def test_library_exports_header():
    lib = load_library()
    names = _declared()
    assert len(names) >= 19
    for n in names:
        assert hasattr(lib, n), n


def test_struct_layout_matches_c(tmp_path):
    src = tmp_path / "layout.c"
    fields = [f for f, _ in MdParams._fields_]
    body = "".join(f'printf("%zu\\n", offsetof(md_params, {f}));' for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "md_engine.h"\n'
                   f'int main(void){{{body}printf("%zu\\n", sizeof(md_params));return 0;}}\n')
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)])
    out = [int(x) for x in subprocess.check_output([str(exe)]).split()]
    assert out[:-1] == [getattr(MdParams, f).offset for f in fields]
    assert out[-1] == ctypes.sizeof(MdParams)


def _random_params(rng):
    p = _base()
    star = rng.random() < 0.5
    lin = (not star) or rng.random() < 0.5
    arm = int(rng.integers(2, 5))
    _apply(p, [("Nga", int(rng.integers(1, 20)) if star else 0),
               ("Ngl", int(rng.integers(1, 20)) if lin else 0),
               ("arm", arm), ("Nma", int(rng.integers(2, 30))), ("Nml", int(rng.integers(2, 60))),
               ("man_l", int(rng.integers(1, 4))), ("man_s", int(rng.integers(1, 4))),
               ("qq", float(rng.choice([-1.0, 0.0, 1.0]))), ("qqi", float(rng.choice([0.0, 1.0, 2.0, 3.0]))),
               ("ion_ratio", float(rng.choice([0.0, 0.5, 1.0]))),
               ("sigmag", float(rng.uniform(0.005, 0.1))), ("Lz", float(rng.uniform(20, 200))),
               ("Z_empty", float(rng.uniform(0.5, 3))), ("tol", float(rng.uniform(1.5, 5))),
               ("tau_rf", float(rng.uniform(1, 1000))),
               ("charged_arm", [int(v) for v in rng.integers(0, 2, arm)])])
    if rng.random() < 0.3:
        p.ovr_alpha = float(rng.uniform(0.2, 1.0))
    if rng.random() < 0.3:
        p.ovr_ng[0], p.ovr_ng[1], p.ovr_ng[2] = 16, 32, 64
    if rng.random() < 0.3:
        p.ovr_skin = float(rng.uniform(0.3, 1.5))
    return p


def test_derivation_matches_oracle():
    rng = np.random.default_rng(1)
    cases = [reference_defaults()] + [_random_params(rng) for _ in range(300)]
    for p in cases:
        a, b = derive(p), orc.derive(p)
        assert a.as_dict() == b.as_dict()


def test_skin_override_widens_the_list_radius():
    """rvl_eff = rcl_s + 1 by the reference's rule (compt_acc.f90:162); a
    caller's skin gives rcl_s + 2 skin in the engine and in the oracle."""
    p = reference_defaults()
    a = derive(p)
    assert a.rvl_eff == a.rcl_s + 1
    q = p.copy()
    q.ovr_skin = 0.8
    b, c = derive(q), orc.derive(q)
    assert b.rvl_eff == b.rcl_s + 1.6 and c.rvl_eff == b.rvl_eff


def test_topology_matches_oracle():
    rng = np.random.default_rng(2)
    for _ in range(100):
        p = derive(_random_params(rng))
        charge, fene, anchor = build_topology(p)
        c2 = np.zeros(max(p.Nq, 1), np.int32)
        f2 = np.zeros(2 * max(p.N_bond, 1), np.int32)
        a2 = np.zeros(max(p.N_anchor, 1), np.int32)
        orc.lib().orc_build_topology(ctypes.byref(p), orc.ip(c2) if p.qq != 0 else None, orc.ip(f2),
                                     orc.ip(a2))
        assert np.array_equal(fene, f2[: 2 * p.N_bond].reshape(2, -1).T)
        assert np.array_equal(anchor, a2[: p.N_anchor])
        if p.qq != 0:
            assert np.array_equal(charge, c2[: p.Nq])


def test_star_topology_shape():
    p = _base()
    _apply(p, [("Nga", 2), ("Ngl", 1), ("arm", 3), ("Nma", 4), ("Nml", 5), ("man_s", 2), ("man_l", 1),
               ("charged_arm", [1, 0, 1])])
    p = derive(p)
    charge, fene, anchor = build_topology(p)
    per = 3 * 4 + 1
    assert list(anchor) == [1, 1 + per, 2 * per + 1]
    deg = np.bincount(fene.ravel(), minlength=p.NN + 1)
    assert deg[1] == 1 and deg[1 + 4] == 3  # anchor, branch point (stem + 2 arms)
    assert p.N_bond == 3 * 4 * 2 + 4
    # charged: every 2nd bead of arm 1 and arm 3, every bead of the chain
    assert list(charge[:4]) == [3, 5, 11, 13]


def test_read_param_files_roundtrip(tmp_path):
    p = reference_defaults()
    write_param_files(p, tmp_path / "s.txt", tmp_path / "f.txt", tmp_path / "q.txt")
    q = MdParams()
    rc = load_library().md_read_params(str(tmp_path / "s.txt").encode(), str(tmp_path / "f.txt").encode(),
                                       str(tmp_path / "q.txt").encode(), ctypes.byref(q))
    assert rc == 0
    for k, _ in MdParams._fields_:
        if k in ("charged_arm",):
            assert list(q.charged_arm)[: p.arm] == list(p.charged_arm)[: p.arm]
        elif k in ("seed",):
            continue
        elif k.startswith(("ovr", "pad")) or k in ("precision",):
            continue
        elif k in ("Nta", "Ntl", "Npe", "N_anchor", "Nq_PE", "Nq_salt_ions", "Nq", "NN", "N_bond",
                   "real_verlet", "ng", "Lx", "Ly", "Zfac", "sigma_sp", "rcl_s", "rcl_sp", "rvl_eff",
                   "alpha", "rcc", "rvc", "gdim"):
            continue
        else:
            a, b = getattr(q, k), getattr(p, k)
            assert (list(a) if hasattr(a, "__len__") else a) == (list(b) if hasattr(b, "__len__") else b), k


def test_read_fortran_exponents(tmp_path):
    s = tmp_path / "s.txt"
    lines = ["0", "75", "2", "5D-2   ! sigmag", "0.8333D0", "-1", "0", "0", "3", "10", "0", "51", "16", "3",
             "1", "0.97D0", "10", "10", "100", "100", "100", "10", "0.005D0"]
    s.write_text("\n".join(lines) + "\n")
    f = tmp_path / "f.txt"
    f.write_text("\n".join(["1", "1", "2", "1.12246D0", "1.81246D0", "1", "2.25", "30", "4", "4", "4", "3",
                            "1", "0.D0", "1.9D0", "800"]) + "\n")
    q = MdParams()
    assert load_library().md_read_params(str(s).encode(), str(f).encode(), None, ctypes.byref(q)) == 0
    assert q.sigmag == 0.05 and q.Beta == 0.8333 and q.dt == 0.005 and q.rcl == 1.12246
    assert q.ordr[0] == 4 and q.Nml == 51 and q.Ngl == 16


@pytest.mark.parametrize("name", ["small", "star3"])
def test_initial_configuration_is_valid(name):
    p = config(name)
    pos = initialize_position(p, 17)
    charge, fene, anchor = build_topology(p)
    a0 = anchor - 1
    assert np.all(pos[a0, 2] == 0)
    assert np.all(np.abs(pos[:, 0]) <= p.Lx / 2) and np.all(np.abs(pos[:, 1]) <= p.Ly / 2)
    movers = np.setdiff1d(np.arange(p.NN), a0)
    assert np.all(pos[movers, 2] >= 0.9) and np.all(pos[movers, 2] <= p.Lz - 0.9)
    d = pos[fene[:, 0] - 1, :3] - pos[fene[:, 1] - 1, :3]
    d[:, 0] -= p.Lx * np.round(d[:, 0] / p.Lx)
    d[:, 1] -= p.Ly * np.round(d[:, 1] / p.Ly)
    assert np.allclose(np.sqrt((d ** 2).sum(1)), p.R_bond)
    assert abs(pos[:, 3].sum()) < 1e-12  # neutral
    assert np.all(pos[charge[: p.Nq_PE] - 1, 3] == p.qq)
    # no two beads closer than the rejection radius of the growth: every pair
    # within the radius, found by a periodic k-d tree in x and y (z open: its
    # period is larger than any distance), then the minimum image squared
    # distance of each such pair as the growth computes it
    from scipy.spatial import cKDTree
    x = pos[:, :3].copy()
    x[:, 0] = np.mod(x[:, 0] + p.Lx / 2, p.Lx)
    x[:, 1] = np.mod(x[:, 1] + p.Ly / 2, p.Ly)
    x[:, 0][x[:, 0] >= p.Lx] = 0.0  # a rounding of mod(-tiny, L) to L
    x[:, 1][x[:, 1] >= p.Ly] = 0.0
    box = [p.Lx, p.Ly, 4 * (p.Lz + 1)]
    pairs = cKDTree(x, boxsize=box).query_pairs(np.sqrt(0.7) + 1e-6, output_type="ndarray")
    if len(pairs):
        dd = pos[pairs[:, 0], :3] - pos[pairs[:, 1], :3]
        dd[:, 0] -= p.Lx * np.round(dd[:, 0] / p.Lx)
        dd[:, 1] -= p.Ly * np.round(dd[:, 1] / p.Ly)
        assert (dd ** 2).sum(1).min() >= 0.7 - 1e-12


def test_initial_configuration_is_seeded():
    p = config("small")
    assert np.array_equal(initialize_position(p, 3), initialize_position(p, 3))
    assert not np.array_equal(initialize_position(p, 3), initialize_position(p, 4))


def test_replica_bytes_without_device():
    """md_replica_bytes sizes a replica on the host (no device call): more
    particles take more bytes, FP32 state less than FP64."""
    from md_brushes_amd import MD_FP32, MD_FP64, initialize_position, replica, replica_bytes
    sizes = {}
    for side in (2, 4):
        for prec in (MD_FP32, MD_FP64):
            p = replica(side, 40, 1, 3, 0.5, 7, prec)
            sizes[side, prec] = replica_bytes(p, initialize_position(p, 7))
    assert all(v > 0 for v in sizes.values())
    assert sizes[4, MD_FP32] > sizes[2, MD_FP32] and sizes[4, MD_FP64] > sizes[2, MD_FP64]
    assert sizes[4, MD_FP32] < sizes[4, MD_FP64]
