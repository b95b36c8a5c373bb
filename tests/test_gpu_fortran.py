"""The Fortran drop-in host (md_brushes_amd/fortran/md_brushes): same
parameter files and main-program flow as the reference, run end to end on
the GPU, and checked against the same run driven through the C-ABI from
Python."""
import os
import subprocess

import numpy as np
import pytest

from md_brushes_amd import MD_NOISE_PHILOX, Engine, derive, initialize_position, initialize_velocity
from md_brushes_amd.configs import _apply, _base, write_param_files

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "md_brushes_amd", "fortran", "md_brushes")


def _params():
    p = _base()
    _apply(p, [("Nga", 4), ("Ngl", 12), ("arm", 3), ("Nma", 10), ("Nml", 31), ("man_l", 2), ("man_s", 2),
               ("qqi", 3.0), ("ion_ratio", 0.5), ("Lz", 60.0), ("sigmag", 0.05), ("EF", 0.3),
               ("StepNum0", 40), ("StepNum", 60), ("DeltaStep1", 10), ("DeltaStep2", 50), ("DeltaStep3", 100),
               ("charged_arm", [1, 1, 0])])
    return p


def test_fortran_host_matches_c_abi(tmp_path):
    assert os.path.exists(EXE), "build the Fortran host first"
    raw = _params()
    write_param_files(raw, tmp_path / "system_data.txt", tmp_path / "force_data.txt", tmp_path / "q_arm.txt")
    (tmp_path / "data").mkdir()
    print("[fortran host] running", flush=True)
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "Finished!" in r.stdout
    h = np.loadtxt(tmp_path / "data" / "height.txt")
    # write_height's nine columns (in_output.f90:1342-1344): step, h_avg,
    # linear hl_max, hl_end, hl_avg, star hs_max, hs_end, hs_branch, hs_avg
    assert h.shape == (10, 9) and np.all(np.isfinite(h))
    assert np.all(h[:, 2] >= h[:, 4]) and np.all(h[:, 5] >= h[:, 8])  # a chain's top above its mean
    # error_analysis at the start (main.f90:30-33 of the reference), on the device
    rms = np.loadtxt(tmp_path / "data" / "rms_force.txt", ndmin=2)
    assert rms.shape == (1, 6) and rms[0, 0] == 0 and 0 < rms[0, 1] < 1e-2 and rms[0, 2] > 0
    assert "Rms force of coulomb force is:" in r.stdout
    # the same run through the C-ABI: the host's mesh rounding, seed and flow
    p = derive(raw)
    for d in range(3):
        k = 4
        while k < p.ng[d]:
            k *= 2
        p.ovr_ng[d] = k
    p.seed = 20261203
    p = derive(p)
    pos = initialize_position(p, p.seed)
    vel = initialize_velocity(p, p.seed + 1)
    e = Engine(p)
    e.upload(pos, np.zeros_like(vel))
    e.build_lists()
    e.upload(pos, vel)
    e.compute_force(0, MD_NOISE_PHILOX)
    e.run(1, 40, preheat=True)
    e.run(41, 60)
    pg, vg, _, _ = e.download()
    assert abs(pg[: p.Npe, 2].mean() - h[-1, 1]) < 1e-5
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    from systems import heights_ref
    ref = heights_ref(pg, p)
    for c, k in enumerate(["h_avg", "hl_max", "hl_end", "hl_avg", "hs_max", "hs_end", "hs_branch", "hs_avg"]):
        assert abs(h[-1, 1 + c] - ref[k]) < 1e-5, k
    e.close()
