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


def _host_params(raw):
    """The run's parameters as the host sets them: the reference's mesh
    rule and spline orders as they stand, its seed."""
    p = derive(raw)
    p.seed = 20261203
    return p


def _run_host(d, raw, env=None):
    write_param_files(raw, d / "system_data.txt", d / "force_data.txt", d / "q_arm.txt")
    (d / "data").mkdir(exist_ok=True)
    r = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, **(env or {})))
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0 and "Finished!" in r.stdout
    return r.stdout


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
    # error_analysis at the start (main.f90:30-33 of the reference), on the
    # device, equals the oracle's at the same configuration (this box has two
    # real-space list cells along x and y, so the production force sums the
    # reference's repeated pairs and its rmsf is large: the reference's own)
    rms = np.loadtxt(tmp_path / "data" / "rms_force.txt", ndmin=2)
    assert rms.shape == (1, 6) and rms[0, 0] == 0 and rms[0, 2] > 0
    assert "Rms force of coulomb force is:" in r.stdout
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    import orc
    # the same run through the C-ABI: the host's parameters, seed and flow
    p = _host_params(raw)
    pos = initialize_position(p, p.seed)
    vel = initialize_velocity(p, p.seed + 1)
    o = orc.Oracle(p)
    o.set_state(pos, vel)
    o.build_lists()
    assert abs(rms[0, 1] - o.error_analysis()[0]["rmsf"]) < 2e-6
    e = Engine(p)
    e.upload(pos, np.zeros_like(vel))
    e.build_lists()
    e.upload(pos, vel)
    e.compute_force(0, MD_NOISE_PHILOX)
    e.run(1, 40, preheat=True)
    e.run(41, 60)
    pg, vg, _, _ = e.download()
    assert abs(pg[: p.Npe, 2].mean() - h[-1, 1]) < 1e-5
    from systems import heights_ref
    ref = heights_ref(pg, p)
    for c, k in enumerate(["h_avg", "hl_max", "hl_end", "hl_avg", "hs_max", "hs_end", "hs_branch", "hs_avg"]):
        assert abs(h[-1, 1 + c] - ref[k]) < 1e-5, k
    e.close()


def test_fortran_exact_resume(tmp_path):
    """A run stopped at a checkpoint and restarted with MD_EXACT_RESUME=1
    (data/state.bin) ends in the state of the uninterrupted run, bit for
    bit; the restart continues at the step after the checkpoint
    (in_output.f90:616)."""
    assert os.path.exists(EXE), "build the Fortran host first"

    def run(d, steps, env=None):
        raw = _params()
        _apply(raw, [("StepNum", steps), ("DeltaStep3", 50)])
        write_param_files(raw, d / "system_data.txt", d / "force_data.txt", d / "q_arm.txt")
        (d / "data").mkdir(exist_ok=True)
        r = subprocess.run([EXE], cwd=d, capture_output=True, text=True, timeout=120,
                           env=dict(os.environ, **(env or {})))
        print(r.stdout[-1500:], r.stderr[-1500:])
        assert r.returncode == 0 and "Finished!" in r.stdout
        return r.stdout

    full, part = tmp_path / "full", tmp_path / "part"
    full.mkdir()
    part.mkdir()
    print("[fortran exact resume] uninterrupted", flush=True)
    run(full, 60)  # steps 1..100, checkpoints at 50 and 100
    print("[fortran exact resume] stop at 50, resume", flush=True)
    run(part, 10)  # steps 1..50
    out = run(part, 60, {"MD_EXACT_RESUME": "1"})
    assert "exact resume from data/state.bin at step" in out
    a = np.fromfile(full / "data" / "state.bin", dtype=np.uint8)
    b = np.fromfile(part / "data" / "state.bin", dtype=np.uint8)
    assert a.size == b.size
    # md_state_save's layout: header, pos, vel, acc, held Coulomb force, list
    # lengths, ELL list, the two cell-count buffers, cell members as binned,
    # each particle's cell, [the walked real-space list: lengths, slots, the
    # two count buffers,] control block
    NN, fp32, ncell1, ccap1, nbcap1, Nq, ncell2, nbcap2 = (
        int(x) for x in np.frombuffer(a[:72].tobytes(), np.int64)[1:9])
    rs = 4 if fp32 else 8
    sizes = [32 * NN, 24 * NN, 24 * NN, 3 * NN * rs, 4 * NN, 4 * nbcap1 * NN, 8 * ncell1,
             4 * ncell1 * ccap1, 4 * NN]
    if nbcap2:
        sizes += [4 * Nq, 4 * nbcap2 * Nq, 8 * ncell2]
    o = 72
    for k, n in enumerate(sizes):
        x, y = a[o:o + n], b[o:o + n]
        if k == 7:  # members in binning order: the order of the atomics is free
            mx = np.frombuffer(x.tobytes(), np.int32).reshape(ncell1, ccap1)
            my = np.frombuffer(y.tobytes(), np.int32).reshape(ncell1, ccap1)
            cnt = np.frombuffer(a[o - 8 * ncell1:o].tobytes(), np.int32).reshape(2, ncell1)
            for c in range(ncell1):
                m = min(max(cnt[0, c], cnt[1, c]), ccap1)
                assert np.array_equal(np.sort(mx[c, :m]), np.sort(my[c, :m]))
        else:
            assert np.array_equal(x, y), k
        o += n
    assert np.array_equal(a[o:], b[o:])  # control block


def test_fortran_text_restart(tmp_path):
    """A restart from the text checkpoint (no MD_EXACT_RESUME) follows the
    reference's restart (main.f90:40-64): positions and velocities read at
    six decimals (in_output.f90:605-618), lists rebuilt, velocities
    rescaled, then a force whose Coulomb part is evaluated afresh (the step
    counter is still 0 there, compt_acc.f90:841) and held by the following
    steps; the run continues at the checkpoint step + 1.  The same sequence
    through the C-ABI reproduces the host's force and trajectory, and its
    held Coulomb force equals the oracle's."""
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    import orc
    from systems import heights_ref
    assert os.path.exists(EXE), "build the Fortran host first"
    raw = _params()
    _apply(raw, [("StepNum", 10), ("DeltaStep3", 50)])
    print("[fortran text restart] steps 1..50", flush=True)
    _run_host(tmp_path, raw)  # checkpoint at step 50
    pos_txt = np.loadtxt(tmp_path / "data" / "pos1.txt")
    vel_txt = np.loadtxt(tmp_path / "data" / "vel1.txt")
    _apply(raw, [("StepNum", 40)])
    print("[fortran text restart] restart, steps 51..80", flush=True)
    out = _run_host(tmp_path, raw)
    assert "exact resume" not in out
    acc_host = np.loadtxt(tmp_path / "data" / "acc.txt")
    vel_host = np.loadtxt(tmp_path / "data" / "vel.txt")
    h = np.loadtxt(tmp_path / "data" / "height.txt")
    assert list(h[-3:, 0]) == [60, 70, 80]

    p = _host_params(raw)
    e = Engine(p)
    e.upload(pos_txt, vel_txt)
    e.build_lists()
    e.rescale_velocity()
    e.compute_force(0, MD_NOISE_PHILOX)
    pos0, vel0, acc0, _ = e.download()
    fc = e.coulomb()
    assert np.abs(vel0 - vel_host).max() < 1e-6
    assert np.abs(acc0 - acc_host).max() < 1e-5
    o = orc.Oracle(p)
    o.set_state(pos_txt)
    o.build_lists()
    ref = o.coulomb()
    charged = np.abs(pos_txt[:, 3]) > 0
    assert np.abs(fc[charged]).max() > 1e-2  # the held Coulomb force is there
    assert orc.rel_rms(fc, ref) < 1e-10
    e.run(51, 30)
    pg = e.download()[0]
    ref_h = heights_ref(pg, p)
    for c, k in enumerate(["h_avg", "hl_max", "hl_end", "hl_avg", "hs_max", "hs_end", "hs_branch", "hs_avg"]):
        assert abs(h[-1, 1 + c] - ref_h[k]) < 1e-5, k
    e.close()
