"""Parity of the HIP engine with the FP64 oracle, through the C-ABI.

Bars: neighbour lists bit-exact (as sorted pair sets); FP64 forces within
1e-11 relative RMS of the oracle; FP32 forces within 1e-4 relative RMS
(the FP32 accuracy gate); one step's drift bit-exact; injected-noise
trajectories step for step within 1e-9 (FP64).
"""
import numpy as np
import pytest

import orc
from md_brushes_amd import MD_FP32, MD_FP64, MD_NOISE_ETA, MD_NOISE_PHILOX, Engine, MDError
from systems import etas, system

pytestmark = pytest.mark.gpu

NAMES = ["small", "default"]


def _engine(name, prec=MD_FP64):
    p, pos, vel = system(name, prec)
    e = Engine(p)
    e.upload(pos, vel)
    e.build_lists()
    return p, pos, vel, e


def _oracle(name):
    p, pos, vel = system(name, MD_FP64)
    o = orc.Oracle(p)
    o.set_state(pos, vel)
    o.build_lists()
    return o


@pytest.mark.parametrize("name", NAMES)
def test_lists_bit_exact(name):
    print(f"[lists] {name}", flush=True)
    p, pos, vel, e = _engine(name)
    o = _oracle(name)
    for which in (0, 1):
        if which == 1 and not p.real_verlet:
            continue
        g = e.lists(which)
        r = o.lists(which)
        r = r[np.lexsort((r[:, 1], r[:, 0]))]
        assert g.shape == r.shape, (which, g.shape, r.shape)
        assert np.array_equal(g, r)
    e.close()


@pytest.mark.parametrize("name", NAMES)
def test_forces_fp64(name):
    print(f"[forces fp64] {name}", flush=True)
    p, pos, vel, e = _engine(name)
    o = _oracle(name)
    fs, fc = e.force_parts()
    ref_s = o.lj_force() + o.fene_force()[0]
    ref_c = o.coulomb()
    assert orc.rel_rms(fs, ref_s) < 1e-11
    assert orc.rel_rms(fc, ref_c) < 1e-10
    e.close()


@pytest.mark.parametrize("name", NAMES)
def test_forces_fp32_gate(name):
    print(f"[forces fp32] {name}", flush=True)
    p, pos, vel, e = _engine(name, MD_FP32)
    o = _oracle(name)
    fs, fc = e.force_parts()
    ref = o.lj_force() + o.fene_force()[0] + o.coulomb()
    err = orc.rel_rms(fs + fc, ref)
    print(f"fp32 relative force rms {err:.3e}")
    assert err < 1e-4
    e.close()


@pytest.mark.parametrize("name", NAMES)
def test_step_with_injected_noise(name):
    print(f"[trajectory] {name}", flush=True)
    p, pos, vel, e = _engine(name)
    o = _oracle(name)
    noise = etas(p.NN, 12)
    assert o.compute_force(0, MD_NOISE_ETA, noise[0]) == 0
    e.compute_force(0, MD_NOISE_ETA, noise[0])
    pg, vg, ag, _ = e.download()
    po, vo, ao, _, _ = o.state()
    assert orc.rel_rms(ag, ao) < 1e-10
    # one step: the drift uses only state both sides hold bit-identically
    acc0 = ao.copy()
    e.upload(pos, vel, acc0)
    e.build_lists()
    o.set_state(pos, vel, acc0)
    o.set_drmax(0, 0)
    o.build_lists()
    for s in range(1, 12):
        assert o.new_position(s, MD_NOISE_ETA, noise[s]) == 0
        assert o.update_verlet_list() == 0
        e.new_position(s, MD_NOISE_ETA, noise[s])
        e.update_verlet_list()
        pg, vg, ag, dg = e.download()
        po, vo, ao, _, do = o.state()
        if s == 1:
            assert np.array_equal(pg, po), np.abs(pg - po).max()
        assert np.abs(pg - po).max() < 1e-9
        assert orc.rel_rms(vg, vo) < 1e-9
        assert np.allclose(dg, do, rtol=0, atol=1e-12)
    e.close()


def _cycle_form(form, monkeypatch):
    """How md_run issues its cycles: the first as a captured graph's replay
    and the others launch by launch on the stream (the default), every
    cycle replayed (MD_GRAPH=1), or every cycle on the stream
    (MD_GRAPH=0)."""
    monkeypatch.delenv("MD_GRAPH", raising=False)
    if form != "mixed":
        monkeypatch.setenv("MD_GRAPH", "1" if form == "graph" else "0")


@pytest.mark.parametrize("cycles", ["mixed", "graph", "stream"])
def test_graph_run_matches_stepwise(cycles, monkeypatch):
    """Cycles (the first replayed as a captured graph and the others issued
    on the stream, every cycle replayed, or every cycle on the stream)
    equal the step-by-step calls."""
    print(f"[{cycles} cycles vs stepwise] default", flush=True)
    _cycle_form(cycles, monkeypatch)
    p, pos, vel, e1 = _engine("default")
    p, pos, vel, e2 = _engine("default")
    e1.compute_force(0, MD_NOISE_PHILOX)
    e2.compute_force(0, MD_NOISE_PHILOX)
    e1.run(1, 25)
    for s in range(1, 26):
        e2.new_position(s, MD_NOISE_PHILOX)
        e2.update_verlet_list()
    a = e1.download()
    b = e2.download()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    e1.close()
    e2.close()


def test_eager_cycles_match_graph(monkeypatch):
    """The default mix of replayed and stream-issued cycles, every cycle
    replayed as a captured graph (MD_GRAPH=1) and every cycle on the stream
    (MD_GRAPH=0) give the same states, across runs of odd length."""
    print("[eager vs graph cycles] default fp32", flush=True)
    out = []
    for form in ("mixed", "graph", "stream"):
        _cycle_form(form, monkeypatch)
        p, pos, vel, e = _engine("default", MD_FP32)
        e.compute_force(0, MD_NOISE_PHILOX)
        e.run(1, 30)
        e.run(31, 17)
        e.run(48, 25)
        out.append(e.download())
        e.close()
    for other in out[1:]:
        for x, y in zip(out[0], other):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("cycles", ["mixed", "graph", "stream"])
def test_odd_multistep_run_matches_stepwise(cycles, monkeypatch):
    """With an odd multistep, consecutive replayed cycles start on
    alternating position buffers (one launch sequence or graph per buffer):
    the run still equals the stepwise sequence, Coulomb at the multiples of
    multistep."""
    print(f"[odd multistep {cycles} cycles vs stepwise] small", flush=True)
    _cycle_form(cycles, monkeypatch)
    from md_brushes_amd import config
    p = config("small")
    p.multistep = 5
    _, pos, vel = system("small")
    es = []
    for _ in range(2):
        e = Engine(p)
        e.upload(pos, vel)
        e.build_lists()
        e.compute_force(0, MD_NOISE_PHILOX)
        es.append(e)
    e1, e2 = es
    e1.run(1, 23)
    e1.run(24, 18)
    for s in range(1, 42):
        e2.new_position(s, MD_NOISE_PHILOX)
        e2.update_verlet_list()
    a = e1.download()
    b = e2.download()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    e1.close()
    e2.close()


@pytest.mark.parametrize("cycles", ["mixed", "graph", "stream"])
def test_preheat_graph_matches_stepwise(cycles, monkeypatch):
    """Pre-heat (main.f90:67-75) replayed in cycles with the gated rescale
    kernels equals the step-by-step sequence with rescale_velocity after
    every 20th step; the run starts off the cycle alignment."""
    print(f"[preheat {cycles} cycles vs stepwise] small", flush=True)
    _cycle_form(cycles, monkeypatch)
    p, pos, vel, e1 = _engine("small")
    p, pos, vel, e2 = _engine("small")
    e1.compute_force(0, MD_NOISE_PHILOX)
    e2.compute_force(0, MD_NOISE_PHILOX)
    e1.run(1, 3)
    e1.run(4, 63, preheat=True)
    e1.run(67, 14)
    for s in range(1, 81):
        e2.new_position(s, MD_NOISE_PHILOX)
        if 4 <= s <= 66 and s % 20 == 0:
            e2.rescale_velocity()
        e2.update_verlet_list()
    a = e1.download()
    b = e2.download()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    e1.close()
    e2.close()


def test_run_is_deterministic():
    print("[determinism] default fp32", flush=True)
    out = []
    for _ in range(2):
        p, pos, vel, e = _engine("default", MD_FP32)
        e.compute_force(0, MD_NOISE_PHILOX)
        e.run(1, 40)
        out.append(e.download())
        e.close()
    for x, y in zip(out[0], out[1]):
        assert np.array_equal(x, y)


def test_philox_run_matches_oracle():
    print("[philox run] small", flush=True)
    p, pos, vel, e = _engine("small")
    o = _oracle("small")
    assert o.compute_force(0, MD_NOISE_PHILOX) == 0
    e.compute_force(0, MD_NOISE_PHILOX)
    assert o.run(1, 20) == 0
    e.run(1, 20)
    pg, vg, _, _ = e.download()
    po, vo, _, _, _ = o.state()
    assert np.abs(pg - po).max() < 1e-8
    e.close()


def test_time_kernels_restores_state():
    """Kernel timing leaves positions, velocities and forces as they were,
    and a run after it stays on the oracle's lists (no bond breaks)."""
    print("[time_kernels restores] default fp32", flush=True)
    p, pos, vel, e = _engine("default", MD_FP32)
    e.compute_force(0, MD_NOISE_PHILOX)
    e.run(1, 20)
    before = e.download()
    ms = e.time_kernels(20)
    after = e.download()
    for x, y in zip(before[:3], after[:3]):
        assert np.array_equal(x, y)
    assert ms["short_range"] > 0 and ms["real"] > 0
    e.run(21, 40)
    e.download()
    e.close()


def test_rescale_velocity():
    print("[rescale] small", flush=True)
    p, pos, vel, e = _engine("small")
    o = _oracle("small")
    v = np.array(vel) * 1.7 + 0.05
    e.upload(pos, v)
    o.set_state(pos, v)
    e.rescale_velocity()
    o.rescale_velocity()
    _, vg, _, _ = e.download()
    _, vo, _, _, _ = o.state()
    assert np.abs(vg - vo).max() < 1e-12
    e.close()


def test_bond_break_reported():
    print("[bond break] small", flush=True)
    p, pos, vel = system("small")
    q = np.array(pos)
    q[5, 2] += 3.0  # stretch a bond past R0 = 1.5
    e = Engine(p)
    e.upload(q, vel)
    e.build_lists()
    e.compute_force(0)
    with pytest.raises(MDError) as ei:
        e.download()
    assert ei.value.code == 2 and "bond" in str(ei.value)
    e.close()


@pytest.mark.parametrize("order", [3, 5, 6])
def test_spme_general_order_fp64(order):
    print(f"[spme order {order}] small", flush=True)
    from md_brushes_amd import config
    p = config("small", order=order)
    _, pos, vel = system("small")
    e = Engine(p)
    e.upload(pos, vel)
    e.build_lists()
    o = orc.Oracle(p)
    o.set_state(pos, vel)
    o.build_lists()
    _, fc = e.force_parts()
    assert orc.rel_rms(fc, o.coulomb()) < 1e-10
    e.close()


@pytest.mark.parametrize("prec", [MD_FP64, MD_FP32])
def test_fft_plane_paths_agree(prec, monkeypatch):
    """The fused x/y FFT planes (spectrum in bit-reversed y order) and the
    separate x and y passes give the same Coulomb forces."""
    print(f"[fft planes vs passes] default prec={prec}", flush=True)
    p, pos, vel = system("default", prec)
    out = []
    for flag in ("1", "0"):
        monkeypatch.setenv("MD_FFT_PLANES", flag)
        e = Engine(p)
        e.upload(pos, vel)
        e.build_lists()
        out.append(e.force_parts()[1])
        e.close()
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("prec", [MD_FP64, MD_FP32])
def test_launch_forms_do_not_change_results(prec, monkeypatch):
    """Every launch form the autotuner may choose (fused planes or x/y
    passes, z-pass lines per workgroup, sweep form, the sweep's workgroup
    count, spreading inside the forward plane kernel, riders dispatched
    ahead of a launch's own workgroups) gives bitwise the same
    Coulomb forces, and autotuning keeps them."""
    print(f"[launch forms] default prec={prec}", flush=True)
    p, pos, vel = system("default", prec)

    def coulomb(env):
        for k in ("MD_FFT_PLANES", "MD_Z_LINES", "MD_REAL_SWEEP", "MD_RIDE", "MD_REAL_GRID", "MD_PLANE_SPREAD",
                  "MD_RIDE_PARTS", "MD_RIDE_LEAD", "MD_RIDE_LEAD_G"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        e = Engine(p)
        e.upload(pos, vel)
        e.build_lists()
        f = e.force_parts()[1]
        return e, f

    e, ref = coulomb({})
    e.close()
    for env in ({"MD_FFT_PLANES": "0", "MD_Z_LINES": "2"}, {"MD_Z_LINES": "16", "MD_REAL_SWEEP": "wave"},
                {"MD_FFT_PLANES": "0", "MD_Z_LINES": "8", "MD_REAL_SWEEP": "wg"},
                {"MD_RIDE": "0", "MD_REAL_SWEEP": "wg"}, {"MD_RIDE": "1", "MD_REAL_SWEEP": "wave"},
                {"MD_REAL_GRID": "2048"}, {"MD_REAL_GRID": "300", "MD_RIDE": "0"},
                {"MD_PLANE_SPREAD": "1"}, {"MD_PLANE_SPREAD": "1", "MD_RIDE": "0", "MD_REAL_SWEEP": "wave"},
                {"MD_PLANE_SPREAD": "1", "MD_REAL_SWEEP": "wave"}, {"MD_PLANE_SPREAD": "1", "MD_REAL_GRID": "300"},
                {"MD_PLANE_SPREAD": "1", "MD_FFT_PLANES": "0"}, {"MD_PLANE_SPREAD": "0"},
                {"MD_PLANE_SPREAD": "1", "MD_RIDE_PARTS": "1"}, {"MD_PLANE_SPREAD": "1", "MD_RIDE_PARTS": "2"},
                {"MD_PLANE_SPREAD": "1", "MD_RIDE": "1", "MD_RIDE_PARTS": "2", "MD_RIDE_LEAD": "512"},
                {"MD_PLANE_SPREAD": "1", "MD_RIDE": "1", "MD_REAL_SWEEP": "wave", "MD_RIDE_LEAD": "768"},
                {"MD_PLANE_SPREAD": "0", "MD_RIDE": "1", "MD_RIDE_LEAD": "256"},
                {"MD_PLANE_SPREAD": "1", "MD_RIDE": "1", "MD_RIDE_LEAD": "512", "MD_RIDE_LEAD_G": "65536"}):
        e, f = coulomb(env)
        e.close()
        assert np.array_equal(f, ref), env
    e, _ = coulomb({})
    choice = e.autotune(2)
    assert choice["z_lines"] in (2, 4, 8, 16) and choice["fused_planes"] in (0, 1)
    assert choice["real_shared_launch"] in (0, 1) and choice["real_grid"] in (1024, 2048)
    assert choice["plane_spreading"] in (0, 1) and choice["ride_parts"] in (1, 2, 3)
    assert choice["ride_lead"] in (0, 256, 512, 768) and choice["ride_lead_gather"] in (0, 512, 1024, 65536)
    assert np.array_equal(e.force_parts()[1], ref)
    e.close()


@pytest.mark.parametrize("prec", [MD_FP64, MD_FP32])
def test_concurrent_real_space_matches_one_stream(prec, monkeypatch):
    """The real-space sweep on its own stream beside the SPME kernels gives
    the states of the one-stream order, in graph replay and stepwise."""
    print(f"[two streams vs one] default prec={prec}", flush=True)
    out = []
    for flag in ("1", "0"):
        monkeypatch.setenv("MD_TWO_STREAMS", flag)
        p, pos, vel, e = _engine("default", prec)
        e.compute_force(0, MD_NOISE_PHILOX)
        e.run(1, 33)
        out.append(e.download())
        out.append(e.force_parts())
        e.close()
    for x, y in zip(out[0] + out[1], out[2] + out[3]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("sweep", ["wave", "wg"])
@pytest.mark.parametrize("name", NAMES)
def test_real_space_sweep_forms(name, sweep, monkeypatch):
    """Both forms of the real-space cell sweep (a wavefront or a workgroup
    per work item) give the oracle's Coulomb forces."""
    print(f"[real-space sweep {sweep}] {name}", flush=True)
    monkeypatch.setenv("MD_REAL_SWEEP", sweep)
    p, pos, vel, e = _engine(name)
    fc = e.force_parts()[1]
    e.close()
    o = _oracle(name)
    assert orc.rel_rms(fc, o.coulomb()) < 1e-10


@pytest.mark.parametrize("prec", [MD_FP64, MD_FP32])
def test_real_space_sweep_forms_identical(prec, monkeypatch):
    """The two sweep forms sum every target in the same order: bitwise equal
    Coulomb forces, so an engine may switch form with its size."""
    print(f"[real-space sweep forms bitwise] default prec={prec}", flush=True)
    p, pos, vel = system("default", prec)
    out = []
    for sweep in ("wave", "wg"):
        monkeypatch.setenv("MD_REAL_SWEEP", sweep)
        e = Engine(p)
        e.upload(pos, vel)
        e.build_lists()
        out.append(e.force_parts()[1])
        e.close()
    assert np.array_equal(out[0], out[1])


def test_error_analysis_matches_oracle():
    """error_analysis on the device (standard Ewald at tol 5 + real space,
    FP64) equals the oracle's restatement: the reference force, its
    parameters and the SPME accuracy figure rmsf."""
    print("[error_analysis] small fp64", flush=True)
    p, pos, vel, e = _engine("small")
    o = _oracle("small")
    g, fse = e.error_analysis()
    r, fse_o, _ = o.error_analysis()
    e.close()
    for k in ("K1", "K2", "K3"):
        assert g[k] == int(r[k]), k
    for k in ("alpha", "rcc", "f_r", "f_k"):
        assert abs(g[k] - r[k]) <= 1e-12 * abs(r[k]), k
    assert orc.rel_rms(fse, fse_o) < 1e-10
    assert abs(g["rmsf"] - r["rmsf"]) < 1e-6 * r["rmsf"]


def test_error_analysis_default_fp32():
    """At the default brush the FP32 engine's Coulomb force is within 5e-4
    of the standard-Ewald reference (the FP64 oracle measures 4.03e-4 there:
    the 64x64x256 order-4 mesh's discretisation error), whose device value
    equals the oracle's; the check leaves the trajectory as it would have
    been."""
    print("[error_analysis] default fp32", flush=True)
    p, pos, vel, e = _engine("default", MD_FP32)
    g, fse = e.error_analysis()
    o = _oracle("default")
    r, fse_o, _ = o.error_analysis()
    assert orc.rel_rms(fse, fse_o) < 1e-10
    assert g["rmsf"] < 5e-4
    # the FP32 figure differs from the FP64 oracle's by the FP32 rounding only
    assert abs(g["rmsf"] - r["rmsf"]) < 0.1 * r["rmsf"]
    e2 = _engine("default", MD_FP32)[3]
    for x in (e, e2):
        x.compute_force(0, MD_NOISE_PHILOX)
    e.run(1, 10)
    e.error_analysis()
    e.run(11, 10)
    e2.run(1, 20)
    for x, y in zip(e.download(), e2.download()):
        assert np.array_equal(x, y)
    e.close()
    e2.close()


@pytest.mark.parametrize("prec", [MD_FP64, MD_FP32])
def test_exact_resume(prec):
    """A run saved mid-way (md_state_save) and loaded into a fresh engine
    (md_state_load) continues bit for bit; a state of another system is
    refused."""
    print(f"[exact resume] default prec={prec}", flush=True)
    p, pos, vel, e1 = _engine("default", prec)
    e1.compute_force(0, MD_NOISE_PHILOX)
    e1.run(1, 47)
    ref = e1.download()
    e1.close()
    p, pos, vel, e2 = _engine("default", prec)
    e2.compute_force(0, MD_NOISE_PHILOX)
    e2.run(1, 23)
    state = e2.save_state()
    e2.close()
    e3 = Engine(p)
    e3.load_state(state)
    e3.run(24, 24)
    got = e3.download()
    e3.close()
    for x, y in zip(ref, got):
        assert np.array_equal(x, y)
    q, qpos, qvel = system("small", prec)
    e4 = Engine(q)
    with pytest.raises(MDError):
        e4.load_state(state)
    e4.close()
    # a truncated state (a checkpoint file cut short) and one with corrupt
    # list capacities are refused before the engine changes
    e5 = Engine(p)
    e5.upload(pos, vel)
    before = e5.download()
    bad = bytearray(state)
    bad[40:48] = np.int64(1 << 40).tobytes()  # nbcap1
    for st in (state[: len(state) // 2], state[:-1], bytes(bad)):
        with pytest.raises(MDError):
            e5.load_state(st)
    for x, y in zip(before, e5.download()):
        assert np.array_equal(x, y)
    e5.close()


def _small_box(prec=MD_FP64, rsk=0.2):
    """A 2x2-chain brush (Lx = 28.3) whose real-space list cells number two
    along x and y: the reference's stencil (compt_acc.f90:538-560) then lists
    the other cell's pairs twice and real_space sums both.  A thin skin
    (rsk = 0.2) makes the real-space list rebuild every few steps."""
    from md_brushes_amd import derive
    from md_brushes_amd.configs import replica
    p = replica(2, 50, 1, 3, 0.5, 20261203, prec)
    p.rsk = rsk
    p = derive(p)
    assert p.real_verlet == 1 and int(p.Lx / p.rvc) == 2
    from md_brushes_amd import initialize_position, initialize_velocity
    return p, initialize_position(p, 31), initialize_velocity(p, 32)


def test_small_box_reference_lists_and_forces():
    """Lists with the reference's repeated pairs are bit-exact; the FP64
    Coulomb force (which walks that list) and the FP32 gate match the
    oracle."""
    print("[small box] lists and forces", flush=True)
    p, pos, vel = _small_box()
    e = Engine(p)
    e.upload(pos, vel)
    e.build_lists()
    o = orc.Oracle(p)
    o.set_state(pos, vel)
    o.build_lists()
    for which in (0, 1):
        g = e.lists(which)
        r = o.lists(which)
        r = r[np.lexsort((r[:, 1], r[:, 0]))]
        assert np.array_equal(g, r), which
    r = o.lists(1)
    assert len(np.unique(r, axis=0)) < len(r)  # the repeats are there
    fs, fc = e.force_parts()
    assert orc.rel_rms(fs, o.lj_force() + o.fene_force()[0]) < 1e-11
    ref_c = o.coulomb()
    assert orc.rel_rms(fc, ref_c) < 1e-10
    e.close()
    q = p.copy()
    q.precision = MD_FP32
    e = Engine(q)
    e.upload(pos, vel)
    e.build_lists()
    fs, fc = e.force_parts()
    err = orc.rel_rms(fs + fc, o.lj_force() + o.fene_force()[0] + ref_c)
    print(f"fp32 relative force rms {err:.3e}")
    assert err < 1e-4
    e.close()


def test_small_box_trajectory():
    """Steps with injected noise (stepwise) and a Philox run (graph-replayed
    cycles) follow the oracle through several real-space list rebuilds and
    Coulomb evaluations."""
    print("[small box] trajectory", flush=True)
    p, pos, vel = _small_box()
    noise = etas(p.NN, 41, seed=3)
    e = Engine(p)
    e.upload(pos, vel)
    e.build_lists()
    o = orc.Oracle(p)
    o.set_state(pos, vel)
    o.build_lists()
    assert o.compute_force(0, MD_NOISE_ETA, noise[0]) == 0
    e.compute_force(0, MD_NOISE_ETA, noise[0])
    for s in range(1, 41):
        assert o.new_position(s, MD_NOISE_ETA, noise[s]) == 0
        assert o.update_verlet_list() == 0
        e.new_position(s, MD_NOISE_ETA, noise[s])
        e.update_verlet_list()
    pg, vg, ag, dg = e.download()
    po, vo, ao, _, do = o.state()
    assert np.abs(pg - po).max() < 1e-8
    assert orc.rel_rms(ag, ao) < 1e-8
    e.close()
    e = Engine(p)
    e.upload(pos, vel)
    e.build_lists()
    o = orc.Oracle(p)
    o.set_state(pos, vel)
    o.build_lists()
    assert o.compute_force(0, MD_NOISE_PHILOX) == 0
    e.compute_force(0, MD_NOISE_PHILOX)
    assert o.run(1, 40) == 0
    e.run(1, 40)
    pg = e.download()[0]
    assert np.abs(pg - o.state()[0]).max() < 1e-8
    e.close()


ENERGY_TERMS = ["lj", "wall", "fene", "real", "recip", "slab", "self", "field", "kinetic", "virial_lj",
                "virial_fene", "virial_real"]


@pytest.mark.parametrize("name", ["small", "default", "small_box"])
def test_energies_match_oracle(name):
    """md_energies (FP64 sums on the device) equals the oracle's energies
    term by term: FP64 engine within 1e-9 relative, FP32 engine within 1e-4
    (the error printed per term)."""
    print(f"[energies] {name}", flush=True)
    if name == "small_box":
        p, pos, vel = _small_box()
    else:
        p, pos, vel = system(name, MD_FP64)
    o = orc.Oracle(p)
    o.set_state(pos, vel)
    o.build_lists()
    ref = o.energies()
    for prec, tol in ((MD_FP64, 1e-9), (MD_FP32, 1e-4)):
        q = p.copy()
        q.precision = prec
        e = Engine(q)
        e.upload(pos, vel)
        e.build_lists()
        got = e.energies()
        e.close()
        errs = {}
        for k in ENERGY_TERMS:
            scale = max(abs(ref[k]), 1e-6 * abs(ref["potential"]), 1e-12)
            errs[k] = abs(got[k] - ref[k]) / scale
        print(("fp64" if prec == MD_FP64 else "fp32"), {k: f"{v:.1e}" for k, v in errs.items()})
        for k, v in errs.items():
            assert v < tol, (k, got[k], ref[k])
        assert abs(got["coulomb"] - (got["real"] + got["recip"] + got["slab"] + got["self"])) < 1e-9 * abs(
            got["coulomb"]) + 1e-12


def test_energies_leave_trajectory():
    """An energy evaluation between runs leaves the trajectory bit for bit."""
    print("[energies] trajectory unchanged", flush=True)
    out = []
    for probe in (False, True):
        p, pos, vel, e = _engine("default", MD_FP32)
        e.compute_force(0, MD_NOISE_PHILOX)
        e.run(1, 20)
        if probe:
            e.energies()
        e.run(21, 20)
        out.append(e.download())
        e.close()
    for x, y in zip(out[0], out[1]):
        assert np.array_equal(x, y)


def _mesh_variant(name, ng, ordr, **kw):
    """A named system with the SPME mesh and per-axis orders given."""
    p, pos, vel = system(name, MD_FP64)
    q = p.copy()
    for d in range(3):
        q.ovr_ng[d] = ng[d]
        q.ordr[d] = ordr[d]
    for k, v in kw.items():
        setattr(q, k, v)
    from md_brushes_amd import derive
    return derive(q), pos, vel


MESHES = {
    # odd sizes on every axis (an odd row count, a lone row per plane)
    "odd_45x45x135_p5": ("small", (45, 45, 135), (5, 5, 5)),
    # prime factors 7, 11, 73 (the reference's rule gives lengths like 219 = 3 x 73)
    "primes_21x22x146_p4": ("small", (21, 22, 146), (4, 4, 4)),
    # an order per axis (force_data.txt holds three, compt_acc.f90:1232-1234)
    "orders_32x30x128_p456": ("small", (32, 30, 128), (4, 5, 6)),
    "orders_36x40x120_p635": ("small", (36, 40, 120), (6, 3, 5)),
}


@pytest.mark.parametrize("case", sorted(MESHES))
@pytest.mark.parametrize("planes", ["1", "0"])
def test_spme_any_mesh_and_orders(case, planes, monkeypatch):
    """Meshes of any length (mixed-radix Stockham passes: radix 2, 3, 4, 5,
    8 butterflies and a DFT pass for other primes) and unequal orders per
    axis: FP64 Coulomb forces within 1e-10 of the oracle, with the fused
    x/y planes and with the separate passes; FP32 within 1e-4."""
    print(f"[spme mesh] {case} planes={planes}", flush=True)
    name, ng, ordr = MESHES[case]
    monkeypatch.setenv("MD_FFT_PLANES", planes)
    p, pos, vel = _mesh_variant(name, ng, ordr)
    o = orc.Oracle(p)
    o.set_state(pos, vel)
    o.build_lists()
    ref = o.coulomb()
    for prec, tol in ((MD_FP64, 1e-10), (MD_FP32, 1e-4)):
        q = p.copy()
        q.precision = prec
        e = Engine(q)
        e.upload(pos, vel)
        e.build_lists()
        fc = e.force_parts()[1]
        e.close()
        err = orc.rel_rms(fc, ref)
        print(f"  prec={prec} rel rms {err:.2e}")
        assert err < tol


@pytest.mark.parametrize("case", sorted(MESHES))
def test_plane_spreading_equals_tiles(case, monkeypatch):
    """Spreading inside the forward plane kernel (a workgroup per z-plane
    over the charges of the order's first z indices) gives bitwise the
    Coulomb forces and energies of the tile spreading, on odd lengths,
    prime factors and an order per axis."""
    print(f"[plane spreading] {case}", flush=True)
    name, ng, ordr = MESHES[case]
    p, pos, vel = _mesh_variant(name, ng, ordr)
    for prec in (MD_FP64, MD_FP32):
        q = p.copy()
        q.precision = prec
        out = []
        for flag in ("0", "1"):
            monkeypatch.setenv("MD_PLANE_SPREAD", flag)
            e = Engine(q)
            e.upload(pos, vel)
            e.build_lists()
            out.append((e.force_parts()[1], e.energies()))
            e.close()
        assert np.array_equal(out[0][0], out[1][0]), prec
        assert out[0][1] == out[1][1], prec


@pytest.mark.parametrize("prec", [MD_FP64, MD_FP32])
def test_reference_ewald_setup(prec):
    """The default brush with force_data.txt's own Ewald set-up (tol 1.9,
    tau_rf 800, order 6, no overrides): the reference's automatic rule gives
    alpha 0.2079, r_c 9.139 and the 45 x 45 x 162 mesh (compt_acc.f90:375-430),
    which the engine runs as it stands.  Forces against the oracle: FP64
    within 1e-10, FP32 within 1e-4; a Philox run follows the oracle."""
    print(f"[reference set-up] default 45x45x162 order 6 prec={prec}", flush=True)
    p, pos, vel = system("default", MD_FP64)
    q = p.copy()
    q.ovr_ng[0] = q.ovr_ng[1] = q.ovr_ng[2] = 0
    q.ovr_alpha = 0.0
    q.ovr_rcc = 0.0
    q.tol, q.tau_rf = 1.9, 800.0
    q.ordr[0] = q.ordr[1] = q.ordr[2] = 6
    from md_brushes_amd import derive
    q = derive(q)
    assert list(q.ng) == [45, 45, 162] and abs(q.alpha - 0.2079) < 5e-5 and q.real_verlet == 1
    o = orc.Oracle(q)
    o.set_state(pos, vel)
    o.build_lists()
    ref_s = o.lj_force() + o.fene_force()[0]
    ref_c = o.coulomb()
    q.precision = prec
    e = Engine(q)
    e.upload(pos, vel)
    e.build_lists()
    fs, fc = e.force_parts()
    if prec == MD_FP64:
        assert orc.rel_rms(fc, ref_c) < 1e-10
        e.compute_force(0, MD_NOISE_PHILOX)
        assert o.compute_force(0, MD_NOISE_PHILOX) == 0
        e.run(1, 20)
        assert o.run(1, 20) == 0
        assert np.abs(e.download()[0] - o.state()[0]).max() < 1e-8
    else:
        err = orc.rel_rms(fs + fc, ref_s + ref_c)
        print(f"fp32 relative force rms {err:.3e}")
        assert err < 1e-4
    e.close()
