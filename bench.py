"""Throughput of the force-and-step loop on one MI355X.

Workload: the default brush of MD_Brushes (8x8 linear chains of 100 charged
beads, counterions and trivalent salt: 21,396 particles, 21,332 charged),
SPME mesh 64x64x256 order 4, Coulomb every 10 steps, dt = 0.005; one step =
new_position + update_verlet_list (main.f90:92-94) with Philox noise, all
state resident in HBM.  Headline: FP32-force steps/s, reported beside the
plain FP64 path and the FP32-vs-FP64 relative force RMS on the same
configuration.  The CPU baseline is the FP64 oracle (a C restatement of the
reference, -O3) timed on this host for a bounded sample, on one thread and
with OpenMP on the host's threads.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from md_brushes_amd import (MD_FP32, MD_FP64, MD_NOISE_PHILOX, Engine, autotune_geometry,  # noqa: E402
                            config, initialize_position, initialize_velocity)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def step_bytes(p):
    """Algorithmic bytes of one step of the FP32 force path (compulsory
    traffic, each array touched once per pass; stored neighbour lists
    excluded): per particle and step 160 B -- the integrator reads x, v
    (FP64, 48 B) and the force (FP32, 12 B), writes x, v (48 B) and the FP32
    position copy (12 B); the short-range pass reads the FP32 position and
    charge (16 B) and the held Coulomb force (12 B) and writes the force
    (12 B) -- plus, once per Coulomb evaluation (every multistep steps),
    72 B per charge and 16 B per mesh point and per half-spectrum point.
    The FP64 path: 200 B per particle, 144 B per charge, 32 B per point."""
    NN, Nq = p.NN, p.Nq
    K1, K2, K3 = p.ng
    M, Mh = K1 * K2 * K3, (K1 // 2 + 1) * K2 * K3
    if p.precision == MD_FP32:
        return 160 * NN, 72 * Nq + 16 * M + 16 * Mh
    return 200 * NN, 144 * Nq + 32 * M + 32 * Mh


def kernel_bytes(p, name):
    """Algorithmic bytes of one launch of a step's kernel: the short-range
    kernel carries the whole per-particle part of the step (it kicks, drifts
    and stores the next positions), the Coulomb chain the per-evaluation
    part."""
    per_particle, per_coulomb = step_bytes(p)
    if name == "short_range":
        return per_particle
    if name == "coulomb":
        return per_coulomb
    return 0


def roofline(ms, plist, multistep):
    """HBM roofline of the kernel with the largest time per step (Coulomb
    kernels run every multistep steps; list kernels only on rebuilds; the
    kick-drift kernel once per run, replayed steps drift in short_range):
    algorithmic bytes of one launch over all replicas / its mean duration
    (HIP events on the engine's stream)."""
    per_step = {k: v * (1.0 / multistep if k in Engine.COULOMB else 1.0)
                for k, v in ms.items() if k not in Engine.LISTS and k not in Engine.RUN_START}
    dom = max(per_step, key=per_step.get)
    b = sum(kernel_bytes(p, dom) for p in plist)
    achieved = b / (ms[dom] * 1e-3)
    return {"bound": "hbm", "kernel": dom, "achieved": round(achieved / 1e9, 2), "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 5), "traffic": None, "bytes_per_launch": b,
            "avg_launch_us": round(ms[dom] * 1e3, 3)}


def roofline_chain(ms, plist):
    """The Coulomb evaluation as a whole (spreading, FFTs, gather, real
    space, store): its algorithmic bytes over the sum of its kernels' mean
    durations."""
    t = sum(ms[k] for k in Engine.COULOMB) * 1e-3
    b = sum(kernel_bytes(p, "coulomb") for p in plist)
    return {"bound": "hbm", "achieved": round(b / t / 1e9, 2), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": round(b / t / HBM_PEAK, 5), "bytes_per_evaluation": b, "us_per_evaluation": round(t * 1e6, 2)}


def roofline_step(p, ms_per_step):
    """The whole step: algorithmic bytes per step over the timed wall time
    per step."""
    a, c = step_bytes(p)
    b = a + c / p.multistep
    return {"bound": "hbm", "achieved": round(b / (ms_per_step * 1e-3) / 1e9, 2), "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": round(b / (ms_per_step * 1e-3) / HBM_PEAK, 5), "bytes_per_step": int(b)}


def prepare(seed, equil_steps):
    """Seeded default brush after FP64 equilibration in the reference's
    pre-heat form (velocities rescaled every 20 steps, main.f90:67-75): the
    random ion start carries a large slab dipole whose relaxation heats the
    system, which the reference's pre-heat absorbs."""
    p64 = config("default", MD_FP64, seed=seed)
    pos = initialize_position(p64, seed)
    vel = initialize_velocity(p64, seed + 1)
    e = Engine(p64)
    e.upload(pos, vel)
    e.build_lists()
    e.compute_force(0, MD_NOISE_PHILOX)
    if equil_steps:
        e.run(1, equil_steps, preheat=True)
    pos, vel, acc, _ = e.download()
    e.close()
    return pos, vel, acc


OBS_STEPS, OBS_EVERY, OBS_BLOCKS = 5000, 50, 10  # observable runs: length, sampling, blocks
ENERGY_TERMS = ("lj", "wall", "fene", "coulomb", "field", "potential", "kinetic", "virial_lj",
                "virial_real")


def observables(state, seed):
    """FP32 forces against FP64 per observable, on the equilibrated default
    brush.  Energies and virials at one configuration (relative difference
    per term), then two runs from the same state with the same Philox noise
    stream, FP64 and FP32, sampled every OBS_EVERY steps: brush height (mean
    z of the polymer beads, in_output.f90:1017-1020), potential energy and
    the beads' z-density profile in 25 bins; each mean's standard error from
    OBS_BLOCKS block averages, and the FP32 - FP64 difference in units of
    the combined error (the runs decorrelate, so this is a statistical
    comparison)."""
    pos, vel, acc = state
    en = {}
    for prec in (MD_FP64, MD_FP32):
        e = Engine(config("default", prec, seed=seed))
        e.upload(pos, vel, acc)
        e.build_lists()
        en[prec] = e.energies()
        e.close()
    same = {k: float(abs(en[MD_FP32][k] - en[MD_FP64][k]) / max(abs(en[MD_FP64][k]), 1e-300))
            for k in ENERGY_TERMS}
    series = {}
    for prec in (MD_FP64, MD_FP32):
        p = config("default", prec, seed=seed)
        e = Engine(p)
        e.upload(pos, vel, acc)
        e.build_lists()
        e.compute_force(0, MD_NOISE_PHILOX)
        edges = np.linspace(0, p.Lz, 26)
        h, prof, pot = [], [], []
        for s in range(1, OBS_STEPS + 1, OBS_EVERY):
            e.run(s, OBS_EVERY)
            z = e.download()[0][: p.Npe, 2]
            h.append(z.mean())
            prof.append(np.histogram(z, edges)[0])
            pot.append(e.energies()["potential"])
        e.close()
        series[prec] = (np.array(h), np.array(prof, float), np.array(pot))

    def mean_se(x):
        m = np.array([b.mean(axis=0) for b in np.array_split(x, OBS_BLOCKS)])
        return x.mean(axis=0), m.std(axis=0, ddof=1) / np.sqrt(OBS_BLOCKS)

    out = {"energies_same_state_rel_diff": same,
           "runs": f"{OBS_STEPS} steps each from the bench state, sampled every {OBS_EVERY}, "
                   f"{OBS_BLOCKS} blocks"}
    for k, name in ((0, "h_avg"), (2, "potential")):
        m64, s64 = mean_se(series[MD_FP64][k])
        m32, s32 = mean_se(series[MD_FP32][k])
        se = float(np.hypot(s64, s32))
        out[name] = {"fp64": float(m64), "fp32": float(m32), "se": se,
                     "z": float((m32 - m64) / se) if se > 0 else 0.0}
    m64, s64 = mean_se(series[MD_FP64][1])
    m32, s32 = mean_se(series[MD_FP32][1])
    se = np.hypot(s64, s32)
    sel = se > 0
    zz = (m32[sel] - m64[sel]) / se[sel]
    out["z_profile"] = {"bins": int(sel.sum()), "chi2_per_bin": float(np.mean(zz ** 2)) if sel.any() else 0.0,
                        "max_abs_z": float(np.abs(zz).max()) if sel.any() else 0.0}
    return out


TUNE_CYCLES = 50  # cycles timed per autotuner candidate


def geometry(p):
    """The engine geometry of derived parameters: spreading tile edges (0 =
    16), LJ skin (0 = the reference's 0.5), real-space cell scale (0 = 1)."""
    return {"tile": list(p.ovr_tile), "skin": p.ovr_skin, "cell_scale": p.ovr_cell_scale,
            "rvl": round(p.rvl_eff, 4)}


WARMUP_MIN = 100  # the first replays of a freshly instantiated cycle graph run slower


def aligned_warmup(warmup, multistep):
    """Warm-up steps: at least WARMUP_MIN, rounded up to whole Coulomb cycles,
    so that the timed steps start at a step n with n % multistep == 1 and
    whole cycles of them are graph-replayed (md_run issues steps before a
    cycle boundary one by one)."""
    w = max(warmup, WARMUP_MIN)
    return -(-w // multistep) * multistep


def time_engine(prec, state, steps, warmup, seed, p=None):
    p = p if p is not None else config("default", prec, seed=seed)
    e = Engine(p)
    pos, vel, acc = state
    e.upload(pos, vel, acc)
    e.build_lists()
    e.compute_force(0, MD_NOISE_PHILOX)
    e.tuned = e.autotune(10)  # launch forms only: results do not depend on them
    w = aligned_warmup(warmup, p.multistep)
    e.run(1, w)
    t0 = time.perf_counter()
    e.run(1 + w, steps)  # ends with a stream synchronize
    dt = time.perf_counter() - t0
    return p, e, steps / dt, dt


COPY_BYTES = 256 << 20  # calibration copy for the HBM byte counters
SPOT_EVERY = 8  # accuracy spot check of every 8th served request


def pmc_workload(seed):
    """Kernels of the default brush (time_kernels launches each kernel of a
    step and of a Coulomb evaluation), then one device copy of COPY_BYTES,
    whose counter readings give the scale from counter units to bytes.  The
    pre-heat is shortened to 500 steps: counter collection serialises every
    dispatch, and the bytes per launch do not depend on how far the brush
    has relaxed."""
    import torch

    state = prepare(seed, 500)
    print("pmc workload: pre-heat done", flush=True)
    p, e, _, _ = time_engine(MD_FP32, state, 100, 10, seed)
    print("pmc workload: FP32 steps done", flush=True)
    e.time_kernels(10)
    e.close()
    a = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(f"pmc workload done: copy of {COPY_BYTES} bytes", flush=True)


KERNEL_SYMBOLS = {  # kernel names of time_kernels -> device symbols (FP32 path)
    "kick_drift": "k_kick_drift", "short_range": "k_short_range<float, true>", "bin_q": "k_bin_q",
    "sort_q": "k_sort_q<float>", "spread": "k_spread<float>", "fft_x_fwd": "k_plane_fwd<float>",
    "plane_spread": "k_plane_spread<float>",
    "fft_z_conv": "k_fftz_conv<float>", "fft_x_inv": "k_plane_inv<float>", "gather": "k_gather<float>",
    "real": ("k_real_cells<float", "k_coul_store<float>"), "bin_lj": "k_bin_lj", "build_lj": "k_build_lj"}


def _counter_rows(path, counter):
    import csv

    rows = []
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            if row.get("Counter_Name") == counter:
                rows.append((row["Kernel_Name"], float(row["Counter_Value"])))
    return rows


def traffic_summary(fetch_csv, write_csv):
    """Per-kernel HBM bytes per launch from a FETCH_SIZE pass and a WRITE_SIZE
    pass (KiB per dispatch), each scaled so that the plain copy of
    COPY_BYTES reads COPY_BYTES fetched and COPY_BYTES written."""
    fetch, write = _counter_rows(fetch_csv, "FETCH_SIZE"), _counter_rows(write_csv, "WRITE_SIZE")

    def copy_reading(rows):  # the copy: largest reading among kernels outside this library
        return max(v for n, v in rows if "mdk::" not in n)

    fcopy, wcopy = copy_reading(fetch), copy_reading(write)
    sf, sw = COPY_BYTES / (fcopy * 1024.0), COPY_BYTES / (wcopy * 1024.0)
    out = {"copy_bytes": COPY_BYTES, "copy_fetch_kib": fcopy, "copy_write_kib": wcopy,
           "fetch_scale": sf, "write_scale": sw, "kernels": {}}
    for name, syms in KERNEL_SYMBOLS.items():
        # a timed step may be several kernels (the sweep and the Coulomb store):
        # the sum of their per-dispatch means
        syms = (syms,) if isinstance(syms, str) else syms
        fk = wk = 0.0
        nd = None
        for sym in syms:
            fv = [v for n, v in fetch if sym in n]
            wv = [v for n, v in write if sym in n]
            if not fv or not wv:
                nd = 0
                break
            fk += sum(fv) / len(fv)
            wk += sum(wv) / len(wv)
            nd = len(fv) if nd is None else min(nd, len(fv))
        if not nd:
            continue
        out["kernels"][name] = {"dispatches": nd, "fetch_kib": round(fk, 1), "write_kib": round(wk, 1),
                                "bytes_per_launch": int(fk * 1024 * sf + wk * 1024 * sw)}
    path = os.path.join(ROOT, "profiles", "traffic_default_brush.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


def traffic_from_profile(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (two rocprofv3 --pmc passes, FETCH_SIZE and WRITE_SIZE, scaled by the
    readings of a plain copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_default_brush.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else k.get("bytes_per_launch")


def cpu_baseline(state, seed, budget_s, threads=1):
    """The FP64 oracle (a C restatement of the reference, -O3, OpenMP over
    its particle, pair-list, mesh and FFT-line loops) on this host with the
    given thread count: whole multistep cycles until about budget_s seconds
    have passed."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import orc  # the oracle: baseline only

    L = orc.fast_lib()
    L.orc_set_threads(threads)
    p = config("default", MD_FP64, seed=seed)
    o = orc.Oracle(p, L)
    pos, vel, acc = state
    o.set_state(pos, vel, acc)
    o.build_lists()
    o.compute_force(0, MD_NOISE_PHILOX)
    steps, t0 = 0, time.perf_counter()
    while True:
        rc = o.run(1 + steps, p.multistep)
        if rc:
            raise RuntimeError(f"oracle run failed: {rc}")
        steps += p.multistep
        el = time.perf_counter() - t0
        if el > budget_s:
            break
    return steps / el, steps, el


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """Threads the host offers this process (OMP_NUM_THREADS where set)."""
    try:
        return max(1, int(os.environ.get("OMP_NUM_THREADS", "")))
    except ValueError:
        return max(1, len(os.sched_getaffinity(0)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--equil", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-obs", action="store_true", help="skip the FP32-vs-FP64 observable runs")
    ap.add_argument("--no-geometry", action="store_true",
                    help="the reference's geometry without the geometry autotuner (profiling runs)")
    ap.add_argument("--no-ewald", action="store_true", help="skip the standard-Ewald accuracy check")
    ap.add_argument("--serve", type=int, default=256, help="requests in the serving leg (0: skip)")
    ap.add_argument("--serve-steps", type=int, default=10000)
    ap.add_argument("--serve-rate", type=float, default=128.0)
    ap.add_argument("--serve-seq", type=int, default=256, help="requests in the sequential FP64 comparator")
    ap.add_argument("--pmc-workload", action="store_true",
                    help="short run for hardware-counter passes: a few hundred steps of the default brush "
                         "and a plain device copy of known size (counter calibration); prints no JSON")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="counter_collection.csv files of the FETCH_SIZE and WRITE_SIZE passes over "
                         "--pmc-workload: writes profiles/traffic_default_brush.json")
    args = ap.parse_args()
    if args.traffic_from:
        traffic_summary(*args.traffic_from)
        return
    if args.pmc_workload:
        pmc_workload(args.seed)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (replicas shard "
              "across GPUs as independent engines, one process per GPU, with no communication)",
              file=sys.stderr)
        sys.exit(2)

    state = prepare(args.seed, args.equil)
    # engine geometry by the autotuner: spreading tiles only (results unchanged)
    geo32, tuner32 = config("default", MD_FP32, seed=args.seed), None
    if not args.no_geometry:
        geo32, tuner32 = autotune_geometry(geo32, state[0], state[1], cycles=TUNE_CYCLES)
    p32, e32, sps32, dt32 = time_engine(MD_FP32, state, args.steps, args.warmup, args.seed, geo32)
    out = {
        "metric": "steps/s at 21,333 particles (FP32 vs FP64, force err<=1e-4); replica p50/p99",
        "value": round(sps32, 2),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_run": aligned_warmup(args.warmup, p32.multistep),
        "ms_per_step": round(1e3 * dt32 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 forces, fp64 state",
        "data": f"synthetic: seeded brush, {args.equil} FP64 pre-heat steps",
        "config": {"workload": "default brush 8x8 linear N=100 + trivalent salt, NN=21396, Nq=21332, "
                               "SPME 64x64x256 p=4, multistep 10",
                   "NN": p32.NN, "Nq": p32.Nq, "mesh": list(p32.ng), "order": p32.ordr[0],
                   "alpha": round(p32.alpha, 5), "r_c": round(p32.rcc, 4), "multistep": p32.multistep,
                   "geometry": geometry(geo32)},
    }
    # accuracy gate on one configuration: FP32 forces against FP64 forces
    pos, vel, acc = state
    e64 = Engine(config("default", MD_FP64, seed=args.seed))
    e64.upload(pos, vel, acc)
    e64.build_lists()
    fs64, fc64 = e64.force_parts()
    e32b = Engine(p32)
    e32b.upload(pos, vel, acc)
    e32b.build_lists()
    fs32, fc32 = e32b.force_parts()
    f64, f32 = fs64 + fc64, fs32 + fc32
    out["fp32_force_rel_rms"] = float(np.sqrt(((f32 - f64) ** 2).sum() / (f64 ** 2).sum()))
    # the Coulomb force against standard Ewald (error_analysis, on the device):
    # the SPME set-up's own accuracy, FP64 and FP32
    if not args.no_ewald:
        out["coulomb_rmsf_vs_ewald"] = {"fp64": e64.error_analysis()[0]["rmsf"],
                                        "fp32": e32b.error_analysis()[0]["rmsf"]}
    e32b.close()
    e64.close()
    if not args.no_obs:
        out["fp32_observables"] = observables(state, args.seed)
    if not args.no_fp64:
        n64 = min(args.steps, 4000)
        geo64 = config("default", MD_FP64, seed=args.seed)
        if not args.no_geometry:
            geo64, _ = autotune_geometry(geo64, state[0], state[1], cycles=TUNE_CYCLES)
        _, e64t, sps64, _ = time_engine(MD_FP64, state, n64, args.warmup, args.seed, geo64)
        out["fp64_steps_per_s"] = round(sps64, 2)
        e64t.close()

    # per-kernel durations (HIP events on the engine's stream)
    ms = e32.time_kernels(20)
    out["config"]["launch_forms"] = e32.tuned
    out["roofline"] = roofline(ms, [p32], p32.multistep)
    out["roofline"]["traffic"] = traffic_from_profile(out["roofline"]["kernel"])
    out["roofline_coulomb"] = roofline_chain(ms, [p32])
    out["roofline_step"] = roofline_step(p32, out["ms_per_step"])
    out["kernel_us"] = {k: round(v * 1e3, 3) for k, v in ms.items()}
    e32.close()
    # the autotuner with the list knobs too (Verlet skin, sweep cell size): the
    # same pairs in another summation order, so reported beside the headline
    if not args.no_geometry:
        geo_l, tuner_l = autotune_geometry(config("default", MD_FP32, seed=args.seed), state[0], state[1],
                                           cycles=TUNE_CYCLES, lists=True)
        out["tuned_lists"] = dict(geometry(geo_l), tuner_steps_per_s=round(tuner_l, 1),
                                  tuner_steps_per_s_tiles_only=round(tuner32, 1),
                                  note="md_autotune_geometry rates over 50 cycles after 10 warm-up cycles; "
                                       "the list knobs change the summation order (forces to rounding)")

    if args.serve > 0:
        from md_brushes_amd.serve import Server, make_stream, sequential_fp64, summarize

        reqs = make_stream(args.serve, args.serve_steps, rate=args.serve_rate, seed=5)
        srv = Server(spot_check_every=SPOT_EVERY)
        done, wall = srv.serve(reqs)
        sv = summarize(done, wall)
        sv["admission"] = {"hbm_budget_bytes": srv.budget, "peak_bytes": srv.peak_used,
                           "peak_replicas": srv.peak_active, "queued": srv.queued}
        nseq = min(len(reqs), args.serve_seq)
        sps_seq, _, rms64 = sequential_fp64(reqs[:nseq], spot_check_every=SPOT_EVERY)
        sv["sequential_fp64_steps_per_s"] = round(sps_seq, 1)
        sv["speedup_vs_sequential"] = round(sv["aggregate_steps_per_s"] / sps_seq, 2)
        rms32 = [r.result["rmsf"] for r in done if r.result and "rmsf" in r.result]
        # error_analysis (rmsf of the Coulomb force against standard Ewald) of
        # every SPOT_EVERY-th request at its end: served FP32 beside the FP64 runs
        sv["accuracy_rmsf"] = {"sample": len(rms32),
                               "served_fp32_median": float(np.median(rms32)) if rms32 else None,
                               "served_fp32_max": float(np.max(rms32)) if rms32 else None,
                               "sequential_fp64_median": float(np.median(rms64)) if rms64 else None,
                               "sequential_fp64_max": float(np.max(rms64)) if rms64 else None}
        sv["stream"] = (f"{args.serve} requests, Poisson rate {args.serve_rate}/s, {args.serve_steps} steps "
                        f"each, chains 2x2..8x8, N 20..100, linear/3-/4-arm stars, salt valence 1/3, "
                        f"Ez in [0,1], FP32 forces; sequential comparator on {nseq}; admission on "
                        f"md_replica_bytes against the budget")
        out["serving"] = sv
        out["replica_latency_p50_s"] = sv["latency_p50_s"]
        out["replica_latency_p99_s"] = sv["latency_p99_s"]
        # the same kernels at batch scale: every request of the stream at once
        from md_brushes_amd import Batch
        bt = Batch(MD_FP32)
        for r in reqs:
            bt.add(r.params, r.pos, r.vel)
        bt.run(2)
        msb = bt.time_kernels(10)
        bt.close()
        rb = roofline(msb, [r.params for r in reqs], reqs[0].params.multistep)
        rb["replicas"] = len(reqs)
        rb["particles"] = int(sum(r.params.NN for r in reqs))
        out["roofline_batch"] = rb
        out["roofline_batch_coulomb"] = roofline_chain(msb, [r.params for r in reqs])
        out["kernel_us_batch"] = {k: round(v * 1e3, 3) for k, v in msb.items()}

    if not args.no_cpu:
        sps_cpu, n, el = cpu_baseline(state, args.seed, args.cpu_seconds, 1)
        out["cpu_baseline"] = {"value": round(sps_cpu, 3), "unit": "steps/s", "cores": 1, "kind": "port",
                               "cpu": cpu_model(), "host_cpus": os.cpu_count(),
                               "sample": f"{n} steps of the same workload (FP64 oracle, C, -O3 "
                                         f"-march=x86-64-v3, one thread) in {el:.1f} s"}
        nt = host_threads()
        sps_omp, n2, el2 = cpu_baseline(state, args.seed, args.cpu_seconds, nt)
        out["cpu_baseline_omp"] = {"value": round(sps_omp, 3), "unit": "steps/s", "cores": nt, "kind": "port",
                                   "sample": f"{n2} steps of the same workload (the same oracle, OpenMP "
                                             f"over {nt} threads) in {el2:.1f} s"}
        out["speedup_vs_cpu"] = round(sps32 / sps_cpu, 1)
        out["speedup_vs_cpu_omp"] = round(sps32 / sps_omp, 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
