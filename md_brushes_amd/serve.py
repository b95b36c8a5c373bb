"""Continuous-batching service for brush replica runs.

A request is one independent brush system (architecture, salt valence,
field strength, seed, initial configuration) to be advanced nsteps Langevin
steps; the answer is its brush height and z-density profile.  Requests that
arrive while others are running are admitted at the next multistep-cycle
boundary when a slot and device memory are free; all admitted replicas
advance together in one graph-replayed launch sequence per cycle, and leave
the engine as soon as they finish.  The reference has no serving mode: it
runs one system per process (main.f90).

Usage: python -m md_brushes_amd.serve --requests 256 --steps 10000
"""
import argparse
import json
import time

import numpy as np

from .abi import (MD_FP32, MD_FP64, MD_NOISE_PHILOX, Batch, Engine, initialize_position, initialize_velocity,
                  replica_bytes)
from .configs import replica


class Request:
    def __init__(self, rid, arrival, params, nsteps, pos, vel, desc):
        self.rid, self.arrival, self.params, self.nsteps = rid, arrival, params, nsteps
        self.pos, self.vel, self.desc = pos, vel, desc
        self.slot = None
        self.need = None
        self.start_step = 1
        self.done_steps = 0
        self.t_admit = self.t_done = None
        self.result = None


def make_stream(n, nsteps, rate, seed, precision=MD_FP32, max_side=8, preheat=None):
    """Seeded stream of n requests with Poisson arrivals of the given rate
    (1/s): chains 2x2..max_side^2, N uniform in 20..100, linear or 3/4-arm
    star, salt valence 1 or 3, E_z uniform in [0,1].  Each runs nsteps steps,
    the first `preheat` of them (default a fifth, a multiple of 20) in the
    reference's pre-heat form, from its random initial configuration."""
    if preheat is None:
        preheat = (nsteps // 5) // 20 * 20
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n):
        t += rng.exponential(1.0 / rate)
        side = int(rng.integers(2, max_side + 1))
        nb = int(rng.integers(20, 101))
        arch = int(rng.choice([1, 3, 4]))
        val = int(rng.choice([1, 3]))
        ez = float(rng.uniform(0, 1))
        rs = int(seed * 1000 + k)
        p = replica(side, nb, arch, val, ez, rs, precision, preheat)
        pos = initialize_position(p, rs)
        vel = initialize_velocity(p, rs + 1)
        out.append(Request(k, t, p, nsteps, pos, vel,
                           {"side": side, "N": nb, "arch": arch, "valence": val, "Ez": round(ez, 4),
                            "NN": p.NN}))
    return out


def footprint(req):
    """Device bytes the replica will hold: md_replica_bytes, the engine's own
    sizing of its buffers for these parameters and this configuration,
    computed on the host."""
    if getattr(req, "need", None) is None:
        req.need = replica_bytes(req.params, req.pos)
    return req.need


class Server:
    """Continuous batching: a request is admitted at a cycle boundary when a
    slot is free and its device bytes fit the HBM budget beside those of the
    running replicas; otherwise it waits in arrival order."""

    def __init__(self, precision=MD_FP32, multistep=10, hbm_budget=64 << 30, max_slots=256, chunk_cycles=10,
                 spot_check_every=0):
        self.batch = Batch(precision, multistep)
        # accuracy spot check (error_analysis on the device) of every n-th
        # request as it finishes: its rmsf against standard Ewald
        self.spot_every = spot_check_every
        self.multistep = multistep
        self.budget = hbm_budget
        self.max_slots = max_slots
        self.chunk = chunk_cycles
        self.active = []
        self.used = 0
        self.peak_used = 0
        self.peak_active = 0
        self.queued = 0  # admissions that had to wait for memory or a slot

    def admit(self, req, now):
        need = footprint(req)
        if need > self.budget:
            raise ValueError(f"request {req.rid} needs {need} device bytes, more than the budget {self.budget}")
        if len(self.active) >= self.max_slots or self.used + need > self.budget:
            return False
        req.slot = self.batch.add(req.params, req.pos, req.vel)
        req.t_admit = now
        self.used += need
        self.peak_used = max(self.peak_used, self.used)
        self.active.append(req)
        self.peak_active = max(self.peak_active, len(self.active))
        return True

    def serve(self, requests, clock=time.perf_counter):
        """Runs the stream; arrival times are offsets from the call.  Returns
        the finished requests (with result, latency) and the wall time."""
        pending = sorted(requests, key=lambda r: r.arrival)
        t0 = clock()
        done = []
        while pending or self.active:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                if not self.admit(pending[0], now):
                    if not getattr(pending[0], "waited", False):
                        pending[0].waited = True
                        self.queued += 1
                    break
                pending.pop(0)
            if not self.active:
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - now, 0.01)))
                continue
            # run up to the next completion, at most `chunk` cycles
            left = min(-(-(r.nsteps - r.done_steps) // self.multistep) for r in self.active)
            nc = max(1, min(self.chunk, left))
            self.batch.run(nc)
            for r in self.active:
                r.done_steps += nc * self.multistep
            finished = [r for r in self.active if r.done_steps >= r.nsteps]
            for r in finished:
                rc, step, msg = self.batch.status(r.slot)
                if rc == 0:
                    h, prof = self.batch.observe(r.slot)
                    r.result = {"height": h, "profile": prof, "status": 0}
                    if self.spot_every and r.rid % self.spot_every == 0:
                        r.result["rmsf"] = self.batch.error_analysis(r.slot)[0]["rmsf"]
                else:
                    r.result = {"status": rc, "message": msg}
                self.batch.remove(r.slot)
                self.used -= r.need
                r.t_done = clock() - t0
                self.active.remove(r)
                done.append(r)
            # every few chunks: evict failed replicas without disturbing the others
            self.chunks_run = getattr(self, "chunks_run", 0) + 1
            if self.chunks_run % 8 == 0:
                for r in list(self.active):
                    rc, _, msg = self.batch.status(r.slot)
                    if rc != 0:
                        r.result = {"status": rc, "message": msg}
                        self.batch.remove(r.slot)
                        self.used -= r.need
                        r.t_done = clock() - t0
                        self.active.remove(r)
                        done.append(r)
        return done, clock() - t0


def run_alone(req, precision=MD_FP32, multistep=10):
    """The same request served alone (for the alone-vs-mixed check)."""
    b = Batch(precision, multistep)
    slot = b.add(req.params, req.pos, req.vel)
    b.run(req.nsteps // multistep)
    h, prof = b.observe(slot)
    pos, vel, _ = b.download(slot)
    b.close()
    return h, prof, pos, vel


def sequential_fp64(requests, multistep=10, spot_check_every=0):
    """Comparator: each request alone on the plain FP64 path, one after the
    other; returns total steps / total seconds, the total seconds, and the
    rmsf of error_analysis at the end of every spot_check_every-th request
    (outside the timing)."""
    total_steps, total_t = 0, 0.0
    rmsf = []
    for r in requests:
        p = r.params.copy()
        p.precision = MD_FP64
        t0 = time.perf_counter()
        e = Engine(p)
        e.upload(r.pos, r.vel)
        e.build_lists()
        e.compute_force(0, MD_NOISE_PHILOX)
        pre = min(r.params.StepNum0, r.nsteps)
        if pre:
            e.run(1, pre, preheat=True)
        e.run(1 + pre, r.nsteps - pre)
        total_t += time.perf_counter() - t0
        total_steps += r.nsteps
        if spot_check_every and r.rid % spot_check_every == 0:
            rmsf.append(e.error_analysis()[0]["rmsf"])
        e.close()
    return total_steps / total_t, total_t, rmsf


def summarize(done, wall):
    lat = np.array([r.t_done - r.arrival for r in done])
    steps = sum(r.nsteps for r in done)
    ok = sum(1 for r in done if r.result and r.result["status"] == 0)
    out = {"requests": len(done), "ok": ok, "replica_steps": steps, "wall_s": round(wall, 3),
           "aggregate_steps_per_s": round(steps / wall, 1),
           "latency_p50_s": round(float(np.percentile(lat, 50)), 4),
           "latency_p99_s": round(float(np.percentile(lat, 99)), 4)}
    failed = [r for r in done if not r.result or r.result["status"] != 0]
    if failed:
        out["failures"] = [{"rid": r.rid, "status": r.result["status"] if r.result else None,
                            "message": r.result.get("message") if r.result else None, **r.desc}
                           for r in failed[:8]]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=256)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--rate", type=float, default=128.0)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--sequential", type=int, default=-1,
                    help="requests in the sequential FP64 comparator (-1: all)")
    ap.add_argument("--batch-only", type=int, default=0, metavar="CYCLES",
                    help="no serving loop: add every request at once and time CYCLES cycles of the batch "
                         "(a fixed workload for kernel traces)")
    args = ap.parse_args()
    reqs = make_stream(args.requests, args.steps, args.rate, args.seed)
    if args.batch_only:
        b = Batch(MD_FP32)
        for r in reqs:
            b.add(r.params, r.pos, r.vel)
        b.run(2)
        b.status(0)
        t0 = time.perf_counter()
        b.run(args.batch_only)
        b.status(0)  # waits for the queued cycles
        dt = time.perf_counter() - t0
        steps = len(reqs) * args.batch_only * b.multistep
        print(json.dumps({"replicas": len(reqs), "cycles": args.batch_only,
                          "replica_steps_per_s": round(steps / dt, 1)}))
        b.close()
        return
    srv = Server()
    done, wall = srv.serve(reqs)
    out = summarize(done, wall)
    n = len(reqs) if args.sequential < 0 else args.sequential
    sps, tseq, _ = sequential_fp64(reqs[:n])
    out["sequential_fp64_steps_per_s"] = round(sps, 1)
    out["speedup_vs_sequential"] = round(out["aggregate_steps_per_s"] / sps, 2)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
