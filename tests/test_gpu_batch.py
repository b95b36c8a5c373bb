"""Replica batches and the serving loop on the GPU: a replica's result is the
same alone and in a mix (bitwise: no cross-replica reductions and
order-independent kernels), a batch replica equals the single-system
engine, observables match the host-side definitions, and a failing replica
is reported without disturbing the others."""
import numpy as np
import pytest

from md_brushes_amd import MD_FP32, MD_FP64, MD_NOISE_PHILOX, Batch, Engine, initialize_position, initialize_velocity
from md_brushes_amd.configs import replica
from md_brushes_amd.serve import Server, make_stream, run_alone

pytestmark = pytest.mark.gpu


def _req(side, nb, arch, val, ez, seed, prec=MD_FP32):
    p = replica(side, nb, arch, val, ez, seed, prec)
    return p, initialize_position(p, seed), initialize_velocity(p, seed + 1)


def test_alone_equals_mixed():
    print("[batch] alone vs mixed", flush=True)
    target = _req(3, 40, 3, 3, 0.4, 11)
    others = [_req(2, 25, 1, 1, 0.9, 12), _req(4, 60, 4, 3, 0.1, 13), _req(2, 80, 1, 3, 0.5, 14)]
    b = Batch(MD_FP32)
    s_alone = b.add(*target)
    b.run(12)
    pa = b.download(s_alone)[0]
    b.close()
    b = Batch(MD_FP32)
    b.add(*others[0])
    b.run(3)
    s_mix = b.add(*target)  # joins while another replica is mid-run
    b.add(*others[1])
    b.run(5)
    b.add(*others[2])
    b.run(7)
    pm = b.download(s_mix)[0]
    b.close()
    assert np.array_equal(pa, pm)


def test_batch_replica_equals_single_engine():
    print("[batch] vs single engine fp64", flush=True)
    p, pos, vel = _req(3, 30, 1, 3, 0.3, 21, MD_FP64)
    e = Engine(p)
    e.upload(pos, vel)
    e.build_lists()
    e.compute_force(0, MD_NOISE_PHILOX)
    e.run(1, 80)
    pe, ve, _, _ = e.download()
    e.close()
    b = Batch(MD_FP64)
    sl = b.add(p, pos, vel)
    b.add(*_req(2, 20, 4, 1, 0.7, 22, MD_FP64))
    b.run(8)
    pb, vb, _ = b.download(sl)
    rc, step, _ = b.status(sl)
    b.close()
    assert rc == 0 and step == 81
    assert np.array_equal(pe, pb) and np.array_equal(ve, vb)


def test_batch_preheat_equals_single_engine():
    """A batch replica with StepNum0 = 60 pre-heats on its own step count,
    beside a replica without pre-heat, exactly as md_run's pre-heat form."""
    print("[batch] preheat vs single engine fp64", flush=True)
    p = replica(3, 30, 1, 3, 0.3, 51, MD_FP64, preheat=60)
    pos, vel = initialize_position(p, 51), initialize_velocity(p, 52)
    e = Engine(p)
    e.upload(pos, vel)
    e.build_lists()
    e.compute_force(0, MD_NOISE_PHILOX)
    e.run(1, 60, preheat=True)
    e.run(61, 40)
    pe, ve, _, _ = e.download()
    e.close()
    b = Batch(MD_FP64)
    b.add(*_req(2, 20, 4, 1, 0.7, 53, MD_FP64))
    b.run(2)
    sl = b.add(p, pos, vel)  # joins two cycles later: its own step count
    b.run(10)
    pb, vb, _ = b.download(sl)
    b.close()
    assert np.array_equal(pe, pb) and np.array_equal(ve, vb)


def test_observables_match_definitions():
    print("[batch] observables", flush=True)
    p, pos, vel = _req(4, 50, 1, 1, 0.2, 31)
    b = Batch(MD_FP32)
    sl = b.add(p, pos, vel)
    b.run(5)
    h, prof = b.observe(sl)
    ps = b.download(sl)[0]
    b.close()
    z = ps[: p.Npe, 2]
    assert abs(h - z.mean()) < 1e-12 * max(1.0, abs(h))
    k = np.ceil(z / (p.Lz / 500)).astype(int)
    ref = np.bincount(k[(k >= 1) & (k <= 500)] - 1, minlength=500)
    assert np.array_equal(prof, ref)


def test_failing_replica_is_isolated():
    print("[batch] failing replica", flush=True)
    good = _req(3, 30, 1, 3, 0.3, 41)
    p, pos, vel = _req(2, 30, 1, 1, 0.3, 42)
    bad = np.array(pos)
    bad[3, 2] += 3.0  # a bond past R0
    b = Batch(MD_FP32)
    sg = b.add(*good)
    with pytest.raises(Exception):
        b.add(p, bad, vel)  # its step-0 force reports the break
    sg2 = b.add(*_req(2, 20, 1, 1, 0.6, 43))
    b.run(4)
    assert b.status(sg)[0] == 0 and b.status(sg2)[0] == 0
    ref = run_alone(_Fake(good, 40))[2]
    assert np.array_equal(b.download(sg)[0], ref)
    b.close()


class _Fake:
    def __init__(self, r, n):
        self.params, self.pos, self.vel = r
        self.nsteps = n


def test_serving_stream_results_match_alone():
    print("[serve] small stream", flush=True)
    reqs = make_stream(8, 200, rate=400.0, seed=3, max_side=3)
    srv = Server(chunk_cycles=4)
    done, wall = srv.serve(reqs)
    assert len(done) == 8 and all(r.result["status"] == 0 for r in done)
    for r in done[:3]:
        h, prof, _, _ = run_alone(r)
        assert h == r.result["height"]
        assert np.array_equal(prof, r.result["profile"])
    lat = [r.t_done - r.arrival for r in done]
    assert min(lat) > 0


def test_batch_error_analysis_spot_check():
    """The accuracy spot check of a served replica equals the single-system
    engine's and leaves every replica's trajectory unchanged."""
    print("[batch] error_analysis spot check", flush=True)
    p, pos, vel = _req(3, 30, 1, 3, 0.3, 31, MD_FP64)
    other = _req(2, 40, 3, 1, 0.6, 32, MD_FP64)
    e = Engine(p)
    e.upload(pos, vel)
    e.build_lists()
    ge, fe = e.error_analysis()
    e.close()
    out = []
    for check in (True, False):
        b = Batch(MD_FP64)
        b.add(*other)
        sl = b.add(p, pos, vel)
        if check:
            gb, fb = b.error_analysis(sl)
            assert gb == ge
            assert np.array_equal(fb, fe)
        b.run(2)
        if check:
            b.error_analysis(sl)
        b.run(2)
        out.append(b.download(sl)[0])
        b.close()
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("name", ["default", "star3", "mixed_small"])
def test_heights_match_definitions(name):
    """md_heights / md_batch_heights equal write_height's definitions
    (in_output.f90:1017-1079) evaluated on the downloaded positions."""
    import sys, os
    sys.path.insert(0, os.path.dirname(__file__))
    from systems import heights_ref, system
    print(f"[heights] {name}", flush=True)
    if name == "mixed_small":  # stars and chains together
        p, pos, vel = _req(4, 40, 3, 3, 0.2, 41, MD_FP64)
        p2 = p.copy()
        from md_brushes_amd.configs import _apply
        from md_brushes_amd import derive
        _apply(p2, [("Ngl", 8), ("Nml", 21)])
        p = derive(p2)
        pos, vel = initialize_position(p, 41), initialize_velocity(p, 42)
    else:
        p, pos, vel = system(name)
    e = Engine(p)
    e.upload(pos, vel)
    e.build_lists()
    e.compute_force(0, MD_NOISE_PHILOX)
    e.run(1, 20)
    got = e.heights()
    pg = e.download()[0]
    e.close()
    ref = heights_ref(pg, p)
    for k, v in ref.items():
        assert abs(got[k] - v) <= 1e-12 * max(1.0, abs(v)), (k, got[k], v)
    b = Batch(p.precision)
    sl = b.add(p, pg, np.zeros_like(vel))
    hb = b.heights(sl)
    b.close()
    for k, v in ref.items():
        assert abs(hb[k] - v) <= 1e-12 * max(1.0, abs(v)), (k, hb[k], v)


def test_replica_bytes_match_allocation():
    """md_replica_bytes (host-side sizing) equals the bytes md_batch_add
    allocates, for linear and star replicas and a small box that walks the
    reference-form real-space list."""
    print("[serve] replica bytes", flush=True)
    from md_brushes_amd import replica_bytes
    b = Batch(MD_FP32)
    for args in ((2, 30, 1, 3, 0.5, 61), (5, 70, 3, 1, 0.2, 62), (8, 100, 1, 3, 0.5, 63)):
        p, pos, vel = _req(*args)
        want = replica_bytes(p, pos)
        sl = b.add(p, pos, vel)
        got, _ = b.info(sl)
        assert got == want, (args, got, want)
    b.close()


def test_admission_queues_on_device_bytes():
    """With an HBM budget that holds only a few replicas, requests that
    arrive together wait for memory (admitted as others finish), the bytes
    in use never exceed the budget, and every result equals the request
    run alone."""
    print("[serve] tiny HBM budget", flush=True)
    from md_brushes_amd.serve import footprint
    reqs = make_stream(10, 100, rate=1e6, seed=9, max_side=4)
    need = sorted(footprint(r) for r in reqs)
    budget = need[-1] + need[-2] + need[-3]
    srv = Server(hbm_budget=budget, chunk_cycles=2)
    done, wall = srv.serve(reqs)
    assert len(done) == 10 and all(r.result["status"] == 0 for r in done)
    assert srv.queued > 0 and srv.peak_used <= budget and srv.peak_active < 10
    assert max(r.t_admit for r in done) > min(r.t_done for r in done)  # admitted after a completion
    for r in done:
        h, prof, _, _ = run_alone(r)
        assert h == r.result["height"]
        assert np.array_equal(prof, r.result["profile"])
