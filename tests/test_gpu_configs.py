"""The benchmark workloads at full size on the GPU: the mixed large brush
(16x16 sites, 4-arm stars and chains of 100, 85,588 particles, SPME
128x128x512 order 6, Lz = 300), the 3-arm star brush over the field sweep
E_z = 0..1, and the 256-request serving stream of 10^4 steps each."""
import numpy as np
import pytest

import orc
from md_brushes_amd import MD_FP32, MD_FP64, MD_NOISE_PHILOX, Engine, autotune_geometry
from md_brushes_amd.serve import Server, make_stream, run_alone, summarize
from systems import system

pytestmark = pytest.mark.gpu


def _parity(p, pos, vel, lists=True):
    o = orc.Oracle(p)
    o.set_state(pos, vel)
    o.build_lists()
    ref_s = o.lj_force() + o.fene_force()[0]
    ref_c = o.coulomb()
    e = Engine(p)
    e.upload(pos, vel)
    e.build_lists()
    if lists:
        for which in (0, 1):
            if which == 1 and not p.real_verlet:
                continue
            r = o.lists(which)
            r = r[np.lexsort((r[:, 1], r[:, 0]))]
            assert np.array_equal(e.lists(which), r), which
    fs, fc = e.force_parts()
    e.close()
    q = p.copy()
    q.precision = MD_FP32
    e = Engine(q)
    e.upload(pos, vel)
    e.build_lists()
    gs, gc = e.force_parts()
    e.close()
    return (orc.rel_rms(fs, ref_s), orc.rel_rms(fc, ref_c), orc.rel_rms(gs + gc, ref_s + ref_c))


def test_mixed_large_brush():
    """Lists bit-exact, FP64 forces against the oracle, the FP32 gate, and
    a thousand FP32 steps without error at 85,588 particles."""
    print("[config] mixed large brush 85,588, 128x128x512 order 6", flush=True)
    p, pos, vel = system("mixed")
    assert p.NN == 85588 and list(p.ng) == [128, 128, 512] and p.ordr[0] == 6
    es, ec, e32 = _parity(p, pos, vel)
    print(f"  fp64 short {es:.1e} coulomb {ec:.1e}; fp32 {e32:.2e}")
    assert es < 1e-11 and ec < 1e-10 and e32 < 1e-4
    q = p.copy()
    q.precision = MD_FP32
    e = Engine(q)
    e.upload(pos, vel)
    e.build_lists()
    e.compute_force(0, MD_NOISE_PHILOX)
    e.run(1, 1000, preheat=True)
    x = e.download()[0]
    e.close()
    assert np.all(np.isfinite(x))


def test_mixed_large_brush_autotuned():
    """The autotuner picks the spreading tiles, the Verlet skin and the
    real-space cell size for the mixed large brush (after 500 pre-heat
    steps); at its choice the lists are bit-exact with the oracle's at the
    same parameters, FP64 forces match, and the FP32 gate holds."""
    print("[config] mixed large brush, autotuned geometry", flush=True)
    p, pos, vel = system("mixed")
    e = Engine(p)
    e.upload(pos, vel)
    e.build_lists()
    e.compute_force(0, MD_NOISE_PHILOX)
    e.run(1, 500, preheat=True)
    x, v = e.download()[:2]
    e.close()
    best, sps = autotune_geometry(p, x, v, cycles=10, lists=True)
    print(f"  tile {list(best.ovr_tile)} skin {best.ovr_skin} cell {best.ovr_cell_scale}: {sps:.0f} steps/s")
    assert sps > 0
    es, ec, e32 = _parity(best, x, v)
    print(f"  fp64 short {es:.1e} coulomb {ec:.1e}; fp32 {e32:.2e}")
    assert es < 1e-11 and ec < 1e-10 and e32 < 1e-4


@pytest.mark.parametrize("ez", [0.0, 0.25, 0.5, 0.75, 1.0])
def test_star_brush_field_sweep(ez):
    """The 3-arm star brush (64 stars of 3 x 33 beads, trivalent salt) at
    each field strength: lists bit-exact, FP64 forces within 1e-10 of the
    oracle, FP32 within 1e-4, and the field force q E_z in the step."""
    print(f"[config] star3 E_z = {ez}", flush=True)
    p, pos, vel = system("star3", MD_FP64, ez=ez)
    assert abs(p.EF - ez) < 1e-15
    es, ec, e32 = _parity(p, pos, vel, lists=ez == 0.5)
    print(f"  fp64 short {es:.1e} coulomb {ec:.1e}; fp32 {e32:.2e}")
    assert es < 1e-11 and ec < 1e-10 and e32 < 1e-4
    # the field enters the step's force: compute_force without noise
    o = orc.Oracle(p)
    o.set_state(pos, vel)
    o.build_lists()
    assert o.compute_force(0) == 0
    e = Engine(p)
    e.upload(pos, vel)
    e.build_lists()
    e.compute_force(0)
    assert orc.rel_rms(e.download()[2], o.state()[2]) < 1e-10
    e.close()


def test_serving_stream_full_size():
    """256 requests with Poisson arrivals, 10^4 steps each: every status 0,
    and a sample of the served results equals the same requests run alone."""
    print("[config] serving stream 256 x 10^4", flush=True)
    reqs = make_stream(256, 10000, rate=128.0, seed=5)
    srv = Server()
    done, wall = srv.serve(reqs)
    sv = summarize(done, wall)
    print(f"  {sv['aggregate_steps_per_s']:.0f} replica-steps/s, p50 {sv['latency_p50_s']} s, "
          f"p99 {sv['latency_p99_s']} s")
    assert len(done) == 256 and all(r.result["status"] == 0 for r in done)
    rng = np.random.default_rng(1)
    for k in rng.choice(256, 4, replace=False):
        r = next(x for x in done if x.rid == k)
        h, prof, _, _ = run_alone(r)
        assert h == r.result["height"]
        assert np.array_equal(prof, r.result["profile"])
