"""Engine geometry overrides and their autotuner, on the GPU, through the
C-ABI: spreading tile edges (results bitwise unchanged), the LJ Verlet skin
(lists bit-exact with the oracle's at the same skin) and the real-space
sweep's cell scale (forces against the oracle), and md_autotune_geometry's
choice."""
import numpy as np
import pytest

import orc
from md_brushes_amd import MD_FP32, MD_FP64, MD_NOISE_PHILOX, Engine, autotune_geometry, derive
from systems import system

pytestmark = pytest.mark.gpu


def _with(p, **kw):
    q = p.copy()
    for k, v in kw.items():
        if k == "ovr_tile":
            q.ovr_tile[0], q.ovr_tile[1], q.ovr_tile[2] = v
        else:
            setattr(q, k, v)
    return derive(q)


def relaxed(name, steps=400, prec=MD_FP64):
    """(params, pos, vel) after a pre-heat run from the seeded start (the
    tuner times production steps, which a fresh start may not survive)."""
    p, pos, vel = system(name, prec)
    e = Engine(p)
    e.upload(pos, vel)
    e.build_lists()
    e.compute_force(0, MD_NOISE_PHILOX)
    e.run(1, steps, preheat=True)
    x, v = e.download()[:2]
    e.close()
    return p, x, v


def _coulomb(p, pos, vel):
    e = Engine(p)
    e.upload(pos, vel)
    e.build_lists()
    fs, fc = e.force_parts()
    e.close()
    return fs, fc


@pytest.mark.parametrize("prec", [MD_FP64, MD_FP32])
def test_tile_edges_do_not_change_results(prec):
    """Spreading tiles of 8 or 4 points per axis (more, smaller workgroups)
    give bitwise the Coulomb forces of the 16-point tiles: the spread sums
    are exact fixed point.  Also on the reference's 45 x 45 x 162 order-6
    mesh, whose tiles are uneven (an edge of 4 falls back to the order)."""
    print(f"[tile edges] default prec={prec}", flush=True)
    p, pos, vel = system("default", prec)
    ref = _coulomb(p, pos, vel)[1]
    for tile in ((8, 8, 8), (16, 16, 8), (4, 4, 4), (8, 16, 4)):
        f = _coulomb(_with(p, ovr_tile=tile), pos, vel)[1]
        assert np.array_equal(f, ref), tile
    q = p.copy()
    q.ovr_ng[0] = q.ovr_ng[1] = q.ovr_ng[2] = 0
    q.ovr_alpha = q.ovr_rcc = 0.0
    q.tol, q.tau_rf = 1.9, 800.0
    q.ordr[0] = q.ordr[1] = q.ordr[2] = 6
    q = derive(q)
    assert list(q.ng) == [45, 45, 162]
    ref = _coulomb(q, pos, vel)[1]
    for tile in ((8, 8, 8), (4, 4, 4)):
        assert np.array_equal(_coulomb(_with(q, ovr_tile=tile), pos, vel)[1], ref), tile


@pytest.mark.parametrize("skin", [0.75, 1.0])
def test_skin_lists_and_forces_match_oracle(skin):
    """A wider LJ skin: the engine's and the oracle's lists at that skin are
    the same pair set (bit-exact), the FP64 short-range force matches the
    oracle's within 1e-11, and it equals the reference-skin force to
    rounding (the extra pairs lie beyond every cut)."""
    print(f"[skin] default skin={skin}", flush=True)
    p, pos, vel = system("default", MD_FP64)
    q = _with(p, ovr_skin=skin)
    assert abs(q.rvl_eff - (q.rcl_s + 2 * skin)) < 1e-15
    o = orc.Oracle(q)
    o.set_state(pos, vel)
    o.build_lists()
    e = Engine(q)
    e.upload(pos, vel)
    e.build_lists()
    r = o.lists(0)
    r = r[np.lexsort((r[:, 1], r[:, 0]))]
    g = e.lists(0)
    assert g.shape[0] > 0 and np.array_equal(g, r)
    fs = e.force_parts()[0]
    e.close()
    ref = o.lj_force() + o.fene_force()[0]
    assert orc.rel_rms(fs, ref) < 1e-11
    fs0 = _coulomb(p, pos, vel)[0]
    assert orc.rel_rms(fs, fs0) < 1e-13


@pytest.mark.parametrize("scale", [1.25, 1.5])
def test_cell_scale_forces_match_oracle(scale):
    """Real-space sweep cells of edge >= scale * r_c: the same pairs in
    another order, so the Coulomb force matches the oracle within 1e-10
    (FP64) and the FP32 gate."""
    print(f"[cell scale] default scale={scale}", flush=True)
    p, pos, vel = system("default", MD_FP64)
    o = orc.Oracle(p)
    o.set_state(pos, vel)
    o.build_lists()
    ref = o.coulomb()
    for prec, tol in ((MD_FP64, 1e-10), (MD_FP32, 1e-4)):
        q = _with(p, ovr_cell_scale=scale, precision=prec)
        assert orc.rel_rms(_coulomb(q, pos, vel)[1], ref) < tol, prec


def test_autotune_geometry_choice_runs_and_matches_oracle():
    """md_autotune_geometry on the small brush returns derived parameters
    with the chosen overrides (tiles within 4..16, skin 0 or one of the
    candidates, cell scale 0 or a candidate) and its measured rate; an
    engine at that geometry follows the oracle (at the same parameters) on
    lists and forces, and a Philox run stays on it to 1e-8."""
    print("[autotune geometry] small", flush=True)
    p, pos, vel = relaxed("small")
    best, sps = autotune_geometry(p, pos, vel, cycles=20, lists=True)
    print(f"  tile {list(best.ovr_tile)} skin {best.ovr_skin} cell {best.ovr_cell_scale} {sps:.0f} steps/s")
    assert sps > 0
    assert all(t in (0, 4, 8, 16) for t in best.ovr_tile)
    assert best.ovr_skin in (0.0, 0.75, 1.0) and best.ovr_cell_scale in (0.0, 1.25, 1.5)
    assert abs(best.rvl_eff - (best.rcl_s + (2 * best.ovr_skin if best.ovr_skin else 1))) < 1e-15
    o = orc.Oracle(best)
    o.set_state(pos, vel)
    o.build_lists()
    e = Engine(best)
    e.upload(pos, vel)
    e.build_lists()
    r = o.lists(0)
    assert np.array_equal(e.lists(0), r[np.lexsort((r[:, 1], r[:, 0]))])
    fs, fc = e.force_parts()
    assert orc.rel_rms(fs, o.lj_force() + o.fene_force()[0]) < 1e-11
    assert orc.rel_rms(fc, o.coulomb()) < 1e-10
    e.compute_force(0, MD_NOISE_PHILOX)
    assert o.compute_force(0, MD_NOISE_PHILOX) == 0
    e.run(1, 20)
    assert o.run(1, 20) == 0
    assert np.abs(e.download()[0] - o.state()[0]).max() < 1e-8
    e.close()
