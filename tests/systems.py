"""Seeded test systems shared by the parity tests (cached per process)."""
import functools

import numpy as np

from md_brushes_amd import MD_FP32, MD_FP64, config, initialize_position, initialize_velocity


@functools.lru_cache(maxsize=None)
def system(name, precision=MD_FP64, seed=20261203, ez=None):
    """(params, pos(NN,4), vel(NN,3)) of a named workload, seeded."""
    p = config(name, precision=precision, ez=ez, seed=seed)
    pos = initialize_position(p, seed)
    vel = initialize_velocity(p, seed + 1)
    pos.setflags(write=False)
    vel.setflags(write=False)
    return p, pos, vel


def etas(NN, nsteps, seed=7):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal((NN, 3)) for _ in range(nsteps)]


def heights_ref(pos, p):
    """The height analysis of write_height (in_output.f90:1017-1079) from
    positions (NN,4), 0 for a molecule kind the system lacks: h_avg,
    hl_max, hl_end, hl_avg, hs_max, hs_end, hs_branch, hs_avg."""
    z = np.asarray(pos)[:, 2]
    per = p.arm * p.Nma + 1
    nta = p.Nga * per
    out = {"h_avg": z[: p.Npe].mean()}
    lin = z[nta: nta + p.Ngl * p.Nml].reshape(p.Ngl, p.Nml) if p.Ngl else np.zeros((0, 1))
    out["hl_max"] = np.maximum(lin.max(axis=1), 0).mean() if p.Ngl else 0.0
    out["hl_end"] = lin[:, -1].mean() if p.Ngl else 0.0
    out["hl_avg"] = lin.mean() if p.Ngl else 0.0
    st = z[:nta].reshape(p.Nga, per) if p.Nga else np.zeros((0, 1))
    out["hs_max"] = np.maximum(st.max(axis=1), 0).mean() if p.Nga else 0.0
    ends = [p.Nma * j for j in range(2, p.arm + 1)]
    out["hs_end"] = st[:, ends].mean() if p.Nga and p.arm > 1 else 0.0
    out["hs_branch"] = st[:, p.Nma].mean() if p.Nga else 0.0
    out["hs_avg"] = st.mean() if p.Nga else 0.0
    return out


__all__ = ["system", "etas", "heights_ref", "MD_FP32", "MD_FP64"]
