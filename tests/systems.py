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


__all__ = ["system", "etas", "MD_FP32", "MD_FP64"]
