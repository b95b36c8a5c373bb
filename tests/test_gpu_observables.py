"""FP32 forces against FP64 per observable: brush height, the z-density
profile of the polymer beads (in_output.f90:1017-1020, 1878-1886) and the
energies, over trajectories of the default brush.  Both runs start from the
same state with the same Philox noise stream; after a relaxation they are
sampled every 50 steps, and block averages give each mean's standard error.
The FP32 run must agree with the FP64 run within four combined standard
errors (the trajectories decorrelate, so this is a statistical bar)."""
import numpy as np
import pytest

from md_brushes_amd import MD_FP32, MD_FP64, MD_NOISE_PHILOX, Engine
from systems import system

pytestmark = pytest.mark.gpu

NSTEPS, RELAX, EVERY, BLOCKS = 10000, 2000, 50, 8


def _samples(prec):
    p, pos, vel = system("default", prec)
    e = Engine(p)
    e.upload(pos, vel)
    e.build_lists()
    e.compute_force(0, MD_NOISE_PHILOX)
    e.run(1, RELAX, preheat=True)
    h, prof, epot = [], [], []
    edges = np.linspace(0, p.Lz, 26)  # 25 coarse bins of the 500-bin profile
    for s in range(RELAX + 1, NSTEPS + 1, EVERY):
        e.run(s, EVERY, preheat=True)
        z = e.download()[0][: p.Npe, 2]
        h.append(z.mean())
        prof.append(np.histogram(z, edges)[0])
        epot.append(e.energies()["potential"])
    e.close()
    return np.array(h), np.array(prof, float), np.array(epot)


def _mean_se(x):
    b = np.array_split(x, BLOCKS)
    m = np.array([bb.mean(axis=0) for bb in b])
    return x.mean(axis=0), m.std(axis=0, ddof=1) / np.sqrt(BLOCKS)


def test_fp32_observables_within_statistical_error():
    print("[fp32 observables] default brush, 10^4 steps each", flush=True)
    h64, p64, e64 = _samples(MD_FP64)
    h32, p32, e32 = _samples(MD_FP32)
    for name, a, b in (("h_avg", h32, h64), ("potential", e32, e64)):
        ma, sa = _mean_se(a)
        mb, sb = _mean_se(b)
        print(f"{name}: fp32 {ma:.4f} +- {sa:.4f}, fp64 {mb:.4f} +- {sb:.4f}")
        assert abs(ma - mb) < 4 * np.hypot(sa, sb) + 1e-12, name
    ma, sa = _mean_se(p32)
    mb, sb = _mean_se(p64)
    se = np.hypot(sa, sb)
    sel = se > 0
    z = (ma[sel] - mb[sel]) / se[sel]
    print(f"profile: {sel.sum()} bins, max |z| {np.abs(z).max():.2f}, chi2/dof {np.mean(z ** 2):.2f}")
    assert np.all(ma[~sel] == mb[~sel])
    assert np.mean(z ** 2) < 4.0
