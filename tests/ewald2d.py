"""Two-dimensional Ewald sum (EW2D) in numpy: an independent reference for
the slab geometry of the reference program (x, y periodic; z open).

The reference treats the open z direction with a 3D Ewald sum over a box
stretched by vacuum plus the Yeh-Berkowitz dipole correction
(compt_acc.f90:1040-1048).  That combination approximates the exact
2D-periodic sum below; its remaining error comes only from in-plane wave
vectors of the image slabs and falls like exp(-2 pi gap / L).

Energy of charges q_i at r_i in a cell of area A = Lx Ly, periodic in x, y
(Parry 1975; Heyes, Barber, Clarke 1977; Yeh & Berkowitz 1999, eq. 2):

  E = 1/2 sum_ij q_i q_j [ sum'_n erfc(a |r_ij + n|) / |r_ij + n|
        + (pi/A) sum_{h != 0} cos(h.s_ij)/h (e^{h z} erfc(h/2a + a z)
                                            + e^{-h z} erfc(h/2a - a z)) ]
      - (pi/A) sum_ij q_i q_j [ z erf(a z) + exp(-a^2 z^2) / (a sqrt(pi)) ]
      - a/sqrt(pi) sum_i q_i^2

with z = z_ij and s_ij the in-plane separation; h runs over the 2D
reciprocal lattice 2 pi (m/Lx, n/Ly).  The h = 0 term is the potential of a
Gaussian-smeared charge sheet, phi'' = -4 pi rho.  Forces are minus the
gradient (central differences here), times c1 = lB kT as in the reference.
"""
import math

import numpy as np
from scipy.special import erf, erfc, erfcx


def energy(x, q, Lx, Ly, alpha, mmax, c1=1.0):
    """EW2D energy of charges q at positions x (N,3)."""
    x = np.asarray(x, float)
    q = np.asarray(q, float)
    N = len(q)
    A = Lx * Ly
    qq = q[:, None] * q[None, :]
    d = x[:, None, :] - x[None, :, :]
    # real space: every image within one cell (erfc(a L) is negligible)
    e_real = 0.0
    eye = np.eye(N, dtype=bool)
    for nx in (-1, 0, 1):
        for ny in (-1, 0, 1):
            dd = d + np.array([nx * Lx, ny * Ly, 0.0])
            r = np.sqrt((dd ** 2).sum(-1))
            if nx == 0 and ny == 0:
                r = np.where(eye, 1.0, r)
                t = np.where(eye, 0.0, erfc(alpha * r) / r)
            else:
                t = erfc(alpha * r) / r
            e_real += 0.5 * np.sum(qq * t)
    # reciprocal space, h != 0: e^{+-hz} erfc(h/2a +- a z) written with
    # erfcx so that no factor overflows
    m = np.arange(-mmax, mmax + 1)
    hx, hy = np.meshgrid(2 * math.pi * m / Lx, 2 * math.pi * m / Ly, indexing="ij")
    hx, hy = hx.ravel(), hy.ravel()
    keep = (hx != 0) | (hy != 0)
    hx, hy = hx[keep], hy[keep]
    h = np.sqrt(hx * hx + hy * hy)
    z = d[..., 2]
    g = np.exp(-(h ** 2)[:, None, None] / (4 * alpha ** 2) - alpha ** 2 * z[None] ** 2)
    u1 = (h / (2 * alpha))[:, None, None] + alpha * z[None]
    u2 = (h / (2 * alpha))[:, None, None] - alpha * z[None]
    ph = np.cos(hx[:, None, None] * d[None, ..., 0] + hy[:, None, None] * d[None, ..., 1])
    e_rec = 0.5 * math.pi / A * np.sum((qq[None] * ph * g * (erfcx(u1) + erfcx(u2))) / h[:, None, None])
    e_h0 = -math.pi / A * np.sum(qq * (z * erf(alpha * z) +
                                       np.exp(-alpha ** 2 * z ** 2) / (alpha * math.sqrt(math.pi))))
    e_self = -alpha / math.sqrt(math.pi) * np.sum(q * q)
    return c1 * (e_real + e_rec + e_h0 + e_self)


def forces(x, q, Lx, Ly, alpha, mmax, c1=1.0, h=1e-5):
    """-grad E by central differences, (N,3)."""
    x = np.asarray(x, float)
    f = np.zeros_like(x)
    for i in range(len(q)):
        for dim in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, dim] += h
            xm[i, dim] -= h
            f[i, dim] = -(energy(xp, q, Lx, Ly, alpha, mmax, c1) -
                          energy(xm, q, Lx, Ly, alpha, mmax, c1)) / (2 * h)
    return f
