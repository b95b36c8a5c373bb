"""Named systems: the reference's shipped parameter files and the benchmark
workloads (brush geometry, ions, SPME mesh) expressed as md_params.

The reference reads system_data.txt (in_output.f90:189-213) and
force_data.txt (compt_acc.f90:140-157); ``reference_defaults`` holds the
values those files ship with, ``write_param_files`` writes files in the same
one-value-per-line format.
"""
import math

from .abi import MD_FP32, MD_FP64, MdParams, derive

# system_data.txt as shipped (23 values, file order)
SYSTEM_DATA = [
    ("uniform_or_random", 1), ("Lz", 200.0), ("Z_empty", 2.0), ("sigmag", 5e-3), ("Beta", 0.8333),
    ("qq", -1.0), ("qqi", 3.0), ("ion_ratio", 0.0), ("arm", 3), ("Nma", 50), ("Nga", 60),
    ("Nml", 101), ("Ngl", 6), ("man_l", 7), ("man_s", 3), ("R_bond", 0.97),
    ("StepNum0", 3000000), ("StepNum", 7000000), ("DeltaStep1", 2000), ("DeltaStep2", 200),
    ("DeltaStep3", 20000), ("multistep", 10), ("dt", 0.005),
]
# force_data.txt as shipped (16 values, file order); ordr is three lines
FORCE_DATA = [
    ("epsilon", 1.0), ("sigma", 1.0), ("sigma_s", 2.0), ("rcl", 1.12246), ("rvl", 1.81246),
    ("rsk", 1.0), ("R0_2", 2.25), ("kFENE", 30.0), ("ordr", (6, 6, 6)), ("lb", 3.0), ("xi", 1.0),
    ("EF", 1.0), ("tol", 1.9), ("tau_rf", 800.0),
]


def _apply(p, items):
    for k, v in items:
        if k == "ordr":
            for d in range(3):
                p.ordr[d] = int(v[d])
        elif k == "charged_arm":
            for a in range(len(v)):
                p.charged_arm[a] = int(v[a])
        elif k == "ovr_ng":
            for d in range(3):
                p.ovr_ng[d] = int(v[d])
        else:
            setattr(p, k, v)
    return p


def reference_defaults():
    """Raw parameters of the shipped files (q_arm.txt: all arms charged)."""
    p = MdParams()
    _apply(p, SYSTEM_DATA)
    _apply(p, FORCE_DATA)
    _apply(p, [("charged_arm", [1] * 16), ("seed", 20261203)])
    return p


def write_param_files(p, system_path, force_path, q_arm_path=None):
    """Writes p's raw values in the reference's file formats."""
    with open(system_path, "w") as f:
        for k, _ in SYSTEM_DATA:
            v = getattr(p, k)
            f.write(f"{v!r}\n" if isinstance(v, float) else f"{v}\n")
    with open(force_path, "w") as f:
        for k, _ in FORCE_DATA:
            if k == "ordr":
                for d in range(3):
                    f.write(f"{p.ordr[d]}\n")
            else:
                f.write(f"{getattr(p, k)!r}\n")
    if q_arm_path:
        with open(q_arm_path, "w") as f:
            for a in range(p.arm):
                f.write(f"{p.charged_arm[a]}\n")


def _base():
    p = reference_defaults()
    _apply(p, [("uniform_or_random", 0), ("ordr", (4, 4, 4)), ("tol", 3.0)])
    return p


# Ewald splitting used with the fixed benchmark meshes: real-space cut
# r_c = 3.0/alpha, alpha = 0.25 at mesh spacing h ~ 1.7 sigma and order 4.
# Measured at the default brush's seeded start with the FP64 oracle's
# error_analysis (SPME + real space against standard Ewald at tol 5): a
# relative Coulomb force RMS of 4.0e-4, the mesh's discretisation error.
ALPHA_H17_P4 = 0.25


def config(name, precision=MD_FP64, ez=None, seed=20261203, order=None):
    """md_params (derived) of a named workload.

    small    4x4 linear chains of 50 beads, every 3rd charged, counterions
    default  8x8 linear chains of 100 charged beads + trivalent salt (21,396)
    star3    8x8 3-arm stars, arms of 33 beads + trivalent salt (21,184)
    mixed    16x16 sites, 4-arm stars (arms 25) and chains of 100 (85,588)
    """
    p = _base()
    if name == "small":
        L = math.sqrt(16 / 0.05)
        _apply(p, [("Nga", 0), ("Ngl", 16), ("Nml", 51), ("man_l", 3), ("man_s", 1), ("qqi", 0.0),
                   ("ion_ratio", 0.0), ("Lz", 75.0), ("Z_empty", 2 * 75.0 / L), ("sigmag", 0.05),
                   ("EF", 0.0), ("ovr_ng", (32, 32, 128)), ("ovr_alpha", ALPHA_H17_P4)])
    elif name == "default":
        _apply(p, [("Nga", 0), ("Ngl", 64), ("Nml", 101), ("man_l", 1), ("man_s", 1), ("qqi", 3.0),
                   ("ion_ratio", 1.0), ("EF", 0.5), ("ovr_ng", (64, 64, 256)),
                   ("ovr_alpha", ALPHA_H17_P4)])
    elif name == "star3":
        _apply(p, [("Nga", 64), ("Ngl", 0), ("arm", 3), ("Nma", 33), ("man_l", 1), ("man_s", 1),
                   ("qqi", 3.0), ("ion_ratio", 1.0), ("EF", 0.5 if ez is None else ez),
                   ("ovr_ng", (64, 64, 256)), ("ovr_alpha", ALPHA_H17_P4)])
    elif name == "mixed":
        _apply(p, [("Nga", 128), ("Ngl", 128), ("arm", 4), ("Nma", 25), ("Nml", 101), ("man_l", 1),
                   ("man_s", 1), ("qqi", 3.0), ("ion_ratio", 1.0), ("Lz", 300.0), ("EF", 0.5),
                   ("ordr", (6, 6, 6)), ("ovr_ng", (128, 128, 512)), ("ovr_alpha", ALPHA_H17_P4)])
    else:
        raise KeyError(name)
    if ez is not None:
        p.EF = ez
    if order is not None:
        p.ordr[0] = p.ordr[1] = p.ordr[2] = order
    p.precision = precision
    p.seed = seed
    return derive(p)


def replica(chains_side, n_beads, arch, valence, ez, seed, precision=MD_FP32, preheat=0):
    """One request of the serving stream: chains_side^2 grafted molecules of
    n_beads beads (linear, or stars with arch arms sharing the beads), all
    charged, counterions plus salt of the given valence at charge ratio 1.
    In a batch, steps 1..preheat follow the pre-heat form (StepNum0 of
    system_data.txt: velocities rescaled every 20 steps, main.f90:67-75)."""
    p = _base()
    n = chains_side * chains_side
    if arch == 1:
        _apply(p, [("Nga", 0), ("Ngl", n), ("Nml", n_beads + 1)])
    else:
        nma = max(2, n_beads // arch)
        _apply(p, [("Nga", n), ("Ngl", 0), ("arm", arch), ("Nma", nma)])
    _apply(p, [("man_l", 1), ("man_s", 1), ("qqi", float(valence)), ("ion_ratio", 1.0), ("EF", ez)])
    # mesh: the smallest power of two with a spacing of at most 1.8 sigma,
    # which keeps the default brush's spacing (113.1/64 = 1.77 sigma in x, y;
    # 426.3/256 = 1.67 in z) for boxes of its size
    q = derive(p)
    ng = []
    for L in (q.Lx, q.Ly, q.Lz * q.Zfac):
        k = 8
        while L / k > 1.8:
            k *= 2
        ng.append(k)
    p.ovr_ng[0], p.ovr_ng[1], p.ovr_ng[2] = ng
    p.ovr_alpha = ALPHA_H17_P4
    p.precision = precision
    p.seed = seed
    p.StepNum0 = preheat
    q = derive(p)
    # The reference's real-space list in a box with fewer than three list
    # cells along x or y names a neighbour cell twice and sums those pairs
    # twice (compt_acc.f90:545-558, 1012-1038): 3% of the Coulomb force at
    # the 2x2 boxes' start.  Those requests take the reference's all-pairs
    # branch instead (real_verlet = 0: each pair once by the minimum image,
    # r_c = 12 <= L/2).
    if int(q.Lx / q.rvc) < 3 or int(q.Ly / q.rvc) < 3:
        q.real_verlet = 0
    return q


PRECISIONS = {"fp64": MD_FP64, "fp32": MD_FP32}
