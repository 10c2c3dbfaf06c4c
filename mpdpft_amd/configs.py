"""Named workloads and the mpDPFT.input writer.

C0: 2D two-species TF + contact gas (N = 100 each, 128^2 on [-16,16]^2).
C1: 3D single-species electrons, softened Coulomb Z = 10, a = 0.5, N = 10,
    Hartree (isolated, zero-padded 256^3) + Dirac + PW92, 128^3 on a 20 bohr cube.
C2: the precision variant of C1 at 256^3 (Z = N = 18, a = 0.4, 24 bohr).
C3: seeded batches of 2D two-species TF + contact gases in anisotropic traps.
C4: 3D cluster of softened Coulomb centres (environment 41) for the adjoint.
"""
import math

from . import default_params

# Names of the 47 positional lines of mpDPFT.input, in order.
LINE_NAMES = [
    "System", "Symmetry", "TaskType", "TaskParameterCount", "TaskHyperBox", "DIM", "Units", "edge", "steps",
    "S", "Abundances", "RelAcc", "InternalAcc", "K", "Resources", "DensityExpression", "mpp", "Mpp", "tauVec",
    "tVec", "TVec", "Environments", "stretchfactor", "Wall", "Noise", "InterpolVQ", "Interactions",
    "incrementalV", "muVec", "DeltamuModifier", "Mixer", "thetaVec", "SCcriterion", "maxSCcount", "method",
    "DelFieldMethod", "MovieQ", "Schedule", "EkinTypes", "alpha", "beta", "gamma", "gammaH", "degeneracy",
    "regularize", "RegularizationThreshold", "Print",
]


def _fmt(v):
    if isinstance(v, float):
        return repr(v)
    return str(v)


def write_input(p, path, extensions=True):
    """Write params as an mpDPFT.input file (47 positional lines + extension keys)."""
    S = p.S

    def vec(name, n):
        return " ".join(_fmt(x) for x in list(getattr(p, name))[:n])

    vals = {
        "System": p.System, "Symmetry": p.Symmetry, "TaskType": p.TaskType,
        "TaskParameterCount": p.TaskParameterCount,
        "TaskHyperBox": vec("TaskHyperBox", p.n_hyperbox) or "0. 0.", "DIM": p.DIM, "Units": p.Units,
        "edge": _fmt(p.edge), "steps": p.steps, "S": S, "Abundances": vec("Abundances", S),
        "RelAcc": _fmt(p.RelAcc), "InternalAcc": _fmt(p.InternalAcc), "K": p.K, "Resources": "0.",
        "DensityExpression": p.DensityExpression, "mpp": vec("mpp", p.n_mpp), "Mpp": vec("Mpp", p.n_Mpp) or "0.",
        "tauVec": vec("tauVec", S), "tVec": vec("tVec", S), "TVec": vec("TVec", S),
        "Environments": vec("Environments", S), "stretchfactor": _fmt(p.stretchfactor), "Wall": _fmt(p.Wall),
        "Noise": _fmt(p.Noise), "InterpolVQ": p.InterpolVQ, "Interactions": vec("Interactions", p.n_interactions),
        "incrementalV": _fmt(p.incrementalV), "muVec": vec("muVec", S), "DeltamuModifier": _fmt(p.DeltamuModifier),
        "Mixer": vec("Mixer", 3), "thetaVec": vec("thetaVec", S), "SCcriterion": _fmt(p.SCcriterion),
        "maxSCcount": p.maxSCcount, "method": p.method, "DelFieldMethod": p.DelFieldMethod, "MovieQ": p.MovieQ,
        "Schedule": vec("Schedule", p.n_schedule), "EkinTypes": vec("EkinTypes", p.n_ekin), "alpha": _fmt(p.alpha),
        "beta": _fmt(p.beta), "gamma": _fmt(p.gamma), "gammaH": _fmt(p.gammaH), "degeneracy": _fmt(p.degeneracy),
        "regularize": vec("regularize", p.n_regularize), "RegularizationThreshold": _fmt(p.RegularizationThreshold),
        "Print": p.Print,
    }
    lines = ["%s\t\t\t//%s" % (vals[n], n) for n in LINE_NAMES]
    if extensions:
        lines += [
            "%d\t\t\t//HartreeMode" % p.HartreeMode,
            "%s\t\t\t//SmoothingKWidth" % _fmt(p.SmoothingKWidth),
            "%d\t\t\t//ConvergenceMode" % p.ConvergenceMode,
            "%d\t\t\t//Precision" % p.Precision,
        ]
        if p.CubeFile:
            lines.append("%d\t\t\t//CubeFile" % p.CubeFile)
        if p.NoiseSeed:
            lines.append("%d\t\t\t//NoiseSeed" % p.NoiseSeed)
        for c in range(p.n_centres):
            lines.append("%s\t\t\t//Centre" % " ".join(_fmt(x) for x in p.centres[c]))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _set(arr, vals):
    for i, v in enumerate(vals):
        arr[i] = v


def c0(steps=128, edge=32.0, N=100.0, g=1.0, convergence_mode=1, tol=1e-8, maxsc=5000):
    """2D two-species repulsive Fermi gas, TF + contact (drop-in inputs of the reference)."""
    p = default_params()
    p.DIM, p.Units, p.edge, p.steps, p.S = 2, 1, edge, steps, 2
    _set(p.Abundances, [N, N])
    p.RelAcc = 1e-11
    p.DensityExpression = 7
    p.n_mpp = 2
    _set(p.mpp, [0.5, 0.5])
    _set(p.Environments, [1, 1])
    p.n_interactions = 1
    p.Interactions[0] = 1
    p.beta = g
    p.degeneracy = 1.0
    _set(p.muVec, [10.0, 10.0])
    p.DeltamuModifier = -3.0
    _set(p.Mixer, [1.0, 8.0, 0.0])
    _set(p.thetaVec, [0.1, 0.1])
    p.SCcriterion = tol
    p.ConvergenceMode = convergence_mode
    p.maxSCcount = maxsc
    p.DelFieldMethod = 5
    p.Print = 0
    return p


def c1(steps=128, edge=20.0, Z=10.0, a=0.5, N=10.0, density_expression=7, hartree_mode=1, kwidth=0.3,
       tol=1e-8, maxsc=2000, T=0.05):
    """3D electrons in a softened Coulomb well: Hartree + Dirac + PW92."""
    p = default_params()
    p.DIM, p.Units, p.edge, p.steps, p.S = 3, 1, edge, steps, 1
    p.Abundances[0] = N
    p.RelAcc = 1e-11
    p.InternalAcc = 1e-6
    p.DensityExpression = density_expression
    p.n_mpp = 2
    _set(p.mpp, [Z, a])
    p.n_Mpp = 3
    _set(p.Mpp, [10.0, 20.0, 1e-6])
    p.TVec[0] = T
    p.Environments[0] = 40
    p.n_interactions = 3
    _set(p.Interactions, [8, 11, -12])
    p.gammaH = 1.0
    p.degeneracy = 2.0
    p.muVec[0] = -1.0
    p.DeltamuModifier = -3.0
    _set(p.Mixer, [1.0, 10.0, 0.0])
    p.thetaVec[0] = 0.05
    p.SCcriterion = tol
    p.ConvergenceMode = 1
    p.maxSCcount = maxsc
    p.DelFieldMethod = 5
    p.HartreeMode = hartree_mode
    p.SmoothingKWidth = kwidth
    p.Print = 0
    return p


def c2(steps=256, edge=24.0, Z=18.0, a=0.4, N=18.0, precision=1, tol=1e-8, maxsc=2000, T=0.05, kwidth=0.3):
    """Precision variant of c1 at 256^3 on a 24 bohr cube, Z = 18, N = 18, a = 0.4;
    precision 1: FP32 transforms and pointwise potentials, then FP64 polish."""
    p = c1(steps=steps, edge=edge, Z=Z, a=a, N=N, density_expression=5, hartree_mode=1, kwidth=kwidth, tol=tol,
           maxsc=maxsc, T=T)
    p.Precision = precision
    return p


class SplitMix64:
    """SplitMix64 (Steele, Lea & Flood 2014): u = (x >> 11) 2^-53."""

    def __init__(self, seed):
        self.state = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * 2.0 ** -53

    def integer(self, a, b):
        return a + int(self.uniform() * (b - a + 1))


def c3(n=512, seed=20260101, grids=(64, 128, 256), tol=1e-8, maxsc=5000):
    """Batched 2D two-species gases: per instance, in this order, N ~ U{20..400}
    (both species), lambda ~ U[1,2], g ~ U[0.2,2.0], grid 64/128/256 with
    probabilities 0.5/0.3/0.2 (grids[0..2]), box 2.5 sqrt(2N); trap
    (x^2 + lambda^2 y^2)/2 as environment 3; TF + contact g; Pulay history 8,
    theta 0.1, converged when max|dV| < tol."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(n):
        N = rng.integer(20, 400)
        lam = 1.0 + rng.uniform()
        g = 0.2 + 1.8 * rng.uniform()
        u = rng.uniform()
        steps = grids[0] if u < 0.5 else (grids[1] if u < 0.8 else grids[2])
        p = c0(steps=steps, edge=2.5 * math.sqrt(2.0 * N), N=float(N), g=g, tol=tol, maxsc=maxsc)
        e = (lam * lam - 1.0) / (lam * lam + 1.0)
        c = (1.0 + lam * lam) / 4.0
        p.n_mpp = 4
        _set(p.mpp, [c, c, e, e])
        _set(p.Environments, [3, 3])
        _set(p.Mixer, [1.0, 8.0, 0.0])
        _set(p.thetaVec, [0.1, 0.1])
        out.append(p)
    return out


def centre_positions(n=32, seed=424242, box=8.0):
    """n points by rejection sampling: x, y, z = -box + 2 box u (SplitMix64),
    kept when |r| < box."""
    rng = SplitMix64(seed)
    out = []
    while len(out) < n:
        r = [-box + 2.0 * box * rng.uniform() for _ in range(3)]
        if math.sqrt(r[0] ** 2 + r[1] ** 2 + r[2] ** 2) < box:
            out.append(r)
    return out


def c4(steps=96, edge=30.0, N=32.0, n_centres=32, Z=1.0, a=0.6, seed=424242, density_expression=5,
       kwidth=0.3, tol=1e-10, maxsc=3000, T=0.05, positions=None):
    """3D cluster: sum of softened Coulomb centres -Z/sqrt(|r - R_c|^2 + a^2)
    (environment 41) at seeded positions inside a sphere of radius 8 bohr;
    the functional of c1 (nAiT or TF, isolated Hartree, Dirac, PW92)."""
    p = c1(steps=steps, edge=edge, N=N, density_expression=density_expression, hartree_mode=1, kwidth=kwidth,
           tol=tol, maxsc=maxsc, T=T)
    p.Environments[0] = 41
    pos = positions if positions is not None else centre_positions(n_centres, seed)
    p.n_centres = len(pos)
    for i, r in enumerate(pos):
        _set(p.centres[i], [r[0], r[1], r[2], Z, a])
    p.RelAcc = 2e-12
    return p
