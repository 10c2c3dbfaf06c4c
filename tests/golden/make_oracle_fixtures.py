"""Write the oracle's converged solves that the GPU parity tests compare with.

The nAiT functional's adaptive Airy quadrature makes an oracle solve take
minutes even at 16^3 (about 110 SC iterations with ~25 density evaluations
each), far too long to repeat inside a GPU test.  This script runs the oracle
(oracle/dpft_oracle.c, FP64, OpenMP) once per case and stores inputs' names,
scalar results and fields as tests/golden/oracle_<case>.npz (numpy, no
pickles).  Cases:

  c1_nait_16  the drop-in functional (nAiT T = 0.05, isolated Hartree, Dirac,
              PW92; Z = 10, a = 0.5, N = 10, 20 bohr cube) at 16^3, converged
              at max|dV| < 1e-8
  c1_nait_16_tight  the same converged at max|dV| < 1e-11 and with mu
              bisected to RelAcc 2e-12 (the fixed point to well below the
              1e-10 comparison)
  c2_nait_16  the precision workload's physics (Z = N = 18, a = 0.4, 24 bohr)
              at 16^3, FP64, converged at max|dV| < 1e-8
  c4_nait_16  the differentiation workload's functional on a 3-centre cluster
              at 16^3: the target density, the converged density at the base
              positions and central differences (h = 1e-4) of the loss
              int (n - n_target)^2 and of the total energy for three
              coordinates, each from solves started at the converged
              potential and stopped at max|dV| < 1e-12

    OMP_NUM_THREADS=8 python tests/golden/make_oracle_fixtures.py c1_nait_16 ...
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mpdpft_amd import configs  # noqa: E402
from oracle import oracle as O  # noqa: E402


def result_dict(r):
    S = r.S
    return {
        "SCcount": np.int64(r.SCcount), "converged": np.int64(r.converged), "LinMix": np.int64(r.LinMix),
        "PulMix": np.int64(r.PulMix), "bisection_steps": np.int64(r.bisection_steps),
        "mu": np.array(list(r.muVec)[:S]), "N": np.array(list(r.FinalAbundances)[:S]),
        "TotalEnergy": np.float64(r.TotalEnergy), "CC": np.float64(r.CC),
        "DispersalEnergies": np.array(list(r.DispersalEnergies)[:S]),
        "InteractionEnergies": np.array(list(r.InteractionEnergies)[:16]),
    }


def c1_nait_16():
    p = configs.c1(steps=16, density_expression=5, hartree_mode=1)
    rc, r, den, V = O.solve(p)
    assert rc == 0 and r.converged
    out = result_dict(r)
    out.update(den=den, V=V)
    return out


def c1_nait_16_tight():
    """C1 at 16^3 converged at max|dV| < 1e-11 with mu bisected to RelAcc
    2e-12: the fixed point itself (RelAcc 1e-11 would leave mu ~1e-10 apart
    between two converged solves)."""
    p = configs.c1(steps=16, density_expression=5, hartree_mode=1, tol=1e-11)
    p.RelAcc = 2e-12
    rc, r, den, V = O.solve(p)
    assert rc == 0 and r.converged
    out = result_dict(r)
    out.update(den=den, V=V)
    return out


def c2_nait_16():
    p = configs.c2(steps=16, precision=0)
    rc, r, den, V = O.solve(p)
    assert rc == 0 and r.converged
    out = result_dict(r)
    out.update(den=den, V=V)
    return out


# The 3-centre cluster of the adjoint tests (Z = 1.5, a = 0.7, N = 4).
POS3 = [[-1.1, 0.3, -0.2], [0.9, -0.6, 0.4], [0.2, 1.0, 0.5]]
C4_COORDS = [(0, 0), (1, 1), (2, 2)]
C4_H = 1e-4


def c4_params(pos, tol):
    p = configs.c4(steps=16, edge=20.0, N=4.0, positions=pos, Z=1.5, a=0.7, density_expression=5, tol=tol,
                   maxsc=3000)
    p.muVec[0] = -1.0
    return p


def c4_weights(p):
    EL = p.steps + 1
    w1 = np.ones(EL)
    w1[0] = w1[-1] = 0.5
    return (w1[:, None, None] * w1[None, :, None] * w1[None, None, :]).reshape(-1) * (p.edge / p.steps) ** 3


def c4_nait_16():
    shifted = [[r[0] + 0.3, r[1] - 0.1, r[2]] for r in POS3]
    rc, rt, nt, _ = O.solve(c4_params(shifted, 1e-10))
    assert rc == 0 and rt.converged
    p = c4_params(POS3, 1e-12)
    rc, r, den, V = O.solve(p)
    assert rc == 0 and r.converged
    w = c4_weights(p)
    fdL, fdE = [], []
    for ci, d in C4_COORDS:
        vals = []
        for sgn in (1, -1):
            pos = [list(q) for q in POS3]
            pos[ci][d] += sgn * C4_H
            rc, rr, dd, _ = O.solve(c4_params(pos, 1e-12), v0=V)
            assert rc == 0 and rr.converged
            vals.append((float(np.sum(w * (dd[0] - nt[0]) ** 2)), rr.TotalEnergy))
            print("  c4 fd", ci, d, sgn, rr.SCcount, flush=True)
        fdL.append((vals[0][0] - vals[1][0]) / (2 * C4_H))
        fdE.append((vals[0][1] - vals[1][1]) / (2 * C4_H))
    out = result_dict(r)
    out.update(den=den, V=V, n_target=nt, coords=np.array(C4_COORDS), h=np.float64(C4_H), fd_loss=np.array(fdL),
               fd_energy=np.array(fdE))
    return out


CASES = {"c1_nait_16": c1_nait_16, "c1_nait_16_tight": c1_nait_16_tight, "c2_nait_16": c2_nait_16, "c4_nait_16": c4_nait_16}


def path(case):
    return os.path.join(HERE, "oracle_%s.npz" % case)


if __name__ == "__main__":
    O.lib()
    for name in sys.argv[1:] or list(CASES):
        t0 = time.perf_counter()
        out = CASES[name]()
        np.savez_compressed(path(name), **out)
        print("wrote %s (%.0f s)" % (path(name), time.perf_counter() - t0), flush=True)
