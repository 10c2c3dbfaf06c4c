"""Print the per-iteration trace of one HIP solve of a named workload
(MPDPFT_VERBOSE lines on stderr) and its result; used to compare SC
trajectories with the oracle's (ORACLE_VERBOSE=1).

    python tests/golden/trace_solve.py c1_16
"""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
os.environ["MPDPFT_VERBOSE"] = "1"
from mpdpft_amd import Context, configs  # noqa: E402

CASES = {
    "c1_16": lambda: configs.c1(steps=16, density_expression=5, hartree_mode=1),
    "c2_16": lambda: configs.c2(steps=16, precision=0),
}

if __name__ == "__main__":
    ctx = Context(CASES[sys.argv[1]]())
    r = ctx.solve()
    print("done SCcount %d bisection %d Lin %d Pul %d E %.15g mu %.15g" % (r.SCcount, r.bisection_steps, r.LinMix,
                                                                           r.PulMix, r.TotalEnergy, r.muVec[0]))
