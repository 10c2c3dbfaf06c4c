"""Write tests/golden/TabFunc_2D3D_eta_l5.dat: the first 6 x 31 lines of the
reference's TabFunc_2D3D_eta.dat (the quasi-2D contact's eta triplets for
l1 = 0..5, every l2), the data the interaction-24 tests run on.  Reads the
reference tree."""
import os
import sys

REF = os.environ.get("MPDPFT_REFERENCE", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "TabFunc_2D3D_eta_l5.dat")

if __name__ == "__main__":
    lines = [l for l in open(os.path.join(REF, "TabFunc_2D3D_eta.dat")).read().splitlines() if l.strip()]
    with open(OUT, "w") as f:
        f.write("\n".join(lines[:6 * 31]) + "\n")
    print("wrote", OUT, file=sys.stderr)
