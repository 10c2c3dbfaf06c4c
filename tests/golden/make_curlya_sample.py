"""Write tests/golden/curlya_ref_sample.txt: every 125th (x, CurlyA) pair of the
reference's TabFunc_CurlyA.dat (x on one line, value on the next), used to pin
this repository's regenerated CurlyA knots.  Reads the reference tree."""
import os
import sys

REF = os.environ.get("MPDPFT_REFERENCE", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "curlya_ref_sample.txt")

if __name__ == "__main__":
    toks = open(os.path.join(REF, "TabFunc_CurlyA.dat")).read().split()
    pairs = list(zip(toks[0::2], toks[1::2]))
    with open(OUT, "w") as f:
        f.write("# index x CurlyA(x) -- from TabFunc_CurlyA.dat (reference data), every 125th pair\n")
        for i in range(0, len(pairs), 125):
            f.write("%d %s %s\n" % (i, pairs[i][0], pairs[i][1]))
    print("wrote", OUT, file=sys.stderr)
