"""Build oracle/_ref/ from the reference tree (test infrastructure only).

Only one piece of the reference's path compiles here: the vendored Eigen 3.4.0
headers (Eigen_Headers/), used for the Pulay Gram inverse.  The rest of the
reference (mpDPFT.cpp) needs FFTW, GSL, libxc, Boost and the missing
optimization.cpp, none of which exist in this image, so it is not built.
Outputs go to oracle/_ref/ (git-ignored, kept beside the built library).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.environ.get("MPDPFT_REFERENCE", "/root/reference")
OUT = os.path.join(HERE, "_ref", "libref_eigen_qr.so")


def build():
    eigen = os.path.join(REF, "Eigen_Headers")
    if not os.path.isdir(eigen):
        return None
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", eigen, "-o", OUT + ".tmp",
           os.path.join(HERE, "ref_eigen_qr.cpp")]
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build() or "reference tree absent; nothing built", file=sys.stderr)
