"""The host-side readers (mpDPFT.input parser, LoadV) and the oracle under
AddressSanitizer + UndefinedBehaviorSanitizer (CPU only).

tests/sanitize_driver.cpp is compiled together with mpdpft_amd/csrc/params.cpp
and oracle/dpft_oracle.c by g++/gcc with -fsanitize=address,undefined and
-fno-sanitize-recover=all, then run on inputs written here: a 2D two-species
input, a 3D nAiT input and one mpDPFT_V.dat file per dimension.  Any
sanitizer report makes the driver exit non-zero."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from mpdpft_amd import configs, tables
from test_import_v import lattice, write_vfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None or shutil.which("gcc") is None:
        pytest.skip("no host compiler")
    out = tmp_path_factory.mktemp("san")
    objs = []
    for src, cc, std in ((os.path.join(ROOT, "oracle", "dpft_oracle.c"), "gcc", "-std=gnu11"),
                         (os.path.join(ROOT, "mpdpft_amd", "csrc", "params.cpp"), "g++", "-std=c++17"),
                         (os.path.join(ROOT, "tests", "sanitize_driver.cpp"), "g++", "-std=c++17")):
        obj = str(out / (os.path.basename(src) + ".o"))
        subprocess.run([cc, std, "-fopenmp", "-I", os.path.join(ROOT, "include"), "-c", src, "-o", obj] + SAN,
                       check=True, capture_output=True, text=True)
        objs.append(obj)
    exe = str(out / "sanitize_driver")
    subprocess.run(["g++", "-fopenmp", "-o", exe] + objs + SAN + ["-lm"], check=True, capture_output=True, text=True)
    return exe


def test_readers_and_oracle_under_sanitizers(driver, tmp_path):
    p2 = configs.c0(steps=32, maxsc=12)
    configs.write_input(p2, str(tmp_path / "c0.input"))
    p3 = configs.c1(steps=16, density_expression=5, hartree_mode=1, maxsc=2)
    configs.write_input(p3, str(tmp_path / "c1.input"))
    a2 = lattice(p2)
    X, Y = np.meshgrid(a2, a2, indexing="ij")
    write_vfile(tmp_path / "v2.dat", [a2, a2], [0.5 * (X ** 2 + Y ** 2), 0.5 * (X ** 2 + Y ** 2) + 0.1 * X])
    a3 = lattice(p3)
    X, Y, Z = np.meshgrid(a3, a3, a3, indexing="ij")
    write_vfile(tmp_path / "v3.dat", [a3, a3, a3], [-10. / np.sqrt(X ** 2 + Y ** 2 + Z ** 2 + 0.25)])
    env = dict(os.environ)
    env["MPDPFT_TABLES"] = tables.ensure(verbose=False)
    env["OMP_NUM_THREADS"] = "2"
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0"
    r = subprocess.run([driver, str(tmp_path)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "all checks passed" in r.stdout
