"""The plot shells (mpDPFT_Plots.sh, mpDPFT_CombinedPlots.tex; ExportData,
mpDPFT.cpp:12802-13063) and the Cube file (mpDPFT_Den_Cube.dat, :12458-12490)
written by mpdpft_run_dir.

Plot shells: every column a figure reads exists in the data file it reads and
holds the field the figure names (ContourData: x, y, |r| then density,
environment, potential per species; CutData: position then the same triple
per species), the mu line carries the run's mu, and the LaTeX file includes
exactly the figures the script makes, after the run's result block.  gnuplot
runs the script when it is installed.

Cube file: the header of a Gaussian cube (nucleus count, three axis lines,
the environment-41 centres as nuclei), then max(1e-14, n dx^3) over
[0, steps)^3 with x slowest, equal to the density of the same solve through
the C ABI (16 printed digits).

Movie frames (MovieQ 1, 2): the cut of every stored frame, checked below."""
import re
import shutil
import subprocess

import numpy as np
import pytest

from mpdpft_amd import Context, configs, run_dir

pytestmark = pytest.mark.gpu


def columns(path):
    for line in open(path):
        if line.strip():
            return len(line.split())


def test_plot_shells_read_the_right_columns(tmp_path):
    print("case plot shells 2D two species", flush=True)
    p = configs.c0(steps=32, edge=8.0, N=6.0, maxsc=40)
    configs.write_input(p, str(tmp_path / "mpDPFT.input"))
    r = run_dir(tmp_path)
    sh = (tmp_path / "mpDPFT_Plots.sh").read_text()
    tex = (tmp_path / "mpDPFT_CombinedPlots.tex").read_text()
    ncon = columns(tmp_path / "mpDPFT_ContourData.dat")
    ncut = columns(tmp_path / "mpDPFT_CutData.dat")
    assert ncon == 3 + 3 * p.S
    blocks = re.findall(r"figure mpDPFT_Plots-(\S+) \S+ \S+ <<'EOF'\n(.*?)\nEOF\n", sh, re.S)
    names = [b[0] for b in blocks]
    fields = {"Den": 0, "Env": 1, "V": 2}
    for name, body in blocks:
        m = re.match(r"2Dcontour(Den|Env|V)(\d+)$", name)
        if m:
            col = int(re.search(r"ContourData.dat' using 1:2:(\d+)", body).group(1))
            assert col == 4 + 3 * int(m.group(2)) + fields[m.group(1)], name
            continue
        for file, col in re.findall(r"'mpDPFT_(\w+)\.dat' using 1:(\d+)", body):
            assert file == "CutData" and 2 <= int(col) <= ncut, (name, col)
        m = re.match(r"(?:LogPlot-)?Cut(Den|Env\+V)(\d+)$", name)
        if m:
            s = int(m.group(2))
            cols = [int(c) for c in re.findall(r"using 1:(\d+)", body)]
            if m.group(1) == "Den":
                assert cols == [2 + 3 * s], name
            else:
                assert cols == [3 + 3 * s, 4 + 3 * s], name
                mu = float(re.search(r"mu\(x\) = (\S+)", body).group(1))
                assert mu == r.muVec[s]
    for s in range(p.S):
        for f in ("2DcontourDen", "2DcontourEnv", "2DcontourV", "CutDen", "LogPlot-CutDen", "CutEnv+V"):
            assert f + str(s) in names
    assert "Overview-CutDen" in names
    assert re.findall(r"\\includegraphics\[[^]]*\]\{mpDPFT_Plots-(\S+)\}", tex) == names
    assert "TotalEnergy" in tex and tex.index("TotalEnergy") < tex.index("includegraphics")
    if shutil.which("gnuplot"):
        out = subprocess.run(["bash", str(tmp_path / "mpDPFT_Plots.sh")], capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr[-2000:]
        for n in names:
            assert (tmp_path / ("mpDPFT_Plots-%s.pdf" % n)).exists()


def test_cube_file(tmp_path):
    print("case cube file 3D centres", flush=True)
    pos = [[-1.1, 0.3, -0.2], [0.9, -0.6, 0.4], [0.2, 1.0, 0.5]]
    p = configs.c4(steps=16, edge=10.0, N=4.0, positions=pos, Z=1.5, a=0.7, density_expression=7, tol=1e-10,
                   maxsc=400)
    p.muVec[0] = -1.0
    p.n_interactions = 1
    p.Interactions[0] = 8
    p.CubeFile = 1
    configs.write_input(p, str(tmp_path / "mpDPFT.input"))
    run_dir(tmp_path)
    assert not (tmp_path / "mpDPFT_Den.dat").exists()
    lines = (tmp_path / "mpDPFT_Den_Cube.dat").read_text().splitlines()
    s, dx = p.steps, p.edge / p.steps
    assert lines[0] == "3 0 0 0"
    for a in range(3):
        t = lines[1 + a].split()
        assert int(t[0]) == s and float(t[1 + a]) == pytest.approx(dx, rel=1e-15)
        assert all(float(t[1 + b]) == 0. for b in range(3) if b != a)
    for k in range(3):
        t = [float(v) for v in lines[4 + k].split()]
        assert t == [1.5] + pos[k]
    vals = np.array([float(v) for v in lines[7:]])
    assert vals.size == s ** 3
    ctx = Context(p)
    ctx.solve()
    n = ctx.density()[0].reshape(s + 1, s + 1, s + 1)[:s, :s, :s].reshape(-1)
    ctx.close()
    ref = np.maximum(1e-14, n * dx ** 3)
    assert np.abs(vals - ref).max() <= 1e-15 * ref.max()
    assert vals.min() >= 1e-14


@pytest.mark.parametrize("movieq", [1, 2])
def test_movie_frames(tmp_path, movieq):
    """MovieQ (StoreMovieData mpDPFT.cpp:12317-12329, the export :12660-12691): one
    row per cut point, the cut coordinate then per frame and species the density,
    environment and potential (4 digits); the frame count in the script matches
    the columns, the first column and the last frame's potential are the CutData
    ones, and MovieQ 2 plots on a log scale."""
    print("case movie", movieq, flush=True)
    p = configs.c0(steps=32, edge=8.0, N=6.0, maxsc=60)
    p.MovieQ = movieq
    configs.write_input(p, str(tmp_path / "mpDPFT.input"))
    run_dir(tmp_path)
    rows = [l.split() for l in (tmp_path / "mpDPFT_MovieData.tmp").read_text().split("\n")]
    assert not (tmp_path / "mpDPFT_MovieData.tmp").read_text().endswith("\n")
    assert len(rows) == p.steps + 1
    ncol = len(rows[0])
    assert all(len(r) == ncol for r in rows) and (ncol - 1) % (3 * p.S) == 0
    frames = (ncol - 1) // (3 * p.S)
    assert frames >= 2
    sh = (tmp_path / "mpDPFT_Movie.sh").read_text()
    assert "do for [frame=0:%d]" % (frames - 1) in sh
    assert ("set logscale y" in sh) == (movieq == 2)
    cut = np.loadtxt(tmp_path / "mpDPFT_CutData.dat")
    m = np.array([[float(v) for v in r] for r in rows])
    assert np.allclose(m[:, 0], cut[:, 0], rtol=1e-3, atol=1e-12)
    last = 1 + (frames - 1) * 3 * p.S
    for s in range(p.S):
        assert np.allclose(m[:, last + 3 * s + 2], cut[:, 3 + 3 * s], rtol=1e-3, atol=1e-9)
