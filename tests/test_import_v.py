"""LoadV/ImportV (mpDPFT.cpp:6430-6502): host-side import of a potential file
and its multilinear interpolation onto a new lattice (no GPU)."""
import itertools

import numpy as np
import pytest

from mpdpft_amd import MpdpftError, configs, import_v


def write_vfile(path, axes, fields):
    """The mpDPFT_V.dat layout: x slowest, coordinates then values, each
    followed by a space (ExportData, mpDPFT.cpp:12504-12520)."""
    D = len(axes)
    with open(path, "w") as f:
        for idx in itertools.product(*[range(len(a)) for a in axes]):
            coords = [axes[d][idx[d]] for d in range(D)]
            vals = [fl[idx] for fl in fields]
            f.write("".join("%.16g " % c for c in coords + vals) + "\n")


def lattice(p):
    x1, x2 = -p.edge / 2, p.edge / 2
    return x1 + (x2 - x1) / p.steps * np.arange(p.steps + 1)


def test_bilinear_exact_inside_and_outside(tmp_path):
    # bilinear fields are reproduced exactly by the cell-wise interpolation,
    # also where the new box extends beyond the imported one (extrapolation)
    a = np.linspace(-4, 4, 9)
    X, Y = np.meshgrid(a, a, indexing="ij")
    f0 = 1 + 0.5 * X - 0.25 * Y + 0.1 * X * Y
    f1 = -2 + X * Y
    path = tmp_path / "mpDPFT_V.dat"
    write_vfile(path, [a, a], [f0, f1])
    p = configs.c0(steps=16)
    p.edge = 10.0
    v, warn = import_v(path, p)
    x = lattice(p)
    Xn, Yn = np.meshgrid(x, x, indexing="ij")
    assert np.allclose(v[0], (1 + 0.5 * Xn - 0.25 * Yn + 0.1 * Xn * Yn).ravel(), rtol=0, atol=1e-13)
    assert np.allclose(v[1], (-2 + Xn * Yn).ravel(), rtol=0, atol=1e-13)


def test_trilinear_exact(tmp_path):
    a = np.linspace(-2, 2, 5)
    X, Y, Z = np.meshgrid(a, a, a, indexing="ij")
    def fn(X, Y, Z):
        return 0.3 - X + 2 * Y + 0.5 * Z + 0.2 * X * Y - 0.1 * X * Z + 0.05 * Y * Z + 0.01 * X * Y * Z
    path = tmp_path / "v3.dat"
    write_vfile(path, [a, a, a], [fn(X, Y, Z)])
    p = configs.c1(steps=16, edge=3.0)
    v, warn = import_v(path, p)
    x = lattice(p)
    Xn, Yn, Zn = np.meshgrid(x, x, x, indexing="ij")
    assert np.abs(v[0] - fn(Xn, Yn, Zn).ravel()).max() < 1e-13


def test_same_grid_is_identity(tmp_path):
    # re-importing a file written on the same lattice returns the field itself
    p = configs.c0(steps=32)
    x = lattice(p)
    rng = np.random.default_rng(7)
    fields = [rng.standard_normal((33, 33)) for _ in range(p.S)]
    path = tmp_path / "same.dat"
    write_vfile(path, [x, x], fields)
    v, warn = import_v(path, p)
    for s in range(p.S):
        assert np.abs(v[s] - fields[s].ravel()).max() < 1e-15


def test_warnings_and_errors(tmp_path):
    a = np.linspace(-16, 16, 9)
    X, Y = np.meshgrid(a, a, indexing="ij")
    p = configs.c0(steps=64)  # fine enough that the integral check stays quiet
    # one extra column: the reference's "dimensional mismatch" (import still runs)
    path = tmp_path / "extra.dat"
    write_vfile(path, [a, a], [X * 0 + 1.0, X * 0 + 1.0, X * 0 + 1.0])
    v, warn = import_v(path, p)
    assert warn == 1 and np.allclose(v, 1.0)
    # integrals that differ by more than 5% (a box of another size)
    path = tmp_path / "box.dat"
    b = np.linspace(-2, 2, 9)
    write_vfile(path, [b, b], [X * 0 + 1.0, X * 0 + 1.0])
    q = configs.c0(steps=16)
    q.edge = 32.0
    _, warn = import_v(path, q)
    assert warn == 1
    # a line count that is not n^DIM
    path = tmp_path / "bad.dat"
    with open(path, "w") as f:
        for i in range(10):
            f.write("%d %d 1 1 \n" % (i, i))
    with pytest.raises(MpdpftError):
        import_v(path, p)
    with pytest.raises(MpdpftError):
        import_v(tmp_path / "missing.dat", p)
