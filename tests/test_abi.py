"""The C ABI library: loads without a GPU, exports every declared symbol,
matches the ctypes layouts, parses mpDPFT.input files, fails loudly."""
import ctypes
import os
import re

import pytest

import mpdpft_amd
from mpdpft_amd import _abi, configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "mpdpft.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mpdpft_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = _abi.load_library()
    names = declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(names) == sorted(_abi.EXPORTS)


def test_layouts_and_version():
    lib = _abi.load_library()
    assert lib.mpdpft_abi_version() == 5
    assert lib.mpdpft_params_sizeof() == ctypes.sizeof(_abi.Params)
    assert lib.mpdpft_result_sizeof() == ctypes.sizeof(_abi.Result)


def test_defaults_are_reference_hardcoded_values():
    p = mpdpft_amd.default_params()
    assert p.TaperOff == 1 and p.RegularizeCoulombSingularity == 1
    assert p.SmearingExtentDx == 6.0 and p.OutPrecision == 16 and p.MaxIteration == 4
    assert p.HartreeMode == 0 and p.ConvergenceMode == 0 and p.Precision == 0


def test_parse_round_trip(tmp_path):
    p = configs.c1(steps=64)
    p.n_hyperbox = 2
    p.TaskHyperBox[0], p.TaskHyperBox[1] = 1.5, 4.0
    p.n_centres = 2
    for c, row in enumerate([(1.0, 2.0, 3.0, 1.0, 0.6), (-1.0, 0.5, 0.0, 2.0, 0.4)]):
        for k, v in enumerate(row):
            p.centres[c][k] = v
    f = tmp_path / "mpDPFT.input"
    configs.write_input(p, f)
    q = mpdpft_amd.parse_input(f)
    for name, _ in _abi.Params._fields_:
        a, b = getattr(p, name), getattr(q, name)
        if hasattr(a, "__len__"):
            if name == "centres":
                assert [list(r) for r in a[:2]] == [list(r) for r in b[:2]]
            else:
                assert list(a) == list(b), name
        else:
            assert a == b, name


def test_parse_reference_style_lines(tmp_path):
    """Values followed by tab-separated // comments; open vectors stop at the comment."""
    p = configs.c0()
    f = tmp_path / "mpDPFT.input"
    configs.write_input(p, f, extensions=False)
    lines = open(f).read().splitlines()
    lines[16] = "0.5 0.5 0.25 0.25\t\t\t//mpp\t\t\t\t--- vector<double> of multi-purpose parameters"
    lines[26] = "1 8 -12\t\t\t//Interactions\t\t\t--- vector<int> of accumulative interaction types"
    open(f, "w").write("\n".join(lines) + "\n")
    q = mpdpft_amd.parse_input(f)
    assert q.n_mpp == 4 and list(q.mpp[:4]) == [0.5, 0.5, 0.25, 0.25]
    assert q.n_interactions == 3 and list(q.Interactions[:3]) == [1, 8, -12]
    assert q.HartreeMode == 0  # no extension lines: reference behaviour


def test_parse_rejects_short_or_bad_files(tmp_path):
    f = tmp_path / "mpDPFT.input"
    f.write_text("0\n0\n0\n")
    with pytest.raises(mpdpft_amd.MpdpftError):
        mpdpft_amd.parse_input(f)
    configs.write_input(configs.c0(), f)
    f.write_text(f.read_text() + "7 //NoSuchKey\n")
    with pytest.raises(mpdpft_amd.MpdpftError):
        mpdpft_amd.parse_input(f)
    with pytest.raises(mpdpft_amd.MpdpftError):
        mpdpft_amd.parse_input(tmp_path / "missing.input")


def test_create_fails_loudly_on_invalid_params():
    p = configs.c0()
    p.DensityExpression = 11  # n7: not on this path
    with pytest.raises(mpdpft_amd.MpdpftError) as e:
        mpdpft_amd.Context(p)
    assert e.value.code in (_abi.EINVAL, _abi.ENODEV)


def test_cli_fails_loudly_without_a_gpu(tmp_path):
    """Without a HIP device the drop-in CLI writes no result files: it exits 1 and
    leaves the reason in mpDPFT_Control.dat (there is no CPU fallback)."""
    import subprocess
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    configs.write_input(configs.c0(steps=32), tmp_path / "mpDPFT.input")
    cli = os.path.join(os.path.dirname(mpdpft_amd.__file__), "bin", "mpdpft_hip")
    out = subprocess.run([cli, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 1
    assert "no HIP device" in (tmp_path / "mpDPFT_Control.dat").read_text()
    assert not (tmp_path / "mpDPFT_Den.dat").exists() and not (tmp_path / "mpDPFT_V.dat").exists()



@pytest.mark.parametrize("field,value", [("n_interactions", 17), ("n_schedule", 17), ("n_regularize", 17),
                                         ("n_ekin", 17), ("n_Mpp", 17), ("n_mpp", 65), ("n_centres", 257),
                                         ("n_interactions", -1), ("abi_version", 1), ("S", 9),
                                         ("n_hyperbox", 17)])
def test_counts_beyond_their_arrays_are_rejected(tmp_path, field, value):
    """A list count larger than its fixed array (or a params record of another
    layout version) is EINVAL at every entry point that takes params, before
    any device work: mpdpft_create, mpdpft_batch_create, mpdpft_import_v."""
    p = configs.c0(steps=16)
    setattr(p, field, value)
    with pytest.raises(mpdpft_amd.MpdpftError) as e:
        mpdpft_amd.Context(p)
    assert e.value.code == _abi.EINVAL
    assert "out of range" in str(e.value) or "abi_version" in str(e.value)
    with pytest.raises(mpdpft_amd.MpdpftError) as e:
        mpdpft_amd.Batch([configs.c0(steps=16), p])
    assert e.value.code == _abi.EINVAL and "instance 1" in str(e.value)
    v = tmp_path / "mpDPFT_V.dat"
    v.write_text("0 0 1 1 \n")
    with pytest.raises(mpdpft_amd.MpdpftError) as e:
        mpdpft_amd.import_v(v, p)
    assert e.value.code == _abi.EINVAL
