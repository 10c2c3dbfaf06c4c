"""TaskType 2 / 3 input checks of the task loop (mpdpft_run_dir), made before
any device work, so they run without a GPU: TaskType 2 varies muVec and needs
RelAcc = 0 (no mu search in EnforceConstraints, mpDPFT.cpp:4578), TaskType 3
varies the abundances and needs RelAcc > 0, both need one TaskHyperBox pair per
species (GetTaskHyperBox :610-618), and task types without a loop here are
rejected by name."""
import pytest

import mpdpft_amd
from mpdpft_amd import _abi, configs


def run(tmp_path, p):
    configs.write_input(p, str(tmp_path / "mpDPFT.input"))
    with pytest.raises(mpdpft_amd.MpdpftError) as e:
        mpdpft_amd.run_dir(tmp_path)
    assert e.value.code == _abi.EINVAL
    for name in ("mpDPFT_ControlTask.dat", "mpDPFT_Control.dat"):
        f = tmp_path / name
        if f.exists():
            return f.read_text()
    return ""


def base(task_type, relacc, box=(3.0, 6.0)):
    p = configs.c0(steps=16)
    p.S = 1
    p.TaskType = task_type
    p.RelAcc = relacc
    p.n_hyperbox = len(box)
    for i, v in enumerate(box):
        p.TaskHyperBox[i] = v
    return p


def test_task3_needs_a_mu_search(tmp_path):
    assert "RelAcc must be positive" in run(tmp_path, base(3, 0.0))


def test_task2_needs_fixed_mu(tmp_path):
    assert "RelAcc must be 0" in run(tmp_path, base(2, 1e-10, (-0.2, 0.1)))


def test_one_hyperbox_pair_per_species(tmp_path):
    assert "one TaskHyperBox pair per species" in run(tmp_path, base(3, 1e-10, (3.0, 6.0, 1.0, 2.0)))


def test_empty_hyperbox(tmp_path):
    assert "is empty" in run(tmp_path, base(3, 1e-10, (6.0, 3.0)))


@pytest.mark.parametrize("task_type", [4, 11, 44])
def test_unbuilt_task_type(tmp_path, task_type):
    assert "TaskType %d is not built" % task_type in run(tmp_path, base(task_type, 1e-10))


def test_hyperbox_round_trips_through_the_input_file(tmp_path):
    p = base(3, 1e-10, (2.5, 7.25))
    configs.write_input(p, str(tmp_path / "mpDPFT.input"))
    q = mpdpft_amd.parse_input(tmp_path / "mpDPFT.input")
    assert q.n_hyperbox == 2 and list(q.TaskHyperBox[:2]) == [2.5, 7.25]


def test_parameter_scans_need_2d_and_a_reference_table(tmp_path):
    p = base(7, 1e-10, (0.3, 0.7))
    p.TaskParameterCount = 3
    p.DIM = 3
    assert "2D reference density" in run(tmp_path, p)
    p.DIM = 2
    assert "TabFunc_ReferenceDensity.dat not found" in run_code(tmp_path, p, _abi.EIO)
    (tmp_path / "TabFunc_ReferenceDensity.dat").write_text("1\n2\n3\n")
    assert "n x n values" in run_code(tmp_path, p, _abi.EIO)
    p.TaskParameterCount = 1
    assert "TaskParameterCount >= 2" in run(tmp_path, p)


def run_code(tmp_path, p, code):
    configs.write_input(p, str(tmp_path / "mpDPFT.input"))
    with pytest.raises(mpdpft_amd.MpdpftError) as e:
        mpdpft_amd.run_dir(tmp_path)
    assert e.value.code == code
    return (tmp_path / "mpDPFT_ControlTask.dat").read_text()
