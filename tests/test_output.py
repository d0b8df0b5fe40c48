"""Output formats and the command-line driver, on the host: the restart
writer's text and its round trip through the restart reader by the
reference's rules (SimulationManager.cpp:1670-1823, 1855-1928), the XYZ
frames (:1937-2061), and fibrenet_run's argument handling and clean
failures (no device in this container)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from fibrenet_amd import FibreNetError, Model, records as R

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
CLI = os.path.join(os.path.dirname(HERE), "fibrenet_amd", "fibrenet_run")
FIELDS = ("position", "image_position", "velocity", "force")


def _g(x):
    """The text of a double at an ostream's default format (%g, 6 digits)."""
    return "%g" % x


def _scrambled(m, seed):
    """The README model's beads with values of every magnitude and sign."""
    rs = np.random.default_rng(seed)
    b = m.beads.copy()
    n = len(b)
    for f in FIELDS:
        v = rs.normal(0.0, 1.0, (n, 3)) * 10.0 ** rs.integers(-7, 7, (n, 3))
        v[0] = (0.0, -0.0, 1234567.891)
        R.set_vec(b, f, v)
    R.set_vec(b, "velocity_old", rs.normal(size=(n, 3)))
    R.set_vec(b, "force_old", rs.normal(size=(n, 3)))
    return b


def _input_with(tmp, lines):
    src = open(os.path.join(GOLD, "readme_input.txt")).read()
    shutil.copy(os.path.join(GOLD, "readme_model.txt"), tmp / "readme_model.txt")
    p = tmp / "input.txt"
    p.write_text(src + "\n" + "\n".join(lines) + "\n")
    return str(p)


def test_restart_writer_text(tmp_path):
    m = Model(os.path.join(GOLD, "readme_input.txt"))
    b = _scrambled(m, 1)
    out = tmp_path / "frame.res"
    m.write_restart(b, 50000, str(out))
    lines = out.read_text().split("\n")
    assert lines[0] == str(m.n_beads)
    s = m.settings
    assert lines[1] == "\t".join(["FRAME", "50000", _g(s["dt"]), _g(s["box_size"]["x"]), _g(s["box_size"]["y"]),
                                  _g(s["box_size"]["z"])])
    assert lines[1] == "FRAME\t50000\t0.0001\t111\t41\t192"
    for i in range(m.n_beads):
        want = [""] + ["C1"] + [_g(b[f][c][i]) for f in FIELDS for c in "xyz"]
        assert lines[2 + i] == "\t".join(want), i
    assert lines[2 + m.n_beads] == "" and len(lines) == m.n_beads + 3
    assert "1.23457e+06" in lines[2] and "\t-0\t" in lines[2]


def test_restart_round_trip_follows_the_reference_reader(tmp_path):
    """A written restart frame read back by restart_file / restart_data all:
    positions, images, velocities and forces at 6 significant digits,
    velocity_old = velocity and force_old = 0 (the restart is lossy, Q26)."""
    m = Model(os.path.join(GOLD, "readme_input.txt"))
    b = _scrambled(m, 2)
    m.write_restart(b, 7, str(tmp_path / "state.res"))
    back = Model(_input_with(tmp_path, ["restart_file\tstate.res", "restart_data\tall"]))
    rd = lambda a: np.array([float(_g(x)) for x in a.ravel()]).reshape(a.shape)
    for f in FIELDS:
        assert np.array_equal(R.vec(back.beads, f), rd(R.vec(b, f))), f
    assert np.array_equal(R.vec(back.beads, "velocity_old"), R.vec(back.beads, "velocity"))
    assert np.all(R.vec(back.beads, "force_old") == 0.0)
    # only the requested fields are restored
    pos_only = Model(_input_with(tmp_path, ["restart_file\tstate.res", "restart_data\tposition"]))
    assert np.array_equal(R.vec(pos_only.beads, "position"), rd(R.vec(b, "position")))
    assert np.array_equal(R.vec(pos_only.beads, "velocity"), R.vec(m.beads, "velocity"))


def test_restart_of_another_model_is_rejected(tmp_path):
    m = Model(os.path.join(GOLD, "readme_input.txt"))
    m.write_restart(m.beads[:], 0, str(tmp_path / "state.res"))
    text = (tmp_path / "state.res").read_text().split("\n")
    (tmp_path / "short.res").write_text("\n".join(text[:-2]) + "\n")        # one bead missing
    with pytest.raises(FibreNetError, match="incorrect number of beads"):
        Model(_input_with(tmp_path, ["restart_file\tshort.res", "restart_data\tall"]))
    (tmp_path / "typed.res").write_text("\n".join(t.replace("\tC1\t", "\tC2\t", 1) if i == 2 else t
                                                  for i, t in enumerate(text)))
    with pytest.raises(FibreNetError, match="type mismatch"):
        Model(_input_with(tmp_path, ["restart_file\ttyped.res", "restart_data\tall"]))


def test_xyz_frames_append(tmp_path):
    m = Model(os.path.join(GOLD, "readme_input.txt"))
    b = _scrambled(m, 3)
    p = str(tmp_path / "run_coord.xyz")
    m.write_xyz(m.beads, 0, "position", p)
    m.write_xyz(b, 50000, "position", p)
    m.write_xyz(b, 50000, "force", str(tmp_path / "run_force.xyz"))
    lines = open(p).read().split("\n")
    n = m.n_beads
    assert lines[0] == str(n) and lines[1] == "FRAME\t0\t0.00010\t111.00000\t41.00000\t192.00000"
    assert lines[2] == "C1\t61.98000\t34.62000\t73.68000"
    assert lines[n + 2] == str(n) and lines[n + 3].startswith("FRAME\t50000\t")
    for i in range(n):
        assert lines[n + 4 + i] == "C1\t" + "\t".join("%.5f" % b["position"][c][i] for c in "xyz")
    f = open(str(tmp_path / "run_force.xyz")).read().split("\n")
    assert f[2] == "C1\t" + "\t".join("%.5f" % b["force"][c][0] for c in "xyz")


def test_xyz_frames_hold_the_exported_types_only(tmp_path):
    m = Model(_input_with(tmp_path, []))
    assert set(m.info["type"]) == {0}
    lines_all = tmp_path / "a.xyz"
    m.write_xyz(m.beads, 0, "velocity", str(lines_all))
    assert open(lines_all).readline().strip() == str(m.n_beads)
    src = open(os.path.join(GOLD, "readme_input.txt")).read()
    src = "\n".join(l for l in src.split("\n") if not l.startswith("output_type")) + "\noutput_type\tC2\n"
    (tmp_path / "c2.txt").write_text(src)
    m2 = Model(str(tmp_path / "c2.txt"))
    m2.write_xyz(m2.beads, 0, "velocity", str(tmp_path / "b.xyz"))
    assert open(tmp_path / "b.xyz").read().split("\n")[:2] == ["0", "FRAME\t0\t0.00010\t111.00000\t41.00000\t192.00000"]


def _cli(*args, cwd=None):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120, cwd=cwd)


def test_cli_usage_errors():
    assert os.path.exists(CLI), "fibrenet_run is not built"
    for args in ((), ("--precision", "f16", "x.txt"), ("--cadence", "slow", "x.txt"), ("--bogus", "x.txt"),
                 ("--device",), ("a.txt", "b.txt")):
        r = _cli(*args)
        assert r.returncode == 2, args
        assert "usage: fibrenet_run" in r.stderr
    assert _cli("--help").returncode == 0


def test_cli_rejects_a_missing_input_and_logs_it(tmp_path):
    r = _cli("--log-dir", str(tmp_path), str(tmp_path / "missing.txt"))
    assert r.returncode == 1
    assert "unable to open input file" in r.stdout
    logs = [f for f in os.listdir(tmp_path) if f.startswith("log_") and f.endswith(".txt")]
    assert len(logs) == 1
    text = open(tmp_path / logs[0]).read()
    assert text.startswith("Starting log file at:") and "unable to open input file" in text


def test_cli_without_a_device_fails_cleanly(tmp_path):
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a HIP device is visible")
    r = _cli("--no-log-file", os.path.join(GOLD, "readme_input.txt"), cwd=str(tmp_path))
    assert r.returncode == 1 and "unable to open the HIP device" in r.stdout
    assert not os.listdir(tmp_path)
