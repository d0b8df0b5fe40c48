"""The run driver, its files and the engine's checkpoints (GPU).

  Run          SimulationManager::Run on a generated input: XYZ frames of
               position, velocity and force at step 0 and after every
               output_freq chunk (the last one shorter), restart stages, the
               final restart frame; the f64 engine's frames match the
               oracle's MinimizeSystemEnergy + RunSimulationLoop chunks to the
               printed digits, and both launch cadences write the same bytes.
  fibrenet_run the command-line driver writes the same files and a log.
  checkpoints  a restored checkpoint continues bit for bit (beads, bond
               statuses, break steps, pair sets), in both precisions.
  group force  fn_group_force equals the host sum of the read-back forces and
               the adjoint forward's sample of the same step.
"""
import os
import subprocess

import numpy as np
import pytest

from fibrenet_amd import FibreNetError, Model, Simulation, generate as G, records as R
from helpers import adjoint_net, engine_for, oracle_for, read_beads, run_engine, small_tensile

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fibrenet_amd", "fibrenet_run")
RUN, FREQ = 95, 30
FRAMES = [0, 30, 60, 90, 95]


def _input(tmp, name="run"):
    net = small_tensile(weak=True)
    net.run_steps = RUN
    path = G.write_files(net, str(tmp), name)
    text = open(path).read()
    text = text.replace("output_freq\t%d" % RUN, "output_freq\t%d" % FREQ)
    text = text.replace("output_data\tposition", "output_data\tall")
    text += "minimize_steps 20\nminimize_substeps 10\nminimize_reset yes\n"
    open(path, "w").write(text)
    return path


def _frames(path):
    """(progress, [(type, x, y, z)]) of every frame of an XYZ file."""
    lines = open(path).read().split("\n")
    out, k = [], 0
    while k < len(lines) and lines[k]:
        n = int(lines[k])
        head = lines[k + 1].split("\t")
        assert head[0] == "FRAME"
        rows = [l.split("\t") for l in lines[k + 2:k + 2 + n]]
        out.append((int(head[1]), head, rows))
        k += n + 2
    return out


def _oracle_frames(path):
    m = Model(path)
    orc = oracle_for(m)
    orc.minimize(20, 10, True, m.grid_update_freq)
    frames = {0: orc.beads.copy()}
    done = 0
    while done < RUN:
        n = min(FREQ, RUN - done)
        orc.run(n, m.grid_update_freq)
        done += n
        frames[done] = orc.beads.copy()
    return m, frames


@pytest.fixture(scope="module")
def run_files(tmp_path_factory):
    out = {}
    for fast in (True, False):
        d = tmp_path_factory.mktemp("fast" if fast else "granular")
        path = _input(d)
        sim = Simulation(path, "f64")
        msgs = []
        sim.set_log(msgs.append)
        sim.run(fast=fast)
        out[fast] = (d, path, sim.output_file, msgs)
        sim.close()
    return out


def test_run_frames_match_the_oracle(need_gpu, run_files):
    d, path, stem, msgs = run_files[True]
    m, ref = _oracle_frames(path)
    for suffix, field in (("_coord.xyz", "position"), ("_vel.xyz", "velocity"), ("_force.xyz", "force")):
        frames = _frames(stem + suffix)
        assert [f[0] for f in frames] == FRAMES, suffix
        for progress, head, rows in frames:
            assert head == ["FRAME", str(progress), "%.5f" % m.settings["dt"]] + \
                ["%.5f" % m.settings["box_size"][c] for c in "xyz"]
            assert len(rows) == m.n_beads
            got = np.array([[float(v) for v in r[1:]] for r in rows])
            want = R.vec(ref[progress], field)
            assert np.all(np.abs(got - want) <= 5.01e-6 + 1e-9 * np.abs(want)), (suffix, progress)
            assert [r[0] for r in rows] == [["C1", "C2", "C3"][t] for t in m.info["type"]]
    assert any("finished 90 out of 95" in x for x in msgs) and any("finished after" in x for x in msgs)


def test_run_restart_frames(need_gpu, run_files):
    d, path, stem, msgs = run_files[True]
    m, ref = _oracle_frames(path)
    lines = open(stem + "_restart.res").read().split("\n")
    assert lines[0] == str(m.n_beads) and lines[1].split("\t")[:2] == ["FRAME", str(RUN)]
    got = np.array([[float(v) for v in l.split("\t")[2:]] for l in lines[2:2 + m.n_beads]])
    want = np.hstack([R.vec(ref[RUN], f) for f in ("position", "image_position", "velocity", "force")])
    assert np.all(np.abs(got - want) <= 1e-5 * np.abs(want) + 1e-300)
    # stages passed 0.25, 0.5 and 0.75 at steps 30, 60 and 90, then the final frame
    assert sum("restart data dump successful" in x for x in msgs) == 4
    # the restart re-reads through the loader (velocity_old = velocity, force_old = 0)
    text = open(path).read() + "restart_file\t%s\nrestart_data\tall\n" % os.path.basename(stem + "_restart.res")
    open(os.path.join(d, "again.txt"), "w").write(text)
    again = Model(os.path.join(d, "again.txt"))
    assert np.array_equal(R.vec(again.beads, "velocity_old"), R.vec(again.beads, "velocity"))
    assert np.allclose(R.vec(again.beads, "position"), R.vec(ref[RUN], "position"), rtol=1e-5)


def test_run_cadences_write_the_same_bytes(need_gpu, run_files):
    fast, gran = run_files[True][2], run_files[False][2]
    for suffix in ("_coord.xyz", "_vel.xyz", "_force.xyz", "_restart.res"):
        assert open(fast + suffix, "rb").read() == open(gran + suffix, "rb").read(), suffix


def test_cli_runs_an_input(need_gpu, tmp_path, run_files):
    path = _input(tmp_path, "run")
    r = subprocess.run([CLI, "--precision", "f64", "--log-dir", str(tmp_path), path], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "finished 95 out of 95" in r.stdout and "Closing program." in r.stdout
    logs = [f for f in os.listdir(tmp_path) if f.startswith("log_")]
    assert len(logs) == 1 and "finished 95 out of 95" in open(tmp_path / logs[0]).read()
    ref = run_files[True][2]
    for suffix in ("_coord.xyz", "_vel.xyz", "_force.xyz", "_restart.res"):
        assert open(str(tmp_path / "run_out") + suffix, "rb").read() == open(ref + suffix, "rb").read(), suffix


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_checkpoint_continues_bit_for_bit(need_gpu, tmp_path, precision):
    net = small_tensile(weak=True)
    ck = str(tmp_path / "state.ckpt")
    a, ia = engine_for(net, precision)
    run_engine(a, net, 41)
    a.checkpoint_save(ck)
    run_engine(a, net, 37)
    b, ib = engine_for(net, precision)
    b.checkpoint_load(ck)
    run_engine(b, net, 37)
    assert read_beads(a, ia).tobytes() == read_beads(b, ib).tobytes()
    assert np.array_equal(a.read_buffer(ia["spring"], R.SPRING)["status"],
                          b.read_buffer(ib["spring"], R.SPRING)["status"])
    assert np.array_equal(a.read_buffer(ia["angle"], R.ANGLE)["status"], b.read_buffer(ib["angle"], R.ANGLE)["status"])
    breaks = a.export_breaks(len(net.springs))
    assert np.array_equal(breaks, b.export_breaks(len(net.springs))) and (breaks >= 0).sum() > 0
    assert np.array_equal(a.export_pairs(), b.export_pairs())
    assert a.stats()["steps"] == b.stats()["steps"] == 78
    # rewinding the same engine repeats the continuation
    a.checkpoint_load(ck)
    run_engine(a, net, 37)
    assert read_beads(a, ia).tobytes() == read_beads(b, ib).tobytes()


def test_checkpoint_of_another_network_is_refused(need_gpu, tmp_path):
    ck = str(tmp_path / "state.ckpt")
    a, _ = engine_for(small_tensile(), "f64")
    a.checkpoint_save(ck)
    other, _ = engine_for(small_tensile(n_fibres=30), "f64")
    with pytest.raises(FibreNetError) as e:
        other.checkpoint_load(ck)
    assert e.value.code == -3
    f32, _ = engine_for(small_tensile(), "f32")
    with pytest.raises(FibreNetError):
        f32.checkpoint_load(ck)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_group_force_is_the_clamp_reaction(need_gpu, precision):
    net = adjoint_net(hr_fl=True)
    group = (net.info["type"] == 1).astype(np.int32)
    eng, ids = engine_for(net, precision)
    eng.run_steps(50, 10)
    got = eng.group_force(group)
    f = R.vec(read_beads(eng, ids), "force")[group == 1, 0]
    assert abs(got - f.sum()) <= 1e-12 * max(1.0, np.abs(f).sum())
    ref, _ = engine_for(net, precision)
    s = ref.adjoint_forward(50, group, [49], grid_update_freq=10)
    assert s[-1] == got
    assert eng.group_force(np.zeros_like(group)) == 0.0
