"""Input/model loaders: the product's C++ loader and the oracle's restatement
against the README fixtures (expected values written out by hand from
README.md:107-120 and :565-606 and the rules of SimulationManager.cpp), and
against each other on generated files and on the reference's quirks."""
import math
import os
import shutil

import numpy as np
import pytest

from fibrenet_amd import FibreNetError, Model, generate as G, records as R
from oracle import loader as OL

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
F32 = lambda x: float(np.float32(x))


def both(path):
    return Model(path), OL.load_input(path)


def same_records(m, o):
    for k in ("beads", "info", "springs", "angles", "pairs"):
        assert getattr(m, k).tobytes() == getattr(o, k).tobytes(), k
    assert m.settings.tobytes() == o.settings.tobytes()
    assert np.array_equal(m.cell_nh, o.cell_nh)
    assert (m.grid_update_freq, m.run_steps, m.minimize_steps, m.minimize_substeps, m.minimize_reset) == (
        o.grid_update_freq, o.run_steps, o.minimize_steps, o.minimize_substeps, o.minimize_reset)


def test_readme_fixture():
    m, o = both(os.path.join(GOLD, "readme_input.txt"))
    same_records(m, o)
    s = m.settings
    # README input: damping 0.03 is stored negated; lambda 0.5 * step 1e-4
    assert s["damping"] == -0.03 and s["dt"] == 1e-4 and s["lambda_dt"] == 0.5 * 1e-4
    assert s["half_dt"] == 1e-4 * 0.5 and s["half_dt2"] == 1e-4 * 1e-4 * 0.5
    assert s["cutoff"] == 1.15 and s["cutoff_sq"] == 1.15 * 1.15
    # 'box_wall_re' is a typo for box_wall_rep, so the box becomes finite (Q11)
    assert s["bc_type"] == 1 and s["wall_contact_stiffness"] == 0.0
    assert (s["cells_x"], s["cells_y"], s["cells_z"]) == (88, 32, 153)
    assert s["grid_step"]["x"] == 111.0 / 88 and s["grid_step"]["z"] == 192.0 / 153
    assert s["half_box_sq"] == 20.5 * 20.5 and s["types_num"] == 3
    assert (m.run_steps, m.minimize_steps, m.minimize_substeps, m.minimize_reset) == (800000, 400000, 1000, True)
    # st pairs (tokens 2000.0 500.0 1.00 1.03): c1=t5, c2=t6, c3=t7-t6, c4=t7^2, c5=t4, c6=2*t6-t7
    assert len(m.pairs) == 6 and np.all(m.pairs["type"] == 3)
    assert np.all(m.pairs["c1"] == 500.0) and np.all(m.pairs["c2"] == 1.0)
    assert np.all(m.pairs["c3"] == 1.03 - 1.0) and np.all(m.pairs["c4"] == 1.03 * 1.03)
    assert np.all(m.pairs["c5"] == 2000.0) and np.all(m.pairs["c6"] == 2 * 1.0 - 1.03)
    assert list(zip(m.pairs["p1"], m.pairs["p2"])) == [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    # beads
    xs = [61.98, 62.48, 62.98, 63.47, 63.97]
    zs = [73.68, 73.73, 73.77, 73.81, 73.85]
    b, q = m.beads, m.info
    assert list(b["position"]["x"]) == xs and list(b["image_position"]["z"]) == zs
    assert np.all(b["position"]["y"] == 34.62)
    assert np.all(q["type"] == 0) and np.all(q["tag"] == 1) and np.all(q["loc_id"] == 0)
    assert np.all(q["radius"] == 0.5) and np.all(q["mass"] == 1.0) and np.all(q["mass_inv"] == 1.0)
    assert np.all(q["lambda_dt_mass_inv"] == 0.5e-4) and np.all(q["half_dt2_mass_inv"] == 1e-4 * 1e-4 * 0.5)
    assert np.all(q["skin_radius"] == 0.6) and np.all(q["skin_radius_sq"] == 0.6 * 0.6)  # modif_bead skin_radius
    # springs and angles
    assert list(m.springs["p1"]) == [0, 1, 2, 3] and list(m.springs["p2"]) == [1, 2, 3, 4]
    assert np.all(m.springs["rest_length"] == 0.5) and np.all(m.springs["stiffness"] == 59654.11)
    assert np.all(m.springs["type"] == 0) and np.all(m.springs["status"] == 1)
    a = m.angles
    assert list(a["b1"]) == [0, 1, 2] and list(a["b3"]) == [2, 3, 4] and list(a["s2"]) == [1, 2, 3]
    assert list(a["angle"]) == [F32("3.13"), F32("3.14"), F32("3.14")]          # std::stof (Q6)
    assert a["angle"][0] != 3.13 and np.all(a["c1"] == 1864.19)


def _write(tmp, name, text):
    p = os.path.join(tmp, name)
    with open(p, "w") as fh:
        fh.write(text)
    return p


def _readme_variant(tmp, edit_input=None, edit_model=None, extra_files=None):
    inp = open(os.path.join(GOLD, "readme_input.txt")).read()
    mod = open(os.path.join(GOLD, "readme_model.txt")).read()
    if edit_input:
        inp = edit_input(inp)
    if edit_model:
        mod = edit_model(mod)
    _write(tmp, "readme_model.txt", mod)
    for k, v in (extra_files or {}).items():
        _write(tmp, k, v)
    return _write(tmp, "in.txt", inp)


def test_restart_file(tmp_path):
    rows = [l for l in open(os.path.join(GOLD, "readme_restart.res")).read().splitlines() if l.count("C1")][:5]
    res = "5\nFRAME\t0\t0.0001\t111\t41\t192\n" + "\n".join(rows) + "\n"
    p = _readme_variant(str(tmp_path), edit_input=lambda s: s + "restart_file r.res\nrestart_data all\n",
                        extra_files={"r.res": res})
    m, o = both(p)
    same_records(m, o)
    first = rows[0].split()
    b = m.beads[0]
    assert b["position"]["x"] == float(first[1]) and b["image_position"]["x"] == float(first[4])
    assert b["velocity"]["z"] == float(first[9]) and b["velocity_old"]["z"] == float(first[9])
    assert b["force"]["x"] == float(first[10]) and b["force_old"]["x"] == 0.0


QUIRKS = {
    # std::stoi of "1e3" is 1 (Q5); HR_FL springs need exactly 9 tokens
    "hr_fl_stoi": (None, lambda s: s.replace("SPRING\tHR\t0\t0\t1\t0.50\t59654.11",
                                              "SPRING\tHR_FL\t0\t0\t1\t0.50\t59654.11\t1e3\t2.9"), True),
    # v_old takes vx on all three axes (Q4)
    "vold_vx": (None, lambda s: s.replace("61.98\t34.62\t73.68\t0.00\t0.00\t0.00", "61.98\t34.62\t73.68\t1.5\t-2\t3"),
                True),
    "no_lambda": (lambda s: s.replace("lambda", "#lambda"), None, False),
    "output_freq_gt_run": (lambda s: s.replace("output_freq         50000", "output_freq 900000"), None, False),
    "missing_pair": (lambda s: s.replace("pair                st\tC2\tC3", "#pair st C2 C3"), None, False),
    "bad_pair_count": (lambda s: s.replace("pair                st\tC2\tC3\t2000.0\t500.0\t1.00\t1.03",
                                           "pair st C2 C3 2000.0 500.0 1.00"), None, True),
    "box_too_small": (lambda s: s.replace("box_unit_cell       1.25", "box_unit_cell 15"), None, False),
    "bead_outside_box": (None, lambda s: s.replace("61.98\t34.62\t73.68\t0.00", "161.98\t34.62\t73.68\t0.00"), False),
    "no_springs": (None, lambda s: "\n".join(l for l in s.splitlines() if not l.startswith("SPRING")), False),
    "pbc_with_wall": (lambda s: s.replace("box_type            finite", "box_type pbc").replace("box_wall_re ",
                                                                                                "box_wall_rep "),
                      None, True),
    "grid_freq": (lambda s: s + "grid_update_freq 6\n", None, True),
    "gravity": (lambda s: s + "gravity 0 0 -9.81\n", None, True),
    "modif_mass_velocity": (lambda s: s + "modif_bead id 2 mass 0.0\nmodif_bead tag 1 velocity 0.1 0 0\n"
                                          "modif_spring id 1 rest_length 0.7\nmodif_spring tag 0 stiffness 5\n",
                            None, True),
    "modif_unknown": (lambda s: s + "modif_bead id 2 charge 1\n", None, False),
    "hr_spring_extra_tokens": (None, lambda s: s.replace("59654.11\r\nSPRING", "59654.11\t7\t8\r\nSPRING", 1), True),
    "angle_bad_count": (None, lambda s: s.replace("3.13\t1864.19", "3.13"), False),
}


@pytest.mark.parametrize("case", sorted(QUIRKS))
def test_quirks_agree(tmp_path, case):
    ei, em, ok = QUIRKS[case]
    p = _readme_variant(str(tmp_path), ei, em)
    if not ok:
        with pytest.raises(FibreNetError):
            Model(p)
        with pytest.raises(OL.LoadError):
            OL.load_input(p)
        return
    m, o = both(p)
    same_records(m, o)
    if case == "hr_fl_stoi":
        assert m.springs["type"][0] == 1 and m.springs["c1"][0] == 1.0 and m.springs["c2"][0] == 2.0
    if case == "vold_vx":
        v = m.beads["velocity_old"][0]
        assert (v["x"], v["y"], v["z"]) == (1.5, 1.5, 1.5) and m.beads["velocity"]["y"][0] == -2.0
    if case == "bad_pair_count":
        assert m.pairs["type"][4] == 0        # the reference keeps an uninitialised entry; here zero
    if case == "pbc_with_wall":
        assert m.settings["bc_type"] == 0 and m.settings["wall_contact_stiffness"] == 100.0
    if case == "modif_mass_velocity":
        assert m.info["mass"][2] == 0.0 and m.info["lambda_dt_mass_inv"][2] == 0.0
        assert np.all(m.beads["velocity_old"]["x"] == 0.1) and m.springs["rest_length"][1] == 0.7
        assert np.all(m.springs["stiffness"] == 5.0)


@pytest.mark.parametrize("workload", ["periodic_hs", "tensile_lj"])
def test_generator_matches_loaders(tmp_path, workload):
    net = (G.periodic_hs(n_fibres=30, n_beads=12, box=(24.0, 24.0, 10.0)) if workload == "periodic_hs"
           else G.tensile_lj(n_fibres=40, n_beads=15, box=(50.0, 50.0, 20.0)))
    p = G.write_files(net, str(tmp_path), workload)
    m, o = both(p)
    same_records(m, o)
    for k in ("beads", "info", "springs", "angles", "pairs"):
        assert getattr(m, k).tobytes() == getattr(net, k).tobytes(), k
    assert m.settings.tobytes() == net.settings.tobytes()
    assert np.array_equal(m.cell_nh, net.cell_nh)


def test_generator_geometry():
    net = G.tensile_lj(n_fibres=200, n_beads=30, box=(100.0, 100.0, 40.0), d_min=0.6)
    p = R.vec(net.beads, "position")
    assert p.min() >= 0.0 and np.all(p <= [100.0, 100.0, 40.0])
    f = p.reshape(200, 30, 3)
    assert np.allclose(np.linalg.norm(np.diff(f, axis=1), axis=2), 1.0)
    # no two beads of different fibres closer than d_min
    from scipy.spatial import cKDTree
    pairs = cKDTree(p).query_pairs(0.6, output_type="ndarray")
    assert np.all(pairs[:, 0] // 30 == pairs[:, 1] // 30)
