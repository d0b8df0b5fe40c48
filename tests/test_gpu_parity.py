"""Parity of the HIP engine with the oracle (GPU).

Tolerances, stated per build:
  f64  - every bead field within 1e-9 of the oracle, relative to the field's
         largest magnitude (floor 1.0).  The builds differ only in the order
         contact forces are summed and in acos/pow ulps (measured: <= 3e-14).
  f32  - one step from a shared state: forces within 2e-5 of the largest force
         (fp32 arithmetic, ulp 6e-8, through LJ powers).  After 200 steps:
         positions/images within 1e-5 absolute (coordinates are O(10)),
         velocities within 2e-3 and forces within 1e-2 of the field's largest
         magnitude - the fp32 trajectory drifts from the fp64 one by ~1e-6 in
         position and LJ near-contacts (r ~ 0.6 sigma) turn that into force.
Cell indices, contact-pair sets and spring-break steps are compared exactly.
The fused (hipGraph) path and the per-kernel path are compared bit for bit.
"""
import copy

import numpy as np
import pytest

from fibrenet_amd import generate as G, records as R
from helpers import (crosslinked, dense_lj, engine_for, field_errors, oracle_for, read_beads, run_engine,
                     small_periodic, small_tensile)

pytestmark = pytest.mark.gpu

TOL = {"f64": {f: 1e-9 for f in ("position", "image_position", "velocity", "force", "velocity_old",
                                 "force_old")},
       "f32": {"position": 1e-5, "image_position": 1e-5, "velocity": 2e-3, "force": 1e-2,
               "velocity_old": 2e-3, "force_old": 1e-2}}

WORKLOADS = {"periodic_hs": lambda: small_periodic(), "tensile_lj": lambda: small_tensile(),
             "tensile_lj_breaking": lambda: small_tensile(weak=True), "crosslinked": lambda: crosslinked(),
             "dense_lj": lambda: dense_lj()}

# Networks whose bonds follow the chain layout (fast path: bonds evaluated in
# the bead kernel by wave shuffles) and one that does not (bond kernels and
# incidence slots).
CHAIN = {"periodic_hs": 1, "tensile_lj": 1, "tensile_lj_breaking": 1, "crosslinked": 0, "dense_lj": 1}


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("workload", sorted(WORKLOADS))
def test_trajectory_matches_oracle(need_gpu, workload, precision):
    net = WORKLOADS[workload]()
    steps = 200
    eng, ids = engine_for(net, precision)
    run_engine(eng, net, steps, fast=True)
    got = read_beads(eng, ids)
    orc = oracle_for(net)
    orc.run(steps, net.grid_update_freq)
    errs = field_errors(got, orc.beads)
    print(workload, precision, errs)
    for f, e in errs.items():
        assert e <= TOL[precision][f], (f, e)
    sp = eng.read_buffer(ids["spring"], R.SPRING)
    if precision == "f64":
        assert np.array_equal(sp["status"], orc.springs["status"])
        assert np.array_equal(eng.export_breaks(len(net.springs)), orc.break_step)


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("workload", sorted(WORKLOADS))
def test_fused_equals_per_kernel(need_gpu, workload, precision):
    net = WORKLOADS[workload]()
    a, ia = engine_for(net, precision)
    b, ib = engine_for(net, precision)
    for n in (1, 7, 40):                      # uneven chunks exercise the chunk boundaries
        run_engine(a, net, n, fast=True)
        run_engine(b, net, n, fast=False)
    ga, gb = read_beads(a, ia), read_beads(b, ib)
    assert ga.tobytes() == gb.tobytes()
    assert np.array_equal(a.read_buffer(ia["spring"], R.SPRING)["status"],
                          b.read_buffer(ib["spring"], R.SPRING)["status"])
    assert np.array_equal(a.export_pairs(), b.export_pairs())


@pytest.mark.parametrize("workload", sorted(WORKLOADS))
def test_chain_layout_detected(need_gpu, workload):
    net = WORKLOADS[workload]()
    eng, _ = engine_for(net, "f64")
    assert eng.get_option("chain_layout") == CHAIN[workload]


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("workload", ["tensile_lj_breaking", "periodic_hs"])
def test_chain_kernel_equals_slot_path(need_gpu, workload, precision):
    """The in-wave chain bead kernel and the bond-kernel + slot path give the
    same bytes (same terms, same summation order)."""
    net = WORKLOADS[workload]()
    a, ia = engine_for(net, precision)
    b, ib = engine_for(net, precision)
    b.set_option("chain", 0)
    for n in (3, 150):
        run_engine(a, net, n, fast=True)
        run_engine(b, net, n, fast=True)
    assert read_beads(a, ia).tobytes() == read_beads(b, ib).tobytes()
    for kind, rec in (("spring", R.SPRING), ("angle", R.ANGLE)):
        assert a.read_buffer(ia[kind], rec).tobytes() == b.read_buffer(ib[kind], rec).tobytes()
    assert np.array_equal(a.export_breaks(len(net.springs)), b.export_breaks(len(net.springs)))


def _inexact_failure_stress():
    net = small_tensile(weak=True)
    net.springs["c2"] = net.springs["c2"] + 0.1          # not exact in float: still parameter sets
    return net


def _status_two():
    net = small_tensile(weak=True)
    net.springs["status"][7] = 2                          # inactive, but not a 0/1 status bit
    return net


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("case", ["sets", "sets_inexact", "indexed"])
def test_chain_parameter_sets_and_indexed_loads_equal_slot_path(need_gpu, case, precision):
    """The chain kernel reads bond parameters through one word per bead (set
    codes and status bits) when every bond status is 0 or 1 and each kind has
    at most 32767 distinct parameter sets, else through the bond indices;
    both give the slot path's bytes."""
    net = {"sets": lambda: small_tensile(weak=True), "sets_inexact": _inexact_failure_stress,
           "indexed": _status_two}[case]()
    a, ia = engine_for(net, precision)
    b, ib = engine_for(net, precision)
    assert a.get_option("chain_sets") == (0 if case == "indexed" else 1)
    b.set_option("chain", 0)
    for n in (5, 200):
        run_engine(a, net, n, fast=True)
        run_engine(b, net, n, fast=True)
    assert read_beads(a, ia).tobytes() == read_beads(b, ib).tobytes()
    sa = a.read_buffer(ia["spring"], R.SPRING)
    assert sa.tobytes() == b.read_buffer(ib["spring"], R.SPRING).tobytes()
    assert np.sum(sa["status"] == 0) > 0                  # springs broke on the way


def test_chain_status_copies_follow_breaks_on_other_paths(need_gpu):
    """Springs broken while the slot path runs are broken for the chain kernel
    that runs next: alternating paths gives the bytes of the chain path alone."""
    net = small_tensile(weak=True)
    a, ia = engine_for(net, "f64")
    b, ib = engine_for(net, "f64")
    for k in range(6):
        a.set_option("chain", k % 2)
        run_engine(a, net, 40, fast=True)
        run_engine(b, net, 40, fast=True)
    assert read_beads(a, ia).tobytes() == read_beads(b, ib).tobytes()
    sa = a.read_buffer(ia["spring"], R.SPRING)
    assert sa.tobytes() == b.read_buffer(ib["spring"], R.SPRING).tobytes()
    assert np.sum(sa["status"] == 0) > 0


@pytest.mark.parametrize("verlet", ["on", "off"])
@pytest.mark.parametrize("workload", sorted(WORKLOADS))
def test_flat_search_equals_per_bead_search(need_gpu, workload, verlet):
    """The wave-flattened neighbour search lists the same pairs as the
    per-bead search, for the exact rule and for kept candidates, in finite
    and periodic boxes (rows that wrap)."""
    net = WORKLOADS[workload]()
    a, ia = engine_for(net, "f64")
    b, ib = engine_for(net, "f64")
    assert a.get_option("flat_search") == 1
    b.set_option("flat_search", 0)
    for e in (a, b):
        e.set_option("verlet", 1 if verlet == "on" else 0)
    for n in (1, 60, 100):
        run_engine(a, net, n, fast=True)
        run_engine(b, net, n, fast=True)
        assert np.array_equal(a.export_pairs(), b.export_pairs())
    assert read_beads(a, ia).tobytes() == read_beads(b, ib).tobytes()
    assert a.stats()["max_list_length"] == b.stats()["max_list_length"]


# Candidate-list settings: kept with the default skin (30% of the reach), kept
# with a tiny skin (rebuilt at most rebuilds), rebuilt every time, and the
# automatic choice.
VERLET = {"on": {"verlet": 1}, "skin5": {"verlet": 1, "skin": 5}, "off": {"verlet": 0}, "auto": {}}


@pytest.mark.parametrize("verlet", sorted(VERLET))
@pytest.mark.parametrize("workload", sorted(WORKLOADS))
def test_cells_pairs_breaks_exact_over_1000_steps(need_gpu, workload, verlet):
    """f64 build against the sequential oracle, every rebuild of 1,000 steps."""
    net = WORKLOADS[workload]()
    eng, ids = engine_for(net, "f64")
    for k, v in VERLET[verlet].items():
        eng.set_option(k, v)
    orc = oracle_for(net)
    g = net.grid_update_freq
    n_rebuilds = 0
    for _ in range(1000 // g):
        run_engine(eng, net, g, fast=True)
        orc.run(g, g)
        assert np.array_equal(eng.export_cells(len(net.beads)), orc.cell_of_bead)
        assert np.array_equal(eng.export_pairs(), orc.pair_set())
        n_rebuilds += 1
    assert np.array_equal(eng.export_breaks(len(net.springs)), orc.break_step)
    print(workload, "rebuilds", n_rebuilds, "breaks", int((orc.break_step >= 0).sum()),
          "pairs", len(orc.pair_set()))


@pytest.mark.parametrize("workload", sorted(WORKLOADS))
def test_f32_single_step_forces(need_gpu, workload):
    """From a shared state (the oracle after 50 steps), one f32 step gives the
    oracle's forces to fp32 rounding."""
    net = WORKLOADS[workload]()
    orc = oracle_for(net)
    orc.run(50, net.grid_update_freq)
    net.beads = orc.beads.copy()
    net.springs = orc.springs.copy()
    net.angles = orc.angles.copy()
    eng, ids = engine_for(net, "f32")
    run_engine(eng, net, 1, fast=True)
    orc.run(1, net.grid_update_freq)
    got = read_beads(eng, ids)
    err = field_errors(got, orc.beads)
    print(workload, err)
    assert err["force"] <= 2e-5 and err["position"] <= 1e-12 + 1e-7


def test_stagewise_f32_rebuild_exact(need_gpu):
    """The f32 build hashes and lists in fp64 from fp64 positions, so fed the
    oracle's state it reproduces the oracle's cells and pairs exactly."""
    net = small_tensile()
    orc = oracle_for(net)
    orc.run(50, net.grid_update_freq)
    net.beads = orc.beads.copy()
    eng, ids = engine_for(net, "f32")
    orc.rebuild()
    eng.launch("ResetGrid"); eng.launch("ResetContactList"); eng.launch("FillGrid"); eng.launch("BuildContactList")
    eng.finish()
    assert np.array_equal(eng.export_cells(len(net.beads)), orc.cell_of_bead)
    assert np.array_equal(eng.export_pairs(), orc.pair_set())


def _cadence_network():
    """The oracle's cadence case (tests/test_oracle.py) as an engine network."""
    from fibrenet_amd import generate as G
    from test_oracle import _lj_coeffs, beads_of, pair, settings
    s = settings(box=(30.0, 30.0, 30.0), unit=10.0, cutoff=2.5, dt=0.01)
    b, q = beads_of([[10.0, 15.0, 15.0], [13.55, 15.0, 15.0]], s=s, mass=0.0)
    for f in ("velocity", "velocity_old"):
        b[f]["x"][1] = -10.0
    net = G.Network()
    net.beads, net.info, net.settings = b, q, s
    net.springs, net.angles = np.zeros(0, R.SPRING), np.zeros(0, R.ANGLE)
    net.pairs = np.array([pair("lj", **_lj_coeffs())], R.PAIR)
    net.cell_nh = G.grid_neighbours(0, 3, 3, 3)
    net.grid_update_freq = 8
    return net


@pytest.mark.parametrize("fast", [True, False])
@pytest.mark.parametrize("calls", [(13,), (17,), (11, 2), (5, 8), (10, 3, 3)])
def test_rebuild_cadence_matches_oracle(need_gpu, calls, fast):
    """Rebuild at the first step of every call and every grid_update_freq steps
    of it: the engine sees the approaching pair exactly when the oracle does."""
    from test_oracle import cadence_case
    net = _cadence_network()
    eng, ids = engine_for(net, "f64")
    for n in calls:
        run_engine(eng, net, n, fast=fast)
    f = R.vec(read_beads(eng, ids), "force")[0]
    ref, _ = cadence_case(calls)
    assert np.allclose(f, ref, rtol=1e-12, atol=0.0) and ((f == 0).all() == (ref == 0).all())


@pytest.mark.parametrize("fast", [True, False])
def test_simulation_minimize_then_run_matches_oracle(need_gpu, tmp_path, fast):
    """fn_sim from files (SimulationManager's loaders), MinimizeSystemEnergy
    with reset over 35 steps in chunks of 10 (the remainder dropped), then a
    run of 37 steps: the f64 engine follows the oracle's MinimizeSystemEnergy
    and RunSimulationLoop to the f64 tolerance."""
    from fibrenet_amd import Model, Simulation, generate as G
    net = small_tensile()
    net.input_extra = net.input_extra + ["minimize_steps 35", "minimize_substeps 10", "minimize_reset yes"]
    path = G.write_files(net, str(tmp_path), "tensile")
    sim = Simulation(path, "f64")
    sim.minimize(fast=fast)
    sim.run_loop(37, fast=fast)
    got = sim.engine.read_buffer(sim.buffer_id("bead"), R.BEAD)
    m = Model(path)
    orc = oracle_for(m)
    orc.minimize(35, 10, True, m.grid_update_freq)
    orc.run(37, m.grid_update_freq)
    errs = field_errors(got, orc.beads)
    for f, e in errs.items():
        assert e <= TOL["f64"][f], (f, e)


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("workload", sorted(WORKLOADS))
def test_candidate_list_is_exact(need_gpu, workload, precision):
    """Keeping the skin-padded candidates between rebuilds gives the same bytes
    as rebuilding them at every rebuild, and the candidates are kept for at
    least some rebuilds."""
    net = WORKLOADS[workload]()
    a, ia = engine_for(net, precision)
    b, ib = engine_for(net, precision)
    a.set_option("verlet", 1)
    b.set_option("verlet", 0)
    builds = 0
    for n in (7, 93, 100):
        run_engine(a, net, n, fast=True)
        run_engine(b, net, n, fast=True)
        builds += a.stats()["candidate_builds"]
        assert np.array_equal(a.export_pairs(), b.export_pairs())
    assert read_beads(a, ia).tobytes() == read_beads(b, ib).tobytes()
    rebuilds = a.stats()["rebuilds"]
    print(workload, precision, "candidate builds", builds, "of", rebuilds, "rebuilds")
    assert 1 <= builds < rebuilds
    assert b.stats()["candidate_builds"] >= 1


def _threshold_pairs(etas, seed=7):
    """Two-bead fibres far out in the 1000 x 1000 x 50 tensile box, in pairs
    whose first beads (and second beads) lie at squared distance
    cutoff_sq + eta: at x near 990 an fp32-rounded coordinate is off by up to
    3e-5, enough to put such pairs on the wrong side of the cutoff."""
    net = G.tensile_lj(n_fibres=2 * len(etas), n_beads=2, box=(1000.0, 1000.0, 50.0), seed=seed, d_min=0.6)
    s = np.array(net.settings, R.SETTINGS).reshape(())
    rs = np.random.Generator(np.random.PCG64(seed))
    pos = []
    for k, eta in enumerate(etas):
        a = np.array([990.0 + rs.uniform(0.0, 0.1), 20.0 + 8.0 * k + rs.uniform(0.0, 0.1), 25.0 + rs.uniform(0.0, 0.1)])
        u = rs.normal(size=3)
        u[2] = 0.0
        u /= np.linalg.norm(u)
        b = a + np.sqrt(float(s["cutoff_sq"]) + eta) * u
        up = np.array([0.0, 0.0, 1.0])
        pos += [a, a + up, b, b + up]
    pos = np.array(pos)
    R.set_vec(net.beads, "position", pos)
    R.set_vec(net.beads, "image_position", pos)
    R.set_vec(net.beads, "velocity", np.zeros_like(pos))
    R.set_vec(net.beads, "velocity_old", np.zeros_like(pos))
    net.info["type"] = 0
    return net, pos, float(s["cutoff_sq"])


def test_filter_decides_pairs_at_the_cutoff_in_fp64(need_gpu):
    """Kept candidates are filtered on fp32-rounded positions with an fp64
    recheck near the threshold: pairs within 1e-3 of the cutoff, where
    rounding alone misjudges some, are listed exactly as the one-pass fp64
    rebuild lists them."""
    etas = [s * e for e in (1e-3, 3e-4, 1e-4, 3e-5, 1e-5, 1e-6, 1e-8, 1e-10) for s in (-1.0, 1.0)] + [0.0]
    net, pos, cut2 = _threshold_pairs(etas)
    # rounding to fp32 flips at least one of these decisions
    p32 = pos.astype(np.float32).astype(np.float64)
    flips = 0
    for k in range(len(etas)):
        for o in (0, 1):
            i, j = 4 * k + o, 4 * k + 2 + o
            exact = np.sum((pos[i] - pos[j]) ** 2) < cut2
            rounded = np.sum((p32[i] - p32[j]) ** 2) < cut2
            flips += exact != rounded
    assert flips >= 1
    for precision in ("f64", "f32"):
        a, _ = engine_for(net, precision)
        b, _ = engine_for(net, precision)
        a.set_option("verlet", 1)
        b.set_option("verlet", 0)
        run_engine(a, net, 1, fast=True)
        run_engine(b, net, 1, fast=True)
        pa, pb = a.export_pairs(), b.export_pairs()
        assert len(pb) > 0 and np.array_equal(pa, pb), precision
        listed = {tuple(sorted(p)) for p in pb.tolist()}
        for k in range(len(etas)):
            for o in (0, 1):
                i, j = 4 * k + o, 4 * k + 2 + o
                assert ((i, j) in listed) == bool(np.sum((pos[i] - pos[j]) ** 2) < cut2), (precision, etas[k], o)


def _ensemble_members():
    """Networks of uneven size and kind: periodic penalty networks in
    different boxes, the tensile network (finite box, LJ, breaking springs)
    and the crosslinked one (bonds outside the chain layout)."""
    from fibrenet_amd import generate as G
    nets = [G.periodic_hs(n_fibres=f, n_beads=b, box=(L, L, 0.5 * L), seed=30 + f, spring_k=k)
            for f, b, L, k in ((12, 15, 16.0, 60.0), (40, 22, 30.0, 150.0), (5, 9, 9.0, 100.0))]
    return nets + [small_tensile(weak=True), crosslinked()]


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("flat", [1, 0])
def test_ensemble_equals_plain_runs(need_gpu, precision, flat):
    """a batched ensemble reproduces each network's own run bit for bit:
    same boxes, pair tables and integrator constants per network, contacts
    only inside a network, the same terms in the same order.  flat: the
    exact search flattened over the wave (k_search_ens_wave) or per bead."""
    from fibrenet_amd import Ensemble
    nets = _ensemble_members()
    guf = 8
    ens = Ensemble(nets, precision)
    ens.set_option("flat_search", flat)
    for n in (5, 60):
        ens.run_steps(n, guf, True, True)
    for k, net in enumerate(nets):
        eng, ids = engine_for(net, precision)
        for n in (5, 60):
            eng.run_steps(n, guf, len(net.springs) > 0, len(net.angles) > 0)
        assert ens.read(k, "bead").tobytes() == read_beads(eng, ids).tobytes(), k
        assert np.array_equal(ens.read(k, "spring")["status"], eng.read_buffer(ids["spring"], R.SPRING)["status"])
        assert np.array_equal(ens.read(k, "angle")["status"], eng.read_buffer(ids["angle"], R.ANGLE)["status"])
    st = ens.stats()
    assert st["steps"] == 65 and st["list_entries"] > 0


def _kept_members(group):
    """Ensembles whose networks can keep candidates under one skin: periodic
    penalty networks (listing reach 1.5, the smallest box half 3), or the
    finite Lennard-Jones tensile and crosslinked networks."""
    from fibrenet_amd import generate as G
    if group == "periodic":
        return [G.periodic_hs(n_fibres=f, n_beads=b, box=box, seed=40 + f, spring_k=k)
                for f, b, box, k in ((12, 15, (16.0, 16.0, 8.0), 60.0), (40, 22, (30.0, 30.0, 15.0), 150.0),
                                     (5, 9, (9.0, 9.0, 6.0), 100.0))]
    return [small_tensile(weak=True), crosslinked()]


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("group,skin", [("periodic", 300), ("periodic", 500), ("finite", 300), ("finite", 1000)])
def test_ensemble_kept_candidates_equal_plain_runs(need_gpu, precision, group, skin):
    """With kept candidates (verlet 1) the ensemble's networks share one
    skin and candidate stencil; each bead's displacement check, cell record
    and filter use its own network's box, grid and pair table.  The lists
    equal the one-pass rebuild's, so every network still reproduces its own
    plain run bit for bit.  At skin 500 the stencil spans seven cells, more
    than the 6-cell periodic axis of the smallest network, which the search
    then visits whole."""
    from fibrenet_amd import Ensemble
    nets = _kept_members(group)
    guf = 8
    ens = Ensemble(nets, precision)
    ens.set_option("verlet", 1)
    ens.set_option("skin", skin)
    for n in (5, 60, 40):
        ens.run_steps(n, guf, True, True)
        if n == 5:                 # the first run builds the candidates
            st = ens.stats()
            assert st["candidate_builds"] >= 1
            assert st["candidate_entries"] >= st["list_entries"] > 0
    for k, net in enumerate(nets):
        eng, ids = engine_for(net, precision)
        for n in (5, 60, 40):
            eng.run_steps(n, guf, len(net.springs) > 0, len(net.angles) > 0)
        assert ens.read(k, "bead").tobytes() == read_beads(eng, ids).tobytes(), k
        assert np.array_equal(ens.read(k, "spring")["status"], eng.read_buffer(ids["spring"], R.SPRING)["status"])


def test_ensemble_rejects_bad_networks(need_gpu):
    from fibrenet_amd import Ensemble, FibreNetError
    net = _ensemble_members()[0]
    bad = copy.deepcopy(net)
    bad.springs = bad.springs.copy()
    bad.springs["p2"][0] = len(bad.beads) + 3
    with pytest.raises(FibreNetError):
        Ensemble([net, bad], "f64")



@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("workload", ["tensile_lj_breaking", "periodic_hs"])
def test_graphs_equal_direct_launches(need_gpu, workload, precision):
    """Captured graphs of any chunk length give the direct launches' results."""
    net = WORKLOADS[workload]()
    spr, ang = len(net.springs) > 0, len(net.angles) > 0
    out = []
    for graphs, chunk in ((0, 128), (1, 8), (1, 64), (1, 128), (1, 256)):
        eng, ids = engine_for(net, precision)
        eng.set_option("graphs", graphs)
        eng.set_option("chunk", chunk)
        for n in (37, 300):
            eng.run_steps(n, net.grid_update_freq, spr, ang)
        out.append((read_beads(eng, ids).tobytes(), eng.export_breaks(len(net.springs)).tobytes(),
                    eng.export_pairs().tobytes()))
    assert all(o == out[0] for o in out[1:])


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("workload", ["tensile_lj_breaking", "crosslinked", "dense_lj"])
def test_block_sizes_equal(need_gpu, workload, precision):
    """The bead kernels' block size (128, 256, 512 threads; the chain kernel
    and the slot kernel) changes nothing: every wave covers the same beads."""
    net = WORKLOADS[workload]()
    spr, ang = len(net.springs) > 0, len(net.angles) > 0
    out = []
    for block in (256, 128, 512):
        eng, ids = engine_for(net, precision)
        eng.set_option("block", block)
        for n in (37, 120):
            eng.run_steps(n, net.grid_update_freq, spr, ang)
        out.append((read_beads(eng, ids).tobytes(), eng.export_breaks(len(net.springs)).tobytes(),
                    eng.export_pairs().tobytes()))
    assert all(o == out[0] for o in out[1:])


def test_block_option_is_checked(need_gpu):
    from fibrenet_amd import FibreNetError
    eng, _ = engine_for(small_periodic(), "f32")
    with pytest.raises(FibreNetError):
        eng.set_option("block", 192)
    assert eng.get_option("block") == 256


@pytest.mark.parametrize("workload", ["tensile_lj_breaking", "periodic_hs"])
def test_autotune_equals_run_steps(need_gpu, workload):
    """The autotuner's run is a plain run bit for bit, whatever it tries."""
    net = WORKLOADS[workload]()
    spr, ang = len(net.springs) > 0, len(net.angles) > 0
    a, ia = engine_for(net, "f32")
    trials, chosen = a.autotune(10000, net.grid_update_freq, spr, ang)   # room for every trial
    a.run_steps(100, net.grid_update_freq, spr, ang)        # continues with the chosen settings
    b, ib = engine_for(net, "f32")
    b.run_steps(10000, net.grid_update_freq, spr, ang)
    b.run_steps(100, net.grid_update_freq, spr, ang)
    assert read_beads(a, ia).tobytes() == read_beads(b, ib).tobytes()
    assert np.array_equal(a.export_breaks(len(net.springs)), b.export_breaks(len(net.springs)))
    assert np.array_equal(a.export_pairs(), b.export_pairs())
    knobs = {t["knob"] for t in trials}
    assert {"chunk", "staged", "verlet", "skin", "block"} <= knobs and all(t["ms_per_step"] > 0 for t in trials)
    assert a.stats()["steps"] == 10100


def test_dense_network_takes_the_long_paths(need_gpu):
    """The dense workload drives the paths the sparse ones leave idle: some
    64 consecutive beads hold more than 1,024 list entries (the compaction
    then writes per bead instead of through LDS), every 60 owned beads of a
    chain wave more than 256 (several contact passes), and the neighbour
    search tests more than 768 candidates per wave (several passes)."""
    net = dense_lj()
    eng, ids = engine_for(net, "f64")
    run_engine(eng, net, 1, fast=True)
    p = eng.export_pairs()
    n = len(net.beads)
    per_bead = np.bincount(p.ravel(), minlength=n)          # two-sided list length of every bead
    wave64 = np.add.reduceat(per_bead, np.arange(0, n, 64))
    chain60 = np.add.reduceat(per_bead, np.arange(0, n, 60))
    print("list entries per 64 beads: max", wave64.max(), "per chain wave: min", chain60[:-1].min())
    assert wave64.max() > 1024
    assert chain60[:-1].min() > 256
    assert per_bead.mean() > 768 / 64


def test_overflowing_run_is_repeated_with_longer_rows(need_gpu):
    """Kept candidates of the dense network (about 100 per bead) overflow the
    initial 64-entry search rows.  The engine keeps the state at the start of
    a run it has no row lengths for, grows the rows and repeats the run, which
    then ends bit for bit as a run made with long enough rows from the start."""
    net = dense_lj()
    a, ia = engine_for(net, "f64")
    b, ib = engine_for(net, "f64")
    for e in (a, b):
        e.set_option("verlet", 1)
    run_engine(b, net, 10, fast=True)              # b's rows grow here
    for kind, rec in (("bead", net.beads), ("spring", net.springs), ("angle", net.angles)):
        b.write_buffer(ib[kind], rec)              # b back at the start, rows kept
    run_engine(a, net, 30, fast=True)              # overflows, is repeated
    run_engine(b, net, 30, fast=True)
    assert read_beads(a, ia).tobytes() == read_beads(b, ib).tobytes()
    assert np.array_equal(a.export_pairs(), b.export_pairs())
    assert a.stats()["broken_springs"] == b.stats()["broken_springs"] > 0


def test_more_than_seven_bead_types_are_refused_at_upload(need_gpu):
    """The contact-list entries carry the pair index in 5 bits: a pair table
    of more than 32 entries (8 or more bead types) is refused when it is
    created, with the limit in the message."""
    from fibrenet_amd import Engine, FibreNetError
    eng = Engine("f64")
    with pytest.raises(FibreNetError) as e:
        eng.create_buffer("pair", np.zeros(36, R.PAIR))
    assert e.value.code == -5 and "at most 7 bead types" in str(e.value)
    assert eng.create_buffer("pair", np.zeros(28, R.PAIR)) >= 0
