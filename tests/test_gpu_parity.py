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
import numpy as np
import pytest

from fibrenet_amd import records as R
from helpers import (crosslinked, engine_for, field_errors, oracle_for, read_beads, run_engine, small_periodic,
                     small_tensile)

pytestmark = pytest.mark.gpu

TOL = {"f64": {f: 1e-9 for f in ("position", "image_position", "velocity", "force", "velocity_old",
                                 "force_old")},
       "f32": {"position": 1e-5, "image_position": 1e-5, "velocity": 2e-3, "force": 1e-2,
               "velocity_old": 2e-3, "force_old": 1e-2}}

WORKLOADS = {"periodic_hs": lambda: small_periodic(), "tensile_lj": lambda: small_tensile(),
             "tensile_lj_breaking": lambda: small_tensile(weak=True), "crosslinked": lambda: crosslinked()}

# Networks whose bonds follow the chain layout (fast path: bonds evaluated in
# the bead kernel by wave shuffles) and one that does not (bond kernels and
# incidence slots).
CHAIN = {"periodic_hs": 1, "tensile_lj": 1, "tensile_lj_breaking": 1, "crosslinked": 0}


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


@pytest.mark.parametrize("workload", sorted(WORKLOADS))
def test_cells_pairs_breaks_exact_over_1000_steps(need_gpu, workload):
    """f64 build against the sequential oracle, every rebuild of 1,000 steps."""
    net = WORKLOADS[workload]()
    eng, ids = engine_for(net, "f64")
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
