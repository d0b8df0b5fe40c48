#!/usr/bin/env python3
"""Benchmark of the FibreNet step on one MI355X.

Workload (default): the tensile test with 2,000,000 beads - 20,000 straight
fibres of 100 beads in a finite 1000x1000x50 box, HR_FL springs (k 100,
E 1000, failure stress uniform in [50,150]), HR angles (k 10), Lennard-Jones
contacts (eps 1, sigma 1, cutoff 2.5) between all three bead types, beads
with x<5 clamped and x>995 pulled at +0.01, dt 5e-4, grid rebuild every 10
steps.  A step is the spring-term, angle-term and bead kernels over the whole
network (plus the grid rebuild every 10th step); inputs are resident in HBM
when the timed region starts.

The settings that cannot change a result (chain bead kernel, staged search,
kept contact candidates, graph chunk length) are chosen by the autotuner,
run on a second engine with the same network from the same warmed-up state.

The network is generated with its fibres at least 2^(1/6) sigma apart
and relaxed by MinimizeSystemEnergy with reset (8,000 steps, reset every
400) before anything is timed; the clamp and pull keep their velocities.
`--workload tensile_run` runs the same network as config 1 asks: 20,000
steps in 100-step RunSimulationLoop calls with the clamp's reaction force
read after each, reporting the dynamics along the way.

Prints one JSON line: bead-steps/s of the timed steps, the roofline of the
dominant kernel (the step kernel with the largest duration, measured with
HIP event nodes in the captured graphs of a second run of the same length,
against the survey's algorithmic bytes of a step) and a CPU baseline (the fp64 oracle
on all host threads with the reference's atomic adds, and on one core, on a
bounded sample of the same network).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def implementation_bytes(net, pairs, tb, aliased, chain):
    """HBM bytes each launch of this implementation moves at least (every
    datum of its own layout read or written once; tb = bytes of a
    force/velocity component, positions fp64).  Reported beside the survey's
    algorithmic bytes (survey_bytes), which the roofline uses.

    chain layout (one launch per step, the bead kernel evaluating the bonds):
      per bead position read and next position written (48 B), image read
      and written unless it is the position array (48 B), force of the
      previous step, velocity_old read and written, this step's force
      written, three mass coefficients (12 tb), radius (tb), list offset and
      count (8 B), chain spring and angle index (8 B); per spring type,
      status, rest length (16 B), stiffness (tb) and the failure constants of
      an HR_FL spring (16 B); per angle type, status (8 B), rest angle and
      stiffness (2 tb); per listed pair two list entries (8 B).

    otherwise (three launches per step):

    spring-term kernel: per spring p1, p2, type, status (16 B), rest length
      (8 B), stiffness (tb), the two failure constants of an HR_FL spring
      (16 B) and the term written (3 tb); per bead its image (24 B).
    angle-term kernel: per angle b1, b2, b3, s1, s2, type, status (28 B),
      rest angle and stiffness (2 tb), the two terms written (6 tb); per
      spring its status (4 B); per bead its image (24 B).
    bead kernel: per bead position read and next position written (48 B),
      image read and written unless it is the position array (48 B), force
      of the previous step, velocity_old read and written, this step's force
      written, three mass coefficients (12 tb), three list offsets (12 B);
      per spring two incidence entries and its term (8 B + 3 tb), per angle
      three entries and its terms (12 B + 6 tb), per listed pair two list
      entries (8 B).
    rebuild: per bead its position read (24 B) and its position, tag and
      loc written in cell order (32 B), per listed pair two list entries
      (8 B)."""
    import numpy as np
    n = len(net.beads)
    s_hr = int(np.sum(net.springs["type"] == 0))
    s_fl = int(np.sum(net.springs["type"] == 1))
    s = s_hr + s_fl
    a = len(net.angles)
    rebuild = 56 * n + 8 * pairs
    if chain:
        bead = ((48 + (0 if aliased else 48) + 13 * tb + 16) * n + (16 + tb) * s + 16 * s_fl
                + (8 + 2 * tb) * a + 8 * pairs)
        return (dict(spring=0, angle=0, bead=bead), rebuild,
                dict(beads=n, springs_hr=s_hr, springs_hr_fl=s_fl, angles=a, pairs=pairs))
    spring = (16 + 8 + tb + 3 * tb) * s + 16 * s_fl + 24 * n if s else 0
    angle = (28 + 2 * tb + 6 * tb) * a + 4 * s + 24 * n if a else 0
    bead = (48 + (0 if aliased else 48) + 12 * tb + 12) * n + (8 + 3 * tb) * s + (12 + 6 * tb) * a + 8 * pairs
    return (dict(spring=spring, angle=angle, bead=bead), rebuild,
            dict(beads=n, springs_hr=s_hr, springs_hr_fl=s_fl, angles=a, pairs=pairs))


def survey_bytes(counts, grid_update_freq, chain):
    """Algorithmic bytes of a step by the survey's formula (fp32 SoA, each
    array touched once per step; extra-precision state such as the fp64
    positions is not counted):

      B_step = 100 N + 16 S_HR + 28 S_FL + 20 A + 8 P + (36 N + 8 P) / grid_update_freq

    with N beads, S_HR / S_FL springs by type, A angles and P listed unique
    contact pairs.  Per kernel: with the chain layout the bead kernel is the
    whole step apart from the rebuild; otherwise the spring-term kernel
    takes 16 S_HR + 28 S_FL, the angle-term kernel 20 A and the bead kernel
    100 N + 8 P.  The rebuild is 36 N + 8 P per rebuild."""
    n, s_hr, s_fl, a, p = (counts[k] for k in ("beads", "springs_hr", "springs_hr_fl", "angles", "pairs"))
    bonds = 16 * s_hr + 28 * s_fl
    if chain:
        kern = dict(bead=100 * n + bonds + 20 * a + 8 * p)
    else:
        kern = dict(spring=bonds, angle=20 * a, bead=100 * n + 8 * p)
    rebuild = 36 * n + 8 * p
    return kern, rebuild, sum(kern.values()) + rebuild / grid_update_freq


def ensemble_bench(n_networks, steps, warmup, loop_steps, precision, guf=8, tune_steps=0):
    """Batched ensemble of config-2 networks against looping the plain engine
    over the same networks: bead-steps/s of each (Sum N_i per step).  With
    tune_steps, the autotuner first runs that many steps on the ensemble
    (the settings it chooses, block size and chunk length among them, also
    go to the looped engines)."""
    from fibrenet_amd import Engine, Ensemble, bind_standard, generate as G
    t0 = time.perf_counter()
    nets = G.ensemble(n_networks)
    gen_s = time.perf_counter() - t0
    n_tot = sum(len(n.beads) for n in nets)
    ens = Ensemble(nets, precision)
    tune = None
    if tune_steps > 0:
        trials, chosen = ens.autotune(tune_steps, guf, True, True)
        tune = {"steps": tune_steps, "chosen": chosen, "trials": trials}
    ens.run_steps(warmup, guf, True, True)
    ens.run_steps(steps, guf, True, True)          # captures the graphs of this run length
    ens.finish()
    t = time.perf_counter()
    ens.run_steps(steps, guf, True, True)
    wall = time.perf_counter() - t
    st = ens.stats()
    ens.close()
    # the plain engine, one network after the other (graphs captured first)
    engines = []
    for net in nets:
        net.cell_nh = G.grid_neighbours(0, int(net.settings["cells_x"]), int(net.settings["cells_y"]),
                                        int(net.settings["cells_z"]))
        e = Engine(precision)
        bind_standard(e, net)
        for k in ("block", "chunk"):
            if tune:
                e.set_option(k, tune["chosen"][k])
        e.run_steps(loop_steps, guf, True, True)
        engines.append(e)
    for e in engines:
        e.finish()
    t = time.perf_counter()
    for e in engines:
        e.run_steps(loop_steps, guf, True, True)
    for e in engines:
        e.finish()
    wall_loop = time.perf_counter() - t
    for e in engines:
        e.close()
    v_ens = n_tot * steps / wall
    v_loop = n_tot * loop_steps / wall_loop
    return {"networks": n_networks, "beads": n_tot, "steps": steps, "ms_per_step": wall * 1e3 / steps,
            "bead_steps_per_s": v_ens, "loop_steps": loop_steps, "loop_bead_steps_per_s": v_loop,
            "speedup_over_loop": v_ens / v_loop, "listed_pairs": st["list_entries"] // 2,
            "grid_update_freq": guf, "generate_s": gen_s, "autotune": tune}


def drift_vs_f64(net, options, marks=(200, 800)):
    """Drift of the fp32 build against the fp64 build of the same integrator:
    both engines run the network from the same start (the fp32 one with the
    measured engine's settings) and their beads are compared at each step
    count in `marks`: the largest and the rms position difference, and the
    largest velocity_old and force differences relative to the largest
    velocity_old and force of the fp64 run.  The marks lie before the
    workload's launched beads dominate (about step 900)."""
    import numpy as np
    from fibrenet_amd import Engine, bind_standard, records as R
    engs = []
    try:
        for prec in ("f32", "f64"):
            e = Engine(prec)
            engs.append((e, bind_standard(e, net)))
        for k, v in options.items():
            engs[0][0].set_option(k, v)
        out, done = [], 0
        for m in marks:
            for e, _ in engs:
                e.run_steps(m - done, net.grid_update_freq, True, True)
            done = m
            b32, b64 = (e.read_buffer(ids["bead"], R.BEAD) for e, ids in engs)
            dp = np.linalg.norm(R.vec(b32, "position") - R.vec(b64, "position"), axis=1)
            dv = np.abs(R.vec(b32, "velocity_old") - R.vec(b64, "velocity_old")).max()
            df = np.abs(R.vec(b32, "force") - R.vec(b64, "force")).max()
            out.append({"step": m, "max_position_diff": float(dp.max()),
                        "rms_position_diff": float(np.sqrt(np.mean(dp ** 2))),
                        "max_velocity_diff_rel": float(dv / np.abs(R.vec(b64, "velocity_old")).max()),
                        "max_force_diff_rel": float(df / np.abs(R.vec(b64, "force")).max()),
                        "breaks_f32": int((engs[0][0].export_breaks(len(net.springs)) >= 0).sum()),
                        "breaks_f64": int((engs[1][0].export_breaks(len(net.springs)) >= 0).sum())})
        return out
    finally:
        for e, _ in engs:
            e.close()


def bead_state(eng, ids):
    """Regime of the network after the timed steps: non-finite beads and the
    largest speed among the finite ones."""
    import numpy as np
    from fibrenet_amd import records as R
    b = eng.read_buffer(ids["bead"], R.BEAD)
    v = R.vec(b, "velocity_old")
    ok = np.isfinite(R.vec(b, "position")).all(1) & np.isfinite(v).all(1)
    return {"nonfinite_beads": int((~ok).sum()), "max_speed": float(np.linalg.norm(v[ok], axis=1).max())}


def adjoint_bench(net, precision, options, warmup, steps, every):
    """Adjoint cost on the benchmark network, on an engine of its own started
    from the warmed-up state (`warmup` steps), where the dynamics are still
    regular: the checkpointing forward over `steps` steps sampling the
    clamp's reaction force 20 times, then the backward sweep (segment
    recompute, per-step records, adjoint kernel) with dL/d(sample) = 1,
    returning the force-scale derivatives (those of the spring, angle and
    contact stiffnesses; the per-bead ones stay on the device).  The
    plain forward is timed on the same engine over the next `steps` steps
    (graph-captured, after a capture run), and the ratios are against it.

    The first backward call allocates the records of a segment (positions,
    bond flags and contact lists of `every` steps, about 3 GB here); the
    repeat figures time a second backward over the same forward run, as in
    each iteration after the first of a fitting loop."""
    import numpy as np
    from fibrenet_amd import Engine, bind_standard
    from fibrenet_amd.api import SCALE_OUTPUTS
    group = (np.asarray(net.info["type"]) == 1).astype(np.int32)
    samples = np.unique(np.linspace(steps // 20 - 1, steps - 1, 20).astype(np.int32))
    guf = net.grid_update_freq
    eng = Engine(precision)
    try:
        bind_standard(eng, net)
        for k, v in options.items():
            eng.set_option(k, v)
        if warmup > 0:
            eng.run_steps(warmup, guf, True, True)
        eng.finish()
        t0 = time.perf_counter()
        eng.adjoint_forward(steps, group, samples, guf, True, True, every)
        t1 = time.perf_counter()
        # the force-scale derivatives: the per-bead ones (96 MB here) stay on the device
        g = eng.adjoint_backward(np.ones(len(samples)), len(net.springs), len(net.angles), len(net.pairs),
                                 SCALE_OUTPUTS)
        t2 = time.perf_counter()
        # the same backward again: its records and buffers are allocated now,
        # as in every call after the first of a fitting loop
        eng.adjoint_backward(np.ones(len(samples)), len(net.springs), len(net.angles), len(net.pairs),
                             SCALE_OUTPUTS)
        t2b = time.perf_counter()
        eng.run_steps(steps, guf, True, True)          # graph capture
        eng.finish()
        t3 = time.perf_counter()
        eng.run_steps(steps, guf, True, True)
        t4 = time.perf_counter()
    finally:
        eng.close()
    fwd, bwd, plain = (t1 - t0) * 1e3 / steps, (t2 - t1) * 1e3 / steps, (t4 - t3) * 1e3 / steps
    bwd2 = (t2b - t2) * 1e3 / steps
    finite = all(bool(np.isfinite(g[k]).all()) for k in SCALE_OUTPUTS)
    return {"steps": steps, "start_step": warmup, "checkpoint_every": every or "auto", "samples": len(samples),
            "plain_forward_ms_per_step": plain, "forward_ms_per_step": fwd, "backward_ms_per_step": bwd,
            "backward_over_forward": bwd / plain, "total_over_forward": (fwd + bwd) / plain,
            "repeat_backward_ms_per_step": bwd2, "repeat_backward_over_forward": bwd2 / plain,
            "gradients_finite": finite,
            "d_spring_scale_sum": float(g["spring"].sum()), "d_angle_scale_sum": float(g["angle"].sum()),
            "d_contact_scale": [float(x) for x in g["pair"]]}


def tensile_fit_bench(args):
    """The differentiated tensile fit: 5,000 beads (100 bent fibres of 50
    beads, finite box 100x100x20), HR springs, HR angles, penalty
    contacts, clamp x < 5 and pull x > 95 at +0.01, 2,000 steps at dt 1e-3,
    the clamp's reaction force sampled at 20 evenly spaced strains; loss =
    squared error against the forward model's samples at hidden stiffnesses
    (spring 120, angle 8, contact 900 against 100, 10, 1000).  Measures the
    adjoint's cost against the plain forward, checks its gradient against
    central differences (h = 1e-3) of the fp64 engine and, as the CPU leg,
    against the fp64 dual-number reference on the host, then fits the three
    stiffnesses by gradient steps."""
    import numpy as np
    import torch
    from fibrenet_amd import Engine, bind_standard, generate as G
    from fibrenet_amd.autograd import TensileRun, tensile_samples

    net = G.tensile_fit()
    n, guf, every = net.run_steps, net.grid_update_freq, 100
    group = (np.asarray(net.info["type"]) == 1).astype(np.int32)
    steps = (np.arange(1, 21) * (n // 20) - 1).astype(np.int32)
    hidden = [1.2, 0.8, 0.9]
    base = [1.0, 1.0, 1.0]

    # plain forward: the captured-graph path over the same steps
    eng = Engine(args.precision)
    ids = bind_standard(eng, net)
    eng.run_steps(n, guf, True, True)
    eng.write_buffer(ids["bead"], net.beads)
    eng.write_buffer(ids["spring"], net.springs)
    eng.write_buffer(ids["angle"], net.angles)
    eng.finish()
    t = time.perf_counter()
    eng.run_steps(n, guf, True, True)
    eng.finish()
    plain_ms = (time.perf_counter() - t) * 1e3 / n
    eng.close()

    run = TensileRun(net, args.precision, group, steps, n, guf, every)
    target = run.forward(hidden)
    run.forward(base)
    run.backward(np.ones(len(steps)))                      # warm
    t = time.perf_counter()
    s0 = run.forward(base)
    t1 = time.perf_counter()
    g = run.backward(2.0 * (s0 - target))
    t2 = time.perf_counter()
    fwd_ms, bwd_ms = (t1 - t) * 1e3 / n, (t2 - t1) * 1e3 / n
    grad = np.array([g["spring"].sum(), g["angle"].sum(), g["pair"].sum()])
    loss0 = float(np.sum((s0 - target) ** 2))

    # central differences of the fp64 engine's loss, and its own adjoint
    run64 = TensileRun(net, "f64", group, steps, n, guf, every)
    target64 = run64.forward(hidden)
    s64 = run64.forward(base)
    g64 = run64.backward(2.0 * (s64 - target64))
    grad64 = np.array([g64["spring"].sum(), g64["angle"].sum(), g64["pair"].sum()])
    h = 1e-3
    fd = []
    for k in range(3):
        sp, sm = list(base), list(base)
        sp[k], sm[k] = 1 + h, 1 - h
        lp = np.sum((run64.forward(sp) - target64) ** 2)
        lm = np.sum((run64.forward(sm) - target64) ** 2)
        fd.append((lp - lm) / (2 * h))
    fd = np.array(fd)

    cpu = None
    if not args.no_cpu_baseline:
        # the CPU leg: the fp64 dual-number reference on the host, one run
        # with the three scale derivatives carried along
        from oracle.dual_reference import DualRun
        N = len(net.beads)
        t = time.perf_counter()
        d = DualRun(net.beads, net.info, net.springs, net.angles, net.pairs, net.settings, np.zeros((N, 3)),
                    np.zeros((N, 3)))
        vals, ders = d.run(n, guf, group, steps)
        cpu_s = time.perf_counter() - t
        gref = ders[:3] @ (2.0 * (vals - target64))
        rel = lambda a, b: (np.abs(a - b) / np.maximum(np.abs(b), 1e-300)).tolist()
        cpu = {"value": N * n / cpu_s, "unit": "bead-steps/s (fp64 forward-mode, 5 directions)", "cores": 1,
               "kind": "port", "sample": "the whole workload: %d beads, %d steps, %.1f s" % (N, n, cpu_s),
               "gradient": gref.tolist(), "rel_err_f64_adjoint": rel(grad64, gref),
               "rel_err_%s_adjoint" % args.precision: rel(grad, gref),
               "samples_rel_err": float(np.max(np.abs(vals - s64)) / np.max(np.abs(vals)))}

    # fit the three stiffnesses by gradient steps on their logarithms
    x = [torch.tensor(0.0, dtype=torch.float64, requires_grad=True) for _ in range(3)]
    opt = torch.optim.Adam(x, lr=0.05)
    tgt = torch.as_tensor(target)
    fit = []
    for it in range(args.fit_iterations):
        opt.zero_grad()
        scales = [torch.exp(v) for v in x]
        loss = ((tensile_samples(run, *scales) - tgt) ** 2).sum()
        loss.backward()
        opt.step()
        fit.append({"iteration": it, "loss": float(loss.detach()),
                    "stiffness": [100.0 * float(torch.exp(x[0].detach())), 10.0 * float(torch.exp(x[1].detach())),
                                  1000.0 * float(torch.exp(x[2].detach()))]})
    out = {"metric": "adjoint/forward cost ratio", "value": bwd_ms / plain_ms, "unit": "x", "n_gpus": 1,
           "steps": n, "warmup": n, "ms_per_step": bwd_ms, "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "%s state, f64 gradients" % args.precision,
           "data": "synthetic (generated bent-fibre network, seed %d)" % net.meta["seed"],
           "config": {"workload": "tensile_fit: 100 bent fibres x 50 beads (%d beads), finite box 100x100x20, "
                                  "HR springs k 100, HR angles k 10, penalty contacts k 1000, clamp x<5, "
                                  "pull x>95 at +0.01, dt 1e-3, 2000 steps, checkpoint every 100, 20 samples"
                                  % len(net.beads), "beads": len(net.beads)},
           "cpu_baseline": cpu,
           "extra": {"plain_forward_ms_per_step": plain_ms, "checkpointed_forward_ms_per_step": fwd_ms,
                     "backward_ms_per_step": bwd_ms, "total_over_forward": (fwd_ms + bwd_ms) / plain_ms,
                     "loss": loss0, "gradient": grad.tolist(), "gradient_f64": grad64.tolist(),
                     "central_differences_f64": fd.tolist(),
                     "fd_rel_err_f64": (np.abs(fd - grad64) / np.abs(grad64)).tolist(),
                     "hidden_stiffness": [120.0, 8.0, 900.0], "fit": fit}}
    print(json.dumps(out))


def tensile_run_bench(args):
    """Config 1 as its run asks: the relaxed 2,000,000-bead tensile network,
    20,000 steps as 200 RunSimulationLoop calls of 100 steps (the rebuild
    cadence restarts with each call, SimulationManager.cpp:151-181), the
    clamp's reaction force (the x-force summed over the clamped beads,
    fn_group_force) read after each call.  The timed region holds the calls
    and the reads; every 2,000 steps the beads are read back outside it
    for the dynamics (non-finite beads, largest speed, broken springs)."""
    import numpy as np
    from fibrenet_amd import Engine, bind_standard, generate as G
    t0 = time.perf_counter()
    net = G.tensile_lj(n_fibres=args.fibres or 20000)
    if args.relax_steps > 0:
        net = G.relax(net, args.relax_steps, args.relax_substeps)
    prep_s = time.perf_counter() - t0
    guf, chunk, total = net.grid_update_freq, 100, args.run_steps
    group = (np.asarray(net.info["type"]) == 1).astype(np.int32)
    chosen = None
    if args.tune_steps > 0:
        tuner = Engine(args.precision)
        bind_standard(tuner, net)
        _, chosen = tuner.autotune(args.tune_steps, guf, True, True)
        tuner.close()
    eng = Engine(args.precision)
    ids = bind_standard(eng, net)
    for k, v in (chosen or {}).items():
        eng.set_option(k, v)
    eng.run_steps(chunk, guf, True, True)            # captures the graphs of a 100-step call
    eng.write_buffer(ids["bead"], net.beads)
    eng.write_buffer(ids["spring"], net.springs)
    eng.finish()
    reaction, dyn, timed = [], [], 0.0
    for k in range(total // chunk):
        t = time.perf_counter()
        eng.run_steps(chunk, guf, True, True)
        reaction.append(eng.group_force(group))
        timed += time.perf_counter() - t
        done = (k + 1) * chunk
        if done % 2000 == 0:
            st = bead_state(eng, ids)
            st.update(step=done, broken_springs=eng.stats()["broken_springs"])
            dyn.append(st)
            print("tensile_run: step %d, %s" % (done, st), file=sys.stderr, flush=True)
    eng.close()
    n = len(net.beads)
    out = {"metric": "bead-steps/s", "value": n * total / timed, "unit": "bead-steps/s", "n_gpus": 1,
           "steps": total, "warmup": 0, "ms_per_step": timed * 1e3 / total, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 forces/velocities, f64 positions" if args.precision == "f32" else "f64",
           "data": "synthetic (generated fibre network, seed %d)" % net.meta["seed"],
           "config": {"workload": "tensile_run: config 1's run, %d steps in %d-step calls with the clamp's "
                                  "reaction force read after each, %d beads, relaxed by %d minimisation steps"
                                  % (total, chunk, n, args.relax_steps), "beads": n},
           "extra": {"options": chosen, "prepare_s": prep_s, "reaction_force_every_100_steps": reaction,
                     "dynamics": dyn, "all_finite": all(d["nonfinite_beads"] == 0 for d in dyn)}}
    print(json.dumps(out))


def batched_fit_bench(args):
    """The batched adjoint: 128 tensile networks of uneven size (fibre counts
    uniform in [50, 300], lengths in [30, 60] beads, HR_FL springs with
    failure stress uniform in [50, 150]), 1,000 steps each, the clamp's
    reaction force sampled at 20 strains per network; per-network loss =
    squared error against the forward model's samples at a hidden spring
    stiffness per network (log-uniform in [50, 200]).  Measures the batched
    adjoint's cost against the batched forward, then fits every network's
    spring stiffness by gradient steps (Adam on its logarithm), and the same
    fit by the fp64 engine as the reference for the fitted values."""
    import numpy as np
    from fibrenet_amd import generate as G
    from fibrenet_amd.autograd import BatchedTensileRun

    nets = G.tensile_batch(args.fit_networks)
    nn = len(nets)
    n, guf, every = nets[0].run_steps, nets[0].grid_update_freq, 100
    steps = (np.arange(1, 21) * (n // 20) - 1).astype(np.int32)
    k_hidden = np.array([net.meta["hidden_spring_k"] for net in nets])
    base = np.ones((nn, 3))
    hidden = base.copy()
    hidden[:, 0] = k_hidden / 100.0
    beads = sum(len(net.beads) for net in nets)

    run = BatchedTensileRun(nets, args.precision, steps, n, guf, every)
    target = run.forward(hidden)
    # the plain batched forward (captured graphs) from the same records
    run.ens.reload(run._scaled(base))
    run.ens.run_steps(n, guf, True, True)
    run.ens.reload(run._scaled(base))
    run.ens.finish()
    t = time.perf_counter()
    run.ens.run_steps(n, guf, True, True)
    run.ens.finish()
    plain_ms = (time.perf_counter() - t) * 1e3 / n
    run.ens.reload(run._scaled(base))
    run.ens.finish()
    t = time.perf_counter()
    s0 = run.ens.adjoint_forward(n, run.group, steps, guf, True, True, every)
    t1 = time.perf_counter()
    g0 = run.backward(2.0 * (s0 - target))
    t2 = time.perf_counter()
    fwd_ms, bwd_ms = (t1 - t) * 1e3 / n, (t2 - t1) * 1e3 / n

    def fit_run(r, tgt, label):
        """Adam on x = log(spring scale), one parameter per network."""
        x = np.zeros(nn)
        m = np.zeros(nn)
        v = np.zeros(nn)
        lr, b1, b2 = 0.05, 0.9, 0.999
        hist = []
        for it in range(args.fit_iterations):
            sc = base.copy()
            sc[:, 0] = np.exp(x)
            res = r.forward(sc) - tgt
            loss = np.sum(res * res, axis=1)
            g = r.backward(2.0 * res)[:, 0]                # dL_k / d log(spring scale of k)
            m = b1 * m + (1 - b1) * g
            v = b2 * v + (1 - b2) * g * g
            x -= lr * (m / (1 - b1 ** (it + 1))) / (np.sqrt(v / (1 - b2 ** (it + 1))) + 1e-12)
            if it % 10 == 0 or it == args.fit_iterations - 1:
                e = np.abs(100.0 * np.exp(x) - k_hidden) / k_hidden
                hist.append({"iteration": it, "loss": float(loss.sum()), "median_rel_err": float(np.median(e)),
                             "max_rel_err": float(e.max())})
                print("batched fit (%s): iteration %d, loss %.4g, median stiffness error %.3g"
                      % (label, it, loss.sum(), np.median(e)), file=sys.stderr, flush=True)
        return 100.0 * np.exp(x), hist

    def err_stats(k):
        e = np.abs(k - k_hidden) / k_hidden
        return {"median": float(np.median(e)), "p90": float(np.quantile(e, 0.9)), "max": float(e.max()),
                "within_1pct": int(np.sum(e < 0.01)), "within_5pct": int(np.sum(e < 0.05))}

    t = time.perf_counter()
    k_fit, fit = fit_run(run, target, args.precision)
    fit_s = time.perf_counter() - t
    run.ens.close()
    # the fp64 reference: the same fit by the fp64 engine against its own targets
    run64 = BatchedTensileRun(nets, "f64", steps, n, guf, every)
    k_fit64, fit64 = fit_run(run64, run64.forward(hidden), "f64")
    run64.ens.close()
    d64 = np.abs(k_fit - k_fit64) / k_fit64
    out = {"metric": "adjoint/forward cost ratio (batched)", "value": bwd_ms / plain_ms, "unit": "x", "n_gpus": 1,
           "steps": n, "warmup": n, "ms_per_step": bwd_ms, "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "%s state, f64 gradients" % args.precision,
           "data": "synthetic (generated bent-fibre networks)",
           "config": {"workload": "batched_fit: %d tensile networks (%d beads), fibres U[50,300] x beads U[30,60], "
                                  "HR_FL springs, penalty contacts, clamp/pull 0.01, dt 1e-3, 1000 steps, "
                                  "checkpoint every 100, 20 samples per network, hidden spring k log-uniform "
                                  "[50,200]" % (nn, beads), "networks": nn, "beads": beads},
           "cpu_baseline": None,
           "extra": {"plain_forward_ms_per_step": plain_ms, "checkpointed_forward_ms_per_step": fwd_ms,
                     "backward_ms_per_step": bwd_ms, "total_over_forward": (fwd_ms + bwd_ms) / plain_ms,
                     "fit_iterations": args.fit_iterations, "fit_seconds": fit_s, "fit": fit, "fit_f64": fit64,
                     "stiffness_rel_err": err_stats(k_fit), "stiffness_rel_err_f64": err_stats(k_fit64),
                     "fit_vs_f64_fit_rel_diff": {"median": float(np.median(d64)), "max": float(d64.max()),
                                                 "within_1pct": int(np.sum(d64 < 0.01))}}}
    print(json.dumps(out))


def host_cpu():
    """Host facts for the CPU baseline: logical CPUs (nproc), the CPUs this
    process may run on, the model name and the OpenMP thread count."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count()
    threads = int(os.environ.get("OMP_NUM_THREADS") or allowed or 1)
    return {"nproc": os.cpu_count(), "allowed_cpus": allowed, "model": model, "omp_threads": threads}


def cpu_baseline(net, steps, threaded):
    """The oracle on the host: `threaded` runs every kernel's work-items on
    all OpenMP threads with the reference's atomic adds, as a CPU OpenCL
    device runs kernel.cl; otherwise one core in work-item order.  Ten
    warm-up steps, then `steps` timed steps (grid rebuilds included)."""
    from oracle.oracle import Oracle
    orc = Oracle(net.beads, net.info, net.springs, net.angles, net.pairs, net.settings, net.cell_nh,
                 threaded=threaded)
    orc.run(10, net.grid_update_freq)
    t = time.perf_counter()
    orc.run(steps, net.grid_update_freq)
    dt = time.perf_counter() - t
    return len(net.beads) * steps / dt, dt


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--precision", default="f32", choices=["f32", "f64"])
    ap.add_argument("--workload", default="tensile_lj",
                    choices=["tensile_lj", "periodic_hs", "ensemble", "tensile_fit", "batched_fit", "tensile_run"])
    ap.add_argument("--run-steps", type=int, default=20000, help="tensile_run: steps of the run")
    ap.add_argument("--fit-iterations", type=int, default=None,
                    help="tensile_fit / batched_fit: gradient steps of the fit (30 / 200)")
    ap.add_argument("--fit-networks", type=int, default=128, help="batched_fit: networks")
    ap.add_argument("--networks", type=int, default=512, help="ensemble: number of networks")
    ap.add_argument("--loop-steps", type=int, default=40, help="ensemble: steps of the looped plain engine")
    ap.add_argument("--ensemble-networks", type=int, default=512,
                    help="networks of the ensemble comparison reported beside the plain workload (0: none; "
                         "512: the batched-ensemble configuration)")
    ap.add_argument("--fibres", type=int, default=None)
    ap.add_argument("--cpu-steps", type=int, default=100,
                    help="timed steps of the all-core CPU baseline (the threaded oracle)")
    ap.add_argument("--cpu-steps-1core", type=int, default=20,
                    help="timed steps of the one-core CPU baseline (the sequential oracle)")
    ap.add_argument("--relax-steps", type=int, default=8000,
                    help="tensile_lj: MinimizeSystemEnergy steps with reset before the run (0: none)")
    ap.add_argument("--relax-substeps", type=int, default=400,
                    help="tensile_lj: steps between the minimisation's energy resets")
    ap.add_argument("--adjoint-steps", type=int, default=200,
                    help="steps of the adjoint cost measurement on the same network (0: none)")
    ap.add_argument("--checkpoint-every", type=int, default=50)
    ap.add_argument("--tune-steps", type=int, default=5000,
                    help="steps of the autotuner, run on a second engine from the warmed-up state (0: none)")
    ap.add_argument("--set", action="append", default=[], metavar="OPTION=VALUE",
                    help="engine option applied to the measured engine after the autotuner (repeatable)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-drift", dest="drift", action="store_false",
                    help="skip the fp32-against-fp64 drift check (tensile_lj / periodic_hs, fp32)")
    ap.add_argument("--copy-bytes", type=float, default=0,
                    help="also run one device copy of this many bytes (counter calibration)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the step is not sharded across GPUs in this build "
              "(see the sharding design); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from fibrenet_amd import Engine, bind_standard, generate as G

    if args.workload == "tensile_fit":
        args.fit_iterations = 30 if args.fit_iterations is None else args.fit_iterations
        tensile_fit_bench(args)
        return
    if args.workload == "batched_fit":
        args.fit_iterations = 200 if args.fit_iterations is None else args.fit_iterations
        batched_fit_bench(args)
        return
    if args.workload == "tensile_run":
        tensile_run_bench(args)
        return

    if args.workload == "ensemble":
        e = ensemble_bench(args.networks, args.steps, args.warmup, args.loop_steps, args.precision,
                           tune_steps=args.tune_steps)
        out = {"metric": "bead-steps/s (batched ensemble)", "value": e["bead_steps_per_s"], "unit": "bead-steps/s",
               "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": e["ms_per_step"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "f32 forces/velocities, f64 positions" if args.precision == "f32" else "f64",
               "data": "synthetic (generated periodic fibre networks)",
               "config": {"workload": "ensemble: %d periodic networks, fibres U[20,400] x beads U[20,80], cubic "
                                      "boxes of side 60-200 at volume fraction 0.05, spring k log-uniform "
                                      "[30,300], penalty contacts, dt 1e-3" % args.networks,
                          "beads": e["beads"], "precision": args.precision},
               "extra": e}
        print(json.dumps(out))
        return

    t0 = time.perf_counter()
    relax_s = None
    if args.workload == "tensile_lj":
        net = G.tensile_lj(n_fibres=args.fibres or 20000)
        if args.relax_steps > 0:
            t1 = time.perf_counter()
            net = G.relax(net, args.relax_steps, args.relax_substeps)
            relax_s = time.perf_counter() - t1
        desc = ("tensile_lj: %d fibres x 100 beads (%d beads), finite box 1000x1000x50, placed >= 2^(1/6) sigma "
                "apart, relaxed by %d minimisation steps (reset every %d), HR_FL springs + HR angles"
                " + LJ contacts (cutoff 2.5), clamp/pull 0.01, dt 5e-4, grid_update_freq 10"
                % (net.meta["n_fibres"], len(net.beads), args.relax_steps, args.relax_substeps))
    else:
        net = G.periodic_hs(n_fibres=args.fibres or 200)
        desc = ("periodic_hs: %d fibres x 50 beads (%d beads), periodic box 100x100x20, HR springs + HR angles"
                " + penalty contacts, dt 1e-3, grid_update_freq 8" % (net.meta["n_fibres"], len(net.beads)))
    gen_s = time.perf_counter() - t0 - (relax_s or 0.0)

    eng = Engine(args.precision)
    ids = bind_standard(eng, net)
    spr, ang = len(net.springs) > 0, len(net.angles) > 0
    if args.warmup > 0:
        eng.run_steps(args.warmup, net.grid_update_freq, spr, ang)
    eng.finish()
    tune = None
    if args.tune_steps > 0:
        # The autotuner on a second engine holding the same network, from the
        # same warmed-up state, so that the timed region below starts where it
        # would without tuning; its result-neutral choices are then applied.
        tuner = Engine(args.precision)
        bind_standard(tuner, net)
        if args.warmup > 0:
            tuner.run_steps(args.warmup, net.grid_update_freq, spr, ang)
        trials, chosen = tuner.autotune(args.tune_steps, net.grid_update_freq, spr, ang)
        tuner.close()
        for k, v in chosen.items():
            eng.set_option(k, v)
        tune = {"steps": args.tune_steps, "chosen": chosen, "trials": trials}
    for kv in args.set:
        k, v = kv.split("=")
        eng.set_option(k, int(v))
    print("bench: %s, warmup %d done, timing %d steps" % (desc, args.warmup, args.steps), file=sys.stderr,
          flush=True)

    # Timed region: K steps in the hipGraph-captured chunks, bracketed by
    # stream synchronisations (run_steps returns after the stream has
    # drained).  A first run of the same length captures the graphs outside
    # the timed region.
    eng.run_steps(args.steps, net.grid_update_freq, spr, ang)   # capture
    eng.finish()
    t = time.perf_counter()
    eng.run_steps(args.steps, net.grid_update_freq, spr, ang)
    wall = time.perf_counter() - t
    # Kernel durations: the same run length again with HIP event nodes around
    # every step kernel and rebuild in the graphs (read after each chunk).
    eng.set_timing(True)
    eng.run_steps(args.steps, net.grid_update_freq, spr, ang)   # capture with event nodes
    eng.set_timing(True)                                          # clears the sums, keeps the graphs
    eng.finish()
    t = time.perf_counter()
    eng.run_steps(args.steps, net.grid_update_freq, spr, ang)
    wall_events = time.perf_counter() - t
    tim = eng.timing()
    st = eng.stats()
    state = bead_state(eng, ids)
    eng.set_timing(False)

    n = len(net.beads)
    pairs = st["list_entries"] // 2
    aliased = bool(st["img_aliases_pos"])
    chain = bool(eng.get_option("chain") and eng.get_option("chain_layout"))
    impl_kernel, impl_rebuild, counts = implementation_bytes(net, pairs, 4 if args.precision == "f32" else 8,
                                                          aliased, chain)
    b_kernel, b_rebuild, b_step = survey_bytes(counts, net.grid_update_freq, chain)
    n_t = max(tim["steps"], 1)
    kern_ms = {"spring": tim["spring_ms"] / n_t, "angle": tim["angle_ms"] / n_t, "bead": tim["bead_ms"] / n_t}
    names = {"spring": "k_spring_terms", "angle": "k_angle_terms",
             "bead": "k_bead_step_chain" if chain else "k_bead_step"}
    if chain:   # the bond kernels are not launched; their event pairs bracket nothing
        kern_ms = {"bead": kern_ms["bead"]}
    dom = max(kern_ms, key=kern_ms.get)
    rebuild_ms = tim["rebuild_ms"] / max(tim["rebuilds"], 1)
    achieved = b_kernel[dom] / (kern_ms[dom] * 1e-3) / 1e9
    value = n * args.steps / wall
    step_gbs = b_step * args.steps / wall / 1e9

    copy = None
    if args.copy_bytes > 0:   # one k_copy_probe launch of known size, for counter calibration
        nb = (int(args.copy_bytes) + 15) // 16 * 16
        copy = {"bytes": nb, "ms": eng.copy_probe(nb)}

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_%s_%s_%s.json" % (names[dom], args.workload, args.precision))
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get("bytes_per_launch")

    adj = None
    if args.adjoint_steps > 0:
        try:
            opts = {k: eng.get_option(k) for k in ("chunk", "chain", "staged", "verlet", "block", "skin")}
            adj = adjoint_bench(net, args.precision, opts, args.warmup, args.adjoint_steps, args.checkpoint_every)
        except Exception as exc:          # the plain measurement stands on its own
            adj = {"error": repr(exc)}

    drift = None
    if args.drift and args.precision == "f32":
        try:
            drift = drift_vs_f64(net, {k: eng.get_option(k) for k in ("chunk", "chain", "staged", "verlet", "block", "skin")})
        except Exception as exc:          # the plain measurement stands on its own
            drift = {"error": repr(exc)}

    ens = None
    if args.ensemble_networks > 0:
        try:
            ens = ensemble_bench(args.ensemble_networks, 400, 40, args.loop_steps, args.precision,
                                 tune_steps=min(args.tune_steps, 1200))
        except Exception as exc:          # the plain measurement stands on its own
            ens = {"error": repr(exc)}

    cpu = None
    if not args.no_cpu_baseline:
        host = host_cpu()
        v, dt = cpu_baseline(net, args.cpu_steps, True)
        v1, dt1 = cpu_baseline(net, args.cpu_steps_1core, False)
        reb = lambda k: (k + net.grid_update_freq - 1) // net.grid_update_freq
        cpu = {"value": v, "unit": "bead-steps/s", "cores": host["omp_threads"], "kind": "port",
               "sample": "the same %d-bead network, %d steps (%d grid rebuilds) after 10 warm-up steps: the fp64 "
                         "oracle (restatement of kernel.cl) with every kernel's work-items on %d OpenMP threads in "
                         "64-item chunks, shared forces added by compare-and-swap as kernel.cl:174-193 does, one "
                         "join per kernel; %.1f s" % (n, args.cpu_steps, reb(args.cpu_steps), host["omp_threads"], dt),
               "host": host,
               "one_core": {"value": v1, "cores": 1,
                            "sample": "%d steps (%d grid rebuilds) of the sequential oracle, %.1f s"
                                      % (args.cpu_steps_1core, reb(args.cpu_steps_1core), dt1)}}

    out = {
        "metric": "bead-steps/s", "value": value, "unit": "bead-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 forces/velocities, f64 positions" if args.precision == "f32" else "f64",
        "data": "synthetic (generated fibre network, seed %d)" % net.meta["seed"],
        "config": {"workload": desc, "beads": n, "springs": len(net.springs), "angles": len(net.angles),
                   "listed_pairs": pairs, "grid_update_freq": net.grid_update_freq, "precision": args.precision,
                   "bond_path": "chain (in-wave)" if chain else "bond kernels + incidence slots"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": names[dom], "algorithmic_bytes_per_launch": b_kernel[dom],
                     "traffic_over_algorithmic": traffic / b_kernel[dom] if traffic else None,
                     "avg_launch_ms": kern_ms[dom], "launches_timed": tim["steps"],
                     "bytes_formula": "100 N + 16 S_HR + 28 S_FL + 20 A + 8 P per step, + (36 N + 8 P) per rebuild",
                     "step_achieved": step_gbs, "step_frac": step_gbs / HBM_PEAK_GBS,
                     "step_algorithmic_bytes": b_step},
        "cpu_baseline": cpu,
        "extra": {"kernels": {names[k]: {"avg_ms": kern_ms[k], "algorithmic_bytes": b_kernel[k],
                                          "implementation_bytes": impl_kernel[k],
                                          "GBps": b_kernel[k] / (kern_ms[k] * 1e-3) / 1e9 if kern_ms[k] > 0 else None}
                              for k in kern_ms},
                  "step_kernels_ms": tim["step_ms"] / n_t,
                  "event_timed_bead_steps_per_s": n * args.steps / wall_events,
                  "event_timed_ms_per_step": wall_events * 1e3 / args.steps,
                  "rebuild_ms": rebuild_ms, "rebuild_algorithmic_bytes": b_rebuild,
                  "rebuild_implementation_bytes": impl_rebuild,
                  "counts": counts, "stats": st, "state_after_timed": state,
                  "graphs": eng.get_option("graphs"),
                  "generate_s": gen_s, "relax_s": relax_s,
                  "gpu_over_cpu": (value / cpu["value"]) if cpu else None,
                  "gpu_over_cpu_1core": (value / cpu["one_core"]["value"]) if cpu else None, "copy_bytes": copy,
                  "ensemble": ens, "adjoint": adj, "drift_vs_f64": drift, "autotune": tune},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
