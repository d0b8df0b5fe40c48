"""The adjoint on the GPU (fn_adjoint_forward / fn_adjoint_backward).

Its forward is run_steps bit for bit; its backward recomputes every segment
from its checkpoint, leaves the engine where the forward ended and repeats
bit for bit; its derivatives equal the fp64 dual-number reference
(oracle/dual_reference.py, forward-mode differentiation of the reference's
own formulas) and central differences of the engine's own fp64 forward.
"""
import copy

import numpy as np
import pytest
import torch

from fibrenet_amd import FibreNetError, records as R
from fibrenet_amd.autograd import TensileRun, tensile_samples
from helpers import ADJ_NETS, adjoint_net, dual_run, engine_for, read_beads

pytestmark = pytest.mark.gpu

N_STEPS = 300
GUF = 10
SAMPLES = np.arange(14, N_STEPS, 15)          # 20 samples


def _group(net):
    return (net.info["type"] == 1).astype(np.int32)


def _weights(n=len(SAMPLES)):
    return np.random.default_rng(5).normal(size=n)


def _dirs(net):
    rs = np.random.default_rng(7)
    N = len(net.beads)
    return rs.normal(size=(N, 3)), rs.normal(size=(N, 3))


def _adjoint(net, prec, w, every=50, n=N_STEPS):
    eng, ids = engine_for(net, prec)
    s = eng.adjoint_forward(n, _group(net), SAMPLES[SAMPLES < n], grid_update_freq=GUF, checkpoint_every=every)
    g = eng.adjoint_backward(w, len(net.springs), len(net.angles), len(net.pairs))
    return eng, ids, s, g


def _directional(net, g):
    up, uv = _dirs(net)
    return np.array([g["spring"].sum(), g["angle"].sum(), g["pair"].sum(), np.sum(g["position"] * up),
                     np.sum(g["velocity_old"] * uv)])


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_forward_is_run_steps(need_gpu, prec):
    net = adjoint_net(hr_fl=True)
    eng, ids = engine_for(net, prec)
    s = eng.adjoint_forward(N_STEPS, _group(net), SAMPLES, grid_update_freq=GUF, checkpoint_every=40)
    ref, rids = engine_for(net, prec)
    ref.run_steps(N_STEPS, GUF)
    assert read_beads(eng, ids).tobytes() == read_beads(ref, rids).tobytes()
    assert np.array_equal(eng.export_breaks(len(net.springs)), ref.export_breaks(len(net.springs)))
    assert np.sum(eng.export_breaks(len(net.springs)) >= 0) > 0
    f = R.vec(read_beads(eng, ids), "force")[_group(net) == 1, 0].sum()
    assert abs(s[-1] - f) <= 1e-12 * max(1.0, abs(f))


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_backward_restores_and_repeats(need_gpu, prec):
    net = adjoint_net(hr_fl=True)
    w = _weights()
    eng, ids = engine_for(net, prec)
    eng.adjoint_forward(N_STEPS, _group(net), SAMPLES, grid_update_freq=GUF, checkpoint_every=50)
    after_fwd = read_beads(eng, ids).tobytes()
    pairs_fwd = eng.export_pairs()
    g1 = eng.adjoint_backward(w, len(net.springs), len(net.angles), len(net.pairs))
    assert read_beads(eng, ids).tobytes() == after_fwd
    assert np.array_equal(eng.export_pairs(), pairs_fwd)
    g2 = eng.adjoint_backward(w, len(net.springs), len(net.angles), len(net.pairs))
    for k in g1:
        assert g1[k].tobytes() == g2[k].tobytes(), k
    # the engine continues as if only the forward had run
    eng.run_steps(20, GUF)
    ref, rids = engine_for(net, prec)
    ref.run_steps(N_STEPS, GUF)
    ref.run_steps(20, GUF)
    assert read_beads(eng, ids).tobytes() == read_beads(ref, rids).tobytes()


def test_checkpoint_spacing_does_not_change_gradients(need_gpu):
    net = adjoint_net(hr_fl=True)
    w = _weights()
    ga = _adjoint(net, "f64", w, every=10)[3]
    gb = _adjoint(net, "f64", w, every=300)[3]
    gc = _adjoint(net, "f64", w, every=0)[3]
    for k in ga:
        assert ga[k].tobytes() == gb[k].tobytes() == gc[k].tobytes(), k


def test_backward_needs_a_current_forward(need_gpu):
    net = adjoint_net()
    eng, ids = engine_for(net, "f64")
    with pytest.raises(RuntimeError):
        eng.adjoint_backward(_weights(), len(net.springs), len(net.angles), len(net.pairs))
    eng.adjoint_forward(50, _group(net), [49], grid_update_freq=GUF)
    eng.run_steps(5, GUF)
    with pytest.raises(RuntimeError):
        eng.adjoint_backward([1.0], len(net.springs), len(net.angles), len(net.pairs))
    with pytest.raises(RuntimeError):
        eng.adjoint_forward(50, _group(net), [20, 10], grid_update_freq=GUF)
    with pytest.raises(RuntimeError):
        eng.adjoint_forward(50, _group(net), [10], grid_update_freq=GUF, checkpoint_every=15)


# (precision, network, tolerance): fp64 differentiates the same trajectory
# as the reference up to rounding; fp32 differentiates its own trajectory,
# whose fp32 forces move it away from the fp64 one.
@pytest.mark.parametrize("prec,kind,tol", [("f64", "hs", 1e-7), ("f64", "hs_fl", 1e-7), ("f32", "hs", 1e-3),
                                           ("f64", "lj", 1e-7), ("f32", "lj", 1e-3), ("f64", "periodic", 1e-7),
                                           ("f32", "periodic", 1e-3), ("f64", "xlink", 1e-7),
                                           ("f32", "xlink", 1e-3), ("f64", "st", 1e-7), ("f32", "st", 1e-3),
                                           ("f64", "types4", 1e-7), ("f32", "types4", 1e-3)])
def test_matches_dual_number_reference(need_gpu, prec, kind, tol):
    net = ADJ_NETS[kind]()
    w = _weights()
    eng, _, s, g = _adjoint(net, prec, w)
    # every network but xlink is in the chain layout (one-launch adjoint step)
    assert eng.get_option("chain_layout") == (0 if kind == "xlink" else 1)
    up, uv = _dirs(net)
    vals, ders = dual_run(net, up, uv).run(N_STEPS, GUF, _group(net), SAMPLES)
    ref = ders @ w
    got = _directional(net, g)
    assert np.max(np.abs(s - vals)) <= (1e-9 if prec == "f64" else 1e-3) * np.max(np.abs(vals))
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-3 * np.max(np.abs(ref)))
    assert np.all(err <= tol), (got, ref, err)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_chain_adjoint_equals_slot_path(need_gpu, prec):
    """The one-launch chain-layout adjoint step (option chain = 1) and the
    bond kernels with incidence slots (chain = 0) give the same bead and bond
    adjoints bit for bit, on a network whose HR_FL springs break; the contact
    scale adjoints are summed over different block rows, so they agree to
    rounding."""
    net = adjoint_net(hr_fl=True)
    w = _weights()
    out = []
    for chain in (1, 0):
        eng, _ = engine_for(net, prec)
        eng.set_option("chain", chain)
        assert eng.get_option("chain_layout") == 1
        eng.adjoint_forward(N_STEPS, _group(net), SAMPLES, grid_update_freq=GUF, checkpoint_every=50)
        assert np.sum(eng.export_breaks(len(net.springs)) >= 0) > 0
        out.append(eng.adjoint_backward(w, len(net.springs), len(net.angles), len(net.pairs)))
    for k in ("position", "velocity_old", "spring", "angle"):
        assert np.all(np.isfinite(out[0][k])), k
        assert out[0][k].tobytes() == out[1][k].tobytes(), k
    np.testing.assert_allclose(out[0]["pair"], out[1]["pair"], rtol=1e-12,
                               atol=1e-12 * np.abs(out[1]["pair"]).max())


def test_matches_central_differences(need_gpu):
    """h = 1e-3 relative on each force scale, on the fp64 engine.  The
    forward rounds spring and contact lengths to float (native_sqrt, as the
    reference does), so it is smooth only down to ~1e-7 relative and the
    differences carry that noise divided by 2h; on this case the dual
    reference's own differences at h = 1e-3 stand within 1.1e-2 of its exact
    derivative, hence the 3e-2 bound."""
    net = adjoint_net()
    w = _weights()
    _, _, s0, g = _adjoint(net, "f64", w)
    got = np.array([g["spring"].sum(), g["angle"].sum(), g["pair"].sum()])
    run = TensileRun(net, "f64", _group(net), SAMPLES, N_STEPS, GUF, checkpoint_every=50)
    h = 1e-3
    fd = []
    for k in range(3):
        sp, sm = [1.0] * 3, [1.0] * 3
        sp[k], sm[k] = 1 + h, 1 - h
        fd.append((run.forward(sp) @ w - run.forward(sm) @ w) / (2 * h))
    fd = np.array(fd)
    assert np.all(np.abs(fd - got) <= 3e-2 * np.abs(got)), (fd, got)
    # and the engine's differences are the reference's differences: the two
    # forwards agree to rounding, so this holds far more tightly
    ref = []
    for k in range(3):
        sp, sm = [1.0] * 3, [1.0] * 3
        sp[k], sm[k] = 1 + h, 1 - h
        lp = dual_run(net, scales=sp).run(N_STEPS, GUF, _group(net), SAMPLES)[0] @ w
        lm = dual_run(net, scales=sm).run(N_STEPS, GUF, _group(net), SAMPLES)[0] @ w
        ref.append((lp - lm) / (2 * h))
    assert np.all(np.abs(fd - np.array(ref)) <= 1e-6 * np.abs(got)), (fd, ref)


def test_torch_function(need_gpu):
    """Squared error against a target made at other stiffnesses: torch's
    gradient through the harness is the adjoint's, and a step along it
    lowers the loss."""
    net = adjoint_net()
    run = TensileRun(net, "f64", _group(net), SAMPLES, N_STEPS, GUF, checkpoint_every=50)
    target = torch.as_tensor(run.forward([1.2, 0.8, 0.9]))
    x = [torch.tensor(1.0, dtype=torch.float64, requires_grad=True) for _ in range(3)]
    loss = ((tensile_samples(run, *x) - target) ** 2).sum()
    grads = torch.autograd.grad(loss, x)
    w = 2.0 * (run.forward([1.0, 1.0, 1.0]) - target.numpy())
    g = run.backward(w)
    want = [g["spring"].sum(), g["angle"].sum(), g["pair"].sum()]
    assert np.allclose([float(v) for v in grads], want, rtol=1e-12, atol=0.0)
    step = 1e-3 / max(abs(float(v)) for v in grads)
    x2 = [float(a.detach()) - step * float(b) for a, b in zip(x, grads)]
    loss2 = float(((torch.as_tensor(run.forward(x2)) - target) ** 2).sum())
    assert loss2 < float(loss.detach())


@pytest.mark.parametrize("prec,tol", [("f64", 1e-7), ("f32", 1e-3)])
def test_tensile_fit_network_gradients(need_gpu, prec, tol):
    """The differentiated tensile fit's 5,000-bead network (300 of its 2,000
    steps): the loss gradient with respect to the three stiffness scales
    against the fp64 dual-number reference."""
    from fibrenet_amd import generate as G
    net = G.tensile_fit()
    n = 300
    steps = np.arange(14, n, 15)
    group = _group(net)
    run = TensileRun(net, prec, group, steps, n, net.grid_update_freq, checkpoint_every=100)
    target = run.forward([1.2, 0.8, 0.9])
    s = run.forward([1.0, 1.0, 1.0])
    g = run.backward(2.0 * (s - target))
    got = np.array([g["spring"].sum(), g["angle"].sum(), g["pair"].sum()])
    vals, ders = dual_run(net).run(n, net.grid_update_freq, group, steps)
    ref = ders[:3] @ (2.0 * (vals - target))
    assert np.all(np.abs(got - ref) <= tol * np.abs(ref)), (got, ref)


def _batch_members():
    """Three small tensile networks of different sizes and loads."""
    return [adjoint_net(seed=31, n_fibres=30, n_beads=10, hr_fl=True),
            adjoint_net(seed=32, n_fibres=44, n_beads=12, box=(26.0, 22.0, 8.0)),
            adjoint_net(seed=33, n_fibres=25, n_beads=14, pull_v=0.8, hr_fl=True)]


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_batched_adjoint_equals_single_networks(need_gpu, prec):
    """The ensemble's adjoint gives every network what its own engine gives:
    samples and the per-bead derivatives bit for bit (the same terms in the
    same order bead by bead: incidence slots in the ensemble, the chain
    kernel's in-lane bonds for a single network), the per-bond and
    per-pair-entry scale derivatives to rounding (their sums are grouped
    differently)."""
    from fibrenet_amd import Ensemble
    nets = _batch_members()
    rs = np.random.default_rng(3)
    w = rs.normal(size=(len(nets), len(SAMPLES)))
    ens = Ensemble(nets, prec)
    group = np.concatenate([_group(n) for n in nets])
    s = ens.adjoint_forward(N_STEPS, group, SAMPLES, grid_update_freq=GUF, checkpoint_every=50)
    off = {k: ens.offsets(k) for k in ("bead", "spring", "angle", "pair")}
    g = ens.adjoint_backward(w, int(off["spring"][-1]), int(off["angle"][-1]), int(off["pair"][-1]))
    for k, net in enumerate(nets):
        eng, ids = engine_for(net, prec)
        sk = eng.adjoint_forward(N_STEPS, _group(net), SAMPLES, grid_update_freq=GUF, checkpoint_every=50)
        gk = eng.adjoint_backward(w[k], len(net.springs), len(net.angles), len(net.pairs))
        b0, b1 = off["bead"][k], off["bead"][k + 1]
        assert np.allclose(s[k], sk, rtol=1e-12, atol=1e-12 * np.max(np.abs(sk)))
        assert g["position"][b0:b1].tobytes() == gk["position"].tobytes()
        assert g["velocity_old"][b0:b1].tobytes() == gk["velocity_old"].tobytes()
        for key in ("spring", "angle", "pair"):
            o0, o1 = off[key][k], off[key][k + 1]
            scale = max(np.max(np.abs(gk[key])), 1e-300)
            assert np.allclose(g[key][o0:o1], gk[key], rtol=1e-10, atol=1e-12 * scale), key


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_batched_adjoint_with_kept_candidates(need_gpu, prec):
    """The ensemble's adjoint with kept candidates (verlet 1, one skin for
    all networks, candidates carried across segments and restored for each
    recomputed segment) gives the one-pass rebuild's samples and
    derivatives bit for bit."""
    from fibrenet_amd import Ensemble
    nets = _batch_members()
    rs = np.random.default_rng(4)
    w = rs.normal(size=(len(nets), len(SAMPLES)))
    group = np.concatenate([_group(n) for n in nets])
    out = []
    for verlet in (0, 1):
        ens = Ensemble(nets, prec)
        ens.set_option("verlet", verlet)
        off = {k: ens.offsets(k) for k in ("spring", "angle", "pair")}
        s = ens.adjoint_forward(N_STEPS, group, SAMPLES, grid_update_freq=GUF, checkpoint_every=50)
        builds = ens.stats()["candidate_builds"]
        g = ens.adjoint_backward(w, int(off["spring"][-1]), int(off["angle"][-1]), int(off["pair"][-1]))
        if verlet == 1:
            assert builds >= 1, builds
        out.append((s.tobytes(), {k: v.tobytes() for k, v in g.items()}))
        ens.close()
    assert out[0][0] == out[1][0]
    for k in out[0][1]:
        assert out[0][1][k] == out[1][1][k], k


def test_batch_reload_restarts_every_network(need_gpu):
    """fn_batch_load puts new records in place: a run from them equals a run
    of a freshly created ensemble bit for bit."""
    from fibrenet_amd import Ensemble
    nets = _batch_members()
    a = Ensemble(nets, "f32")
    a.run_steps(120, GUF, True, True)
    a.reload(nets)
    a.run_steps(120, GUF, True, True)
    b = Ensemble(nets, "f32")
    b.run_steps(120, GUF, True, True)
    for k in range(len(nets)):
        assert a.read(k, "bead").tobytes() == b.read(k, "bead").tobytes()
        assert a.read(k, "spring").tobytes() == b.read(k, "spring").tobytes()


def test_backward_outputs_selection(need_gpu):
    """Asking for the force-scale derivatives alone returns exactly those,
    bit for bit as in the full set, and unknown names are refused."""
    net = adjoint_net()
    w = _weights()
    eng, ids = engine_for(net, "f64")
    eng.adjoint_forward(N_STEPS, _group(net), SAMPLES, grid_update_freq=GUF, checkpoint_every=50)
    full = eng.adjoint_backward(w, len(net.springs), len(net.angles), len(net.pairs))
    part = eng.adjoint_backward(w, len(net.springs), len(net.angles), len(net.pairs), ("spring", "angle", "pair"))
    assert sorted(part) == ["angle", "pair", "spring"]
    for k in part:
        assert part[k].tobytes() == full[k].tobytes(), k
    with pytest.raises(ValueError):
        eng.adjoint_backward(w, len(net.springs), len(net.angles), len(net.pairs), ("stiffness",))


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("kind", ["hs_fl", "lj"])
def test_kept_candidates_give_the_same_gradients(need_gpu, prec, kind):
    """With kept contact candidates (verlet 1) the adjoint's forward run
    carries its candidates from segment to segment and each recomputed
    segment restores the ones its start held, so both rebuild candidates at
    the same steps and filter them to the one-pass rebuild's lists: samples,
    gradients and the end state equal the exact rebuild's bit for bit, also
    with the widest skin, whose candidates outlive several segments."""
    net = ADJ_NETS[kind]()
    w = _weights()
    out = []
    for verlet, skin in ((0, None), (1, None), (1, 1000)):
        eng, ids = engine_for(net, prec)
        eng.set_option("verlet", verlet)
        if skin is not None:
            eng.set_option("skin", skin)
        s = eng.adjoint_forward(N_STEPS, _group(net), SAMPLES, grid_update_freq=GUF, checkpoint_every=50)
        builds = eng.stats()["candidate_builds"]
        g = eng.adjoint_backward(w, len(net.springs), len(net.angles), len(net.pairs))
        out.append((s.tobytes(), {k: v.tobytes() for k, v in g.items()}, read_beads(eng, ids).tobytes()))
        if verlet == 1:
            assert builds >= 1, builds
        assert eng.get_option("verlet") == verlet
    for o in out[1:]:
        assert out[0][0] == o[0]
        for k in out[0][1]:
            assert out[0][1][k] == o[1][k], k
        assert out[0][2] == o[2]


def test_skin_change_invalidates_the_adjoint_run(need_gpu):
    """The recomputation restores the forward run's candidates, which hold
    for its skin only: a new skin leaves no run to differentiate."""
    net = ADJ_NETS["lj"]()
    eng, _ = engine_for(net, "f32")
    eng.set_option("verlet", 1)
    eng.adjoint_forward(60, _group(net), SAMPLES[:3], grid_update_freq=GUF, checkpoint_every=20)
    eng.set_option("skin", eng.get_option("skin"))          # unchanged: the run stays
    eng.adjoint_backward(_weights()[:3], len(net.springs), len(net.angles), len(net.pairs))
    eng.set_option("skin", eng.get_option("skin") + 100)
    with pytest.raises(FibreNetError):
        eng.adjoint_backward(_weights()[:3], len(net.springs), len(net.angles), len(net.pairs))
