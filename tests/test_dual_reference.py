"""Pins of the fp64 dual-number reference (oracle/dual_reference.py), the
checker of the adjoint: its values follow the C oracle's trajectory, and
its derivatives equal central differences of its own forward (taken with
exact square roots, where the forward is smooth and the differences
converge as h^2)."""
import copy

import numpy as np
import pytest

from fibrenet_amd import records as R
from helpers import ADJ_NETS, adjoint_net, dual_run, oracle_for, rel_err

STEPS = 150
GUF = 10


def _samples(net):
    group = (net.info["type"] == 1).astype(np.int32)
    return group, np.arange(9, STEPS, 10)


@pytest.mark.parametrize("kind", sorted(ADJ_NETS))
def test_values_follow_the_oracle(kind):
    net = ADJ_NETS[kind]()
    hr_fl = kind == "hs_fl"
    group, steps = _samples(net)
    d = dual_run(net)
    vals, _ = d.run(300, GUF, group, np.arange(14, 300, 15))
    o = oracle_for(net)
    o.run(300, GUF)
    st = d.state()
    for f in ("position", "image_position", "velocity_old", "force"):
        assert rel_err(st[f], R.vec(o.beads, f)) < 1e-12, f
    li, lj = d.list_i, d.list_j
    m = li < lj
    assert np.array_equal(np.unique(np.stack([li[m], lj[m]], 1), axis=0), o.pair_set())
    assert np.array_equal(d.sp_status, o.springs["status"])
    assert np.array_equal(d.an_status, o.angles["status"])
    if hr_fl:
        assert np.sum(d.sp_status == 0) > 0          # the case exercises failure
    # the last sample is the x-force on the clamp that the oracle ends with
    f_last = R.vec(o.beads, "force")[group == 1, 0].sum()
    assert abs(vals[-1] - f_last) < 1e-10 * max(1.0, abs(f_last))


@pytest.mark.parametrize("kind", ["hs", "st", "lj"])
def test_derivatives_equal_central_differences(kind):
    """Along the three scale directions and random position and velocity
    directions, for penalty, two-step harmonic and Lennard-Jones contacts."""
    net = ADJ_NETS[kind]()
    N = len(net.beads)
    group, steps = _samples(net)
    rs = np.random.default_rng(5)
    up, uv = rs.normal(size=(N, 3)), rs.normal(size=(N, 3))
    w = rs.normal(size=len(steps))
    _, ders = dual_run(net, up, uv, round_lengths=False).run(STEPS, GUF, group, steps)
    got = ders @ w

    def loss(scales=(1.0, 1.0, 1.0), dpos=0.0, dvel=0.0):
        b = net.beads.copy()
        R.set_vec(b, "position", R.vec(b, "position") + dpos * up)
        R.set_vec(b, "image_position", R.vec(b, "image_position") + dpos * up)
        R.set_vec(b, "velocity_old", R.vec(b, "velocity_old") + dvel * uv)
        nn = copy.copy(net)
        nn.beads = b
        v, _ = dual_run(nn, round_lengths=False, scales=scales).run(STEPS, GUF, group, steps)
        return float(v @ w)

    h = 1e-5
    fd = []
    for k in range(3):
        sp, sm = [1.0] * 3, [1.0] * 3
        sp[k], sm[k] = 1 + h, 1 - h
        fd.append((loss(sp) - loss(sm)) / (2 * h))
    fd.append((loss(dpos=h) - loss(dpos=-h)) / (2 * h))
    fd.append((loss(dvel=h) - loss(dvel=-h)) / (2 * h))
    fd = np.array(fd)
    assert np.all(np.abs(fd - got) <= 1e-6 * np.maximum(np.abs(got), 1.0)), (fd, got)


def test_a_dropped_term_fails_the_difference_check():
    """The check above is sharp: a derivative with the damping term of the
    kick left out (a plausible slip) misses the differences by far."""
    net = adjoint_net()
    group, steps = _samples(net)
    w = np.random.default_rng(5).normal(size=len(steps))
    d = dual_run(net, round_lengths=False)
    orig = d.step

    def step_without_damping_derivative():
        v, f, f_old = orig()
        damp = float(d.s["damping"])
        for c in range(3):      # remove d(damping * v) from the force's derivative
            f[c].d = f[c].d - v[c].d * damp
        return v, f, f_old

    d.step = step_without_damping_derivative
    _, ders = d.run(STEPS, GUF, group, steps)
    h = 1e-5
    lp = dual_run(net, round_lengths=False, scales=(1 + h, 1, 1)).run(STEPS, GUF, group, steps)[0] @ w
    lm = dual_run(net, round_lengths=False, scales=(1 - h, 1, 1)).run(STEPS, GUF, group, steps)[0] @ w
    fd = (lp - lm) / (2 * h)
    assert abs(ders[0] @ w - fd) > 1e-3 * abs(fd)
