"""Pins of the C oracle (oracle/fibrenet_oracle.c) by mathematics the
reference fixes: closed forms of the integrator, gradients of the textbook
potentials (by central finite differences, independent of the restated
formulas), Newton's third law and momentum, brute-force contact sets, the
README fixture, and the capacity quirks of the reference's buckets.
"""
import math
import os

import numpy as np
import pytest

from oracle import loader as OL
from oracle import records as R
from oracle.oracle import Oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def settings(box=(30.0, 30.0, 30.0), bc=0, unit=10.0, cutoff=1.0, dt=0.01, lam=0.5, wall_k=0.0, damping=0.0,
             gravity=(0.0, 0.0, 0.0), types=1):
    s = np.zeros((), R.SETTINGS)
    s["types_num"], s["bc_type"], s["wall_contact_stiffness"] = types, bc, wall_k
    s["cutoff"], s["cutoff_sq"], s["box_unit_cell"] = cutoff, cutoff * cutoff, unit
    s["dt"], s["half_dt"], s["half_dt2"], s["lambda_dt"] = dt, dt * 0.5, dt * dt * 0.5, lam * dt
    s["damping"] = -damping
    s["gravity"]["x"], s["gravity"]["y"], s["gravity"]["z"] = gravity
    for c, v in zip("xyz", box):
        s["box_size"][c] = v
    n = [int(math.floor(v / unit)) for v in box]
    s["cells_x"], s["cells_y"], s["cells_z"] = n
    s["cells_xy"] = n[0] * n[1]
    for c, v, k in zip("xyz", box, n):
        s["grid_step"][c] = v / k
    s["half_box_sq"] = (min(box) * 0.5) ** 2
    return s


def beads_of(pos, img=None, vel=None, force=None, mass=1.0, s=None, types=0, tags=None, locs=0, radius=0.5,
             skin=1.0):
    pos = np.atleast_2d(np.asarray(pos, np.float64))
    n = len(pos)
    b = np.zeros(n, R.BEAD)
    for f, v in (("position", pos), ("image_position", pos if img is None else img)):
        b[f]["x"], b[f]["y"], b[f]["z"] = np.asarray(v).T
    if vel is not None:
        v = np.broadcast_to(np.asarray(vel, np.float64), (n, 3))
        for f in ("velocity", "velocity_old"):
            b[f]["x"], b[f]["y"], b[f]["z"] = v.T
    if force is not None:
        v = np.broadcast_to(np.asarray(force, np.float64), (n, 3))
        for f in ("force", "force_old"):
            b[f]["x"], b[f]["y"], b[f]["z"] = v.T
    q = np.zeros(n, R.BEAD_INFO)
    q["type"] = types
    q["tag"] = np.arange(n) if tags is None else tags
    q["loc_id"] = locs
    q["radius"], q["skin_radius"] = radius, skin
    m = np.broadcast_to(np.asarray(mass, np.float64), (n,))
    q["mass"] = m
    inv = np.where(m > 0, 1.0 / np.where(m > 0, m, 1), 0.0)
    q["mass_inv"] = inv
    q["lambda_dt_mass_inv"] = s["lambda_dt"] * inv
    q["half_dt_mass_inv"] = s["half_dt"] * inv
    q["half_dt2_mass_inv"] = s["half_dt2"] * inv
    return b, q


def pair(kind, a=0, b=0, **c):
    p = np.zeros((), R.PAIR)
    p["type"], p["p1"], p["p2"] = {"lj": 1, "hs": 2, "st": 3}[kind], a, b
    for k, v in c.items():
        p[k] = v
    return p


def springs_of(rows):
    sp = np.zeros(len(rows), R.SPRING)
    for i, r in enumerate(rows):
        for k, v in r.items():
            sp[k][i] = v
        sp["status"][i] = 1
    return sp


def angles_of(rows):
    a = np.zeros(len(rows), R.ANGLE)
    for i, r in enumerate(rows):
        for k, v in r.items():
            a[k][i] = v
        a["status"][i] = 1
    return a


def vec(a, f):
    return R.xyz(a, f)


NOPAIR = np.array([pair("hs", c1=0.0)], R.PAIR)


# ---- integrator ------------------------------------------------------------------------------

def test_free_bead_constant_force_is_exact():
    """Velocity-Verlet is exact for a constant force: x = x0 + v0 t + g t^2/(2m)."""
    g = np.array([0.3, -0.2, 0.1])
    m, dt, n = 2.0, 0.01, 50
    s = settings(dt=dt, gravity=tuple(g))
    x0, v0 = np.array([10.0, 11.0, 12.0]), np.array([0.5, -0.25, 0.125])
    b, q = beads_of(x0, vel=v0, force=g, mass=m, s=s)
    o = Oracle(b, q, springs_of([]), angles_of([]), NOPAIR, s)
    o.run(n)
    t = n * dt
    assert np.allclose(vec(o.beads, "position")[0], x0 + v0 * t + 0.5 * g / m * t * t, rtol=0, atol=1e-12)
    assert np.allclose(vec(o.beads, "velocity_old")[0], v0 + g / m * t, rtol=0, atol=1e-13)
    # the stored velocity is the predicted one of the last step: v_old(n-1) + lambda*dt/m * g
    assert np.allclose(vec(o.beads, "velocity")[0], v0 + g / m * (t - dt) + 0.5 * dt / m * g, atol=1e-13)
    assert np.allclose(vec(o.beads, "image_position")[0], vec(o.beads, "position")[0], atol=0)


def test_damped_bead_matches_linear_recurrence():
    """With damping only, (x, v_old, f) evolves by a fixed 3x3 map per axis;
    the damping force uses the predicted velocity (SimulationManager.cpp:664,
    kernel.cl:205-233)."""
    gamma, m, dt, lam, n = 3.0, 1.5, 0.02, 0.5, 40
    s = settings(dt=dt, lam=lam, damping=gamma)
    x0, v0 = np.array([5.0, 6.0, 7.0]), np.array([1.0, -2.0, 0.5])
    b, q = beads_of(x0, vel=v0, mass=m, s=s)
    o = Oracle(b, q, springs_of([]), angles_of([]), NOPAIR, s)
    o.run(n)
    a, bb, h, D = lam * dt / m, dt * dt * 0.5 / m, dt * 0.5 / m, -gamma
    M = np.array([[1.0, dt, bb], [0.0, 1.0 + h * D, h * (D * a + 1.0)], [0.0, D, D * a]])
    st = np.stack([x0, v0, np.zeros(3)])
    ref = np.linalg.matrix_power(M, n) @ st
    assert np.allclose(vec(o.beads, "position")[0], ref[0], atol=1e-12)
    assert np.allclose(vec(o.beads, "velocity_old")[0], ref[1], atol=1e-12)
    assert np.allclose(vec(o.beads, "force")[0], ref[2], atol=1e-12)


def test_massless_bead_moves_at_constant_velocity():
    """m = 0 zeroes every coefficient (SimulationManager.cpp:1440-1445): a
    clamped or pulled bead ignores forces."""
    s = settings(dt=0.01, gravity=(5.0, 5.0, 5.0))
    b, q = beads_of([[10.0, 10.0, 10.0]], vel=[0.01, 0.0, 0.0], force=[100.0, 0, 0], mass=0.0, s=s)
    o = Oracle(b, q, springs_of([]), angles_of([]), NOPAIR, s)
    o.run(100)
    assert np.allclose(vec(o.beads, "position")[0], [10.0 + 0.01 * 100 * 0.01, 10.0, 10.0], atol=1e-13)


def _lj_coeffs(eps=1.0, sig=1.0, rc=2.5):
    fr6 = (sig * sig / (rc * rc)) ** 3
    return dict(c1=sig, c2=eps, c3=rc, c4=48.0 * eps * fr6 * (fr6 - 0.5) / (rc * rc))


def cadence_case(calls, g=8):
    """Bead B (massless, x-velocity -10, dt 0.01) closes on bead A (massless,
    at rest) from 3.55 away; after step k (0-based) they are 3.55 - 0.1 (k+1)
    apart, inside the LJ cutoff 2.5 from step 10 on.  The pair is listed only
    at a rebuild, so A feels a force at the end iff some rebuild at or after
    step 10 happened.  Returns A's final force and the final distance."""
    s = settings(box=(30.0, 30.0, 30.0), unit=10.0, cutoff=2.5, dt=0.01)
    b, q = beads_of([[10.0, 15.0, 15.0], [13.55, 15.0, 15.0]], s=s, mass=0.0)
    for f in ("velocity", "velocity_old"):
        b[f]["x"][1] = -10.0
    o = Oracle(b, q, springs_of([]), angles_of([]), np.array([pair("lj", **_lj_coeffs())], R.PAIR), s)
    for n in calls:
        o.run(n, g)
    p = vec(o.beads, "position")
    return vec(o.beads, "force")[0], p[1, 0] - p[0, 0]


@pytest.mark.parametrize("calls, listed", [((13,), False),       # rebuilds at steps 0, 8: pair still out
                                           ((17,), True),        # rebuild at step 16 lists it
                                           ((11, 2), True),      # the second call rebuilds at its first step
                                           ((5, 8), False),      # ... at global steps 0, 5 only
                                           ((10, 3, 3), True)])
def test_rebuild_cadence_restarts_per_call(calls, listed):
    """RunSimulationLoop rebuilds at the first step of every call and then
    every grid_update_freq steps of that call (SimulationManager.cpp:151-181,
    Q13); a listed LJ pair then acts with the closed-form force."""
    f, d = cadence_case(calls)
    if not listed:
        assert np.array_equal(f, [0.0, 0.0, 0.0])
        return
    c = _lj_coeffs()
    h = 1e-6
    F = lambda x: -(_lj_v(x + h, c["c2"], c["c1"]) - _lj_v(x - h, c["c2"], c["c1"])) / (2 * h)
    mag = F(d) / d - F(c["c3"]) / c["c3"]
    assert abs(d - (3.55 - 0.1 * sum(calls))) < 1e-12
    assert np.allclose(f, [-mag, 0.0, 0.0], rtol=1e-6, atol=1e-12)


def test_minimize_resets_each_chunk():
    """MinimizeSystemEnergy with reset (SimulationManager.cpp:111-143): zeroed
    forces and velocities before the loop and after each of steps // sub_steps
    chunks (the remainder is dropped).  From rest with f = f_old = 0 a bead
    under constant gravity g moves g dt^2 s (s - 1) / (2 m) per chunk of s
    steps: the first drift sees no force, then each step adds one more
    g dt^2 / m."""
    g, dt, m, sub = -2.0, 0.01, 1.5, 10
    s = settings(box=(50.0, 50.0, 50.0), dt=dt, gravity=(0.0, 0.0, g))
    b, q = beads_of([[20.0, 20.0, 40.0]], s=s, mass=m, force=[0.0, 0.0, g])
    o = Oracle(b, q, springs_of([]), angles_of([]), NOPAIR, s)
    o.minimize(35, sub, True)
    z = vec(o.beads, "position")[0, 2]
    assert abs(z - (40.0 + 3 * g * dt * dt * sub * (sub - 1) / (2 * m))) < 1e-12
    for f in ("velocity", "velocity_old", "force", "force_old"):
        assert np.array_equal(vec(o.beads, f)[0], [0.0, 0.0, 0.0]), f


def test_minimize_without_reset_is_plain_run():
    """Without reset the chunks are plain steps (the file's force column seeds
    the first drift, Q32): constant-force Velocity-Verlet is exact."""
    g, dt, m = -2.0, 0.01, 1.5
    s = settings(box=(50.0, 50.0, 50.0), dt=dt, gravity=(0.0, 0.0, g))
    b, q = beads_of([[20.0, 20.0, 40.0]], s=s, mass=m, force=[0.0, 0.0, g])
    o = Oracle(b, q, springs_of([]), angles_of([]), NOPAIR, s)
    o.minimize(35, 10, False)
    t = 30 * dt
    assert abs(vec(o.beads, "position")[0, 2] - (40.0 + g * t * t / (2 * m))) < 1e-12


def test_periodic_reset_and_walls():
    s = settings(box=(30.0, 30.0, 30.0), bc=0, dt=0.01)
    b, q = beads_of([[0.001, 29.999, 15.0]], vel=[-1.0, 1.0, 0.0], s=s)
    o = Oracle(b, q, springs_of([]), angles_of([]), NOPAIR, s)
    o.advance_position()
    p = vec(o.beads, "position")[0]
    assert p[0] == 30.0 and p[1] == 0.0          # reset to the opposite face, not modulo (Q10)
    assert np.allclose(vec(o.beads, "image_position")[0], [0.001 - 0.01, 29.999 + 0.01, 15.0])
    o.rebuild()
    assert o.cell_of_bead[0] == -1                # a bead exactly at L is outside the grid
    s = settings(box=(30.0, 30.0, 30.0), bc=1, wall_k=100.0, dt=0.01)
    b, q = beads_of([[0.0, 29.999, 15.0]], vel=[-10.0, 1.0, 0.0], s=s)
    o = Oracle(b, q, springs_of([]), angles_of([]), NOPAIR, s)
    o.advance_position()
    p = vec(o.beads, "position")[0]
    f = vec(o.beads, "force")[0]
    assert p[0] < 0 and p[1] > 30.0
    assert np.isclose(f[0], -p[0] * 100.0) and np.isclose(f[1], -(p[1] - 30.0) * 100.0) and f[2] == 0.0


# ---- bonds -----------------------------------------------------------------------------------

def _spring_system(d, l0=1.0, k=100.0, typ=0, c1=0.0, c2=0.0):
    s = settings()
    p1 = np.array([15.0, 15.0, 15.0])
    b, q = beads_of([p1, p1 - d], s=s)
    sp = springs_of([dict(type=typ, p1=0, p2=1, rest_length=l0, stiffness=k, c1=c1, c2=c2)])
    return Oracle(b, q, sp, angles_of([]), NOPAIR, s)


def test_spring_force_is_hookean():
    d = np.array([0.6, -0.9, 0.7])
    L = np.linalg.norm(d)
    o = _spring_system(d, l0=1.0, k=100.0)
    o.resolve_springs()
    f1, f2 = vec(o.beads, "force")
    expected = 100.0 * (1.0 - L) * d / L       # k (L0 - L) u, u from p2 to p1
    assert np.allclose(f1, expected, rtol=1e-6)  # length rounded through float sqrt (Q1)
    assert np.array_equal(f2, -f1)
    assert np.dot(f1, d) < 0                      # stretched: p1 pulled toward p2


@pytest.mark.parametrize("L,sigma_f,breaks", [(1.1, 90.0, True), (1.1, 91.0, False), (0.9, 111.0, True),
                                              (0.9, 112.0, False)])
def test_spring_failure_threshold(L, sigma_f, breaks):
    """HR_FL breaks when |E (L - L0) / L| > sigma_f and then applies no force."""
    d = np.array([L, 0.0, 0.0])
    o = _spring_system(d, l0=1.0, k=100.0, typ=1, c1=1000.0, c2=sigma_f)
    o.resolve_springs(step=7)
    f1 = vec(o.beads, "force")[0]
    stress = 1000.0 * (L - 1.0) / L
    assert (abs(stress) > sigma_f) == breaks
    if breaks:
        assert o.springs["status"][0] == 0 and o.break_step[0] == 7 and np.all(f1 == 0)
    else:
        assert o.springs["status"][0] == 1 and o.break_step[0] == -1
        assert np.isclose(f1[0], 100.0 * (1.0 - L), rtol=1e-6)


def _angle_energy(pos, k, th0):
    ba, bc = pos[0] - pos[1], pos[2] - pos[1]
    c = np.dot(ba, bc) / (np.linalg.norm(ba) * np.linalg.norm(bc))
    return k * (math.acos(max(-1.0, min(1.0, c))) - th0) ** 2


@pytest.mark.parametrize("theta,th0", [(2.0, 2.6), (2.9, 2.4), (1.2, 1.9)])
def test_angle_force_is_minus_gradient(theta, th0):
    """Angle forces equal -grad of k (theta - theta0)^2 (checked by central
    differences) and sum to zero."""
    k = 10.0
    pos = np.array([[15.0 + 1.3, 15.0, 15.0], [15.0, 15.0, 15.0],
                    [15.0 + 0.8 * math.cos(theta), 15.0 + 0.8 * math.sin(theta), 15.2]])
    s = settings()
    b, q = beads_of(pos, s=s)
    an = angles_of([dict(type=0, b1=0, b2=1, b3=2, s1=0, s2=0, angle=th0, c1=k)])
    o = Oracle(b, q, springs_of([dict(p1=0, p2=1, rest_length=1.0, stiffness=0.0)]), an, NOPAIR, s)
    o.resolve_angles()
    f = vec(o.beads, "force")
    h = 1e-6
    for i in range(3):
        for c in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, c] += h
            pm[i, c] -= h
            g = (_angle_energy(pp, k, th0) - _angle_energy(pm, k, th0)) / (2 * h)
            assert abs(f[i, c] + g) <= 1e-5 * max(1.0, abs(g)), (i, c, f[i, c], -g)
    assert np.allclose(f.sum(0), 0.0, atol=1e-12)


def test_angle_zero_at_rest_and_collinear():
    s = settings()
    pos = np.array([[16.0, 15.0, 15.0], [15.0, 15.0, 15.0], [14.0, 15.0, 15.0]])   # theta = pi
    b, q = beads_of(pos, s=s)
    an = angles_of([dict(type=0, b1=0, b2=1, b3=2, angle=2.0, c1=10.0)])
    o = Oracle(b, q, springs_of([dict(p1=0, p2=1)]), an, NOPAIR, s)
    o.resolve_angles()
    assert np.all(vec(o.beads, "force") == 0.0)    # cross product vanishes


def test_hr_fl_angle_dies_with_its_spring():
    s = settings()
    pos = np.array([[16.5, 15.0, 15.0], [15.0, 15.0, 15.0], [14.2, 15.6, 15.0]])
    b, q = beads_of(pos, s=s)
    sp = springs_of([dict(type=1, p1=0, p2=1, rest_length=1.0, stiffness=100.0, c1=1000.0, c2=50.0),
                     dict(type=0, p1=1, p2=2, rest_length=1.0, stiffness=100.0)])
    an = angles_of([dict(type=1, b1=0, b2=1, b3=2, s1=0, s2=1, angle=2.0, c1=10.0),
                    dict(type=0, b1=0, b2=1, b3=2, s1=0, s2=1, angle=2.0, c1=10.0)])
    o = Oracle(b, q, sp, an, NOPAIR, s)
    o.resolve_springs()
    assert o.springs["status"][0] == 0                   # stress 333 > 50
    o.resolve_angles()
    assert list(o.angles["status"]) == [0, 1]            # HR_FL angle dies, HR angle survives (Q23)


# ---- contacts --------------------------------------------------------------------------------

def _two_bead_contact(r, p, cutoff, radius=0.5):
    s = settings(cutoff=cutoff, unit=3.0)
    d = np.array([0.36, 0.48, 0.8])   # unit vector
    p1 = np.array([15.2, 15.1, 15.3])
    b, q = beads_of([p1, p1 - r * d], s=s, radius=radius, skin=radius + 0.5)
    o = Oracle(b, q, springs_of([]), angles_of([]), np.array([p], R.PAIR), s)
    o.rebuild()
    o.resolve_contacts()
    f = vec(o.beads, "force")
    return f, d, len(o.pair_set())


def _lj_v(r, eps, sig):
    return 4 * eps * ((sig / r) ** 12 - (sig / r) ** 6)


@pytest.mark.parametrize("r", [0.95, 1.12, 1.6, 2.4])
def test_lj_contact(r):
    """The reference applies (F(r)/r - F(rc)/rc) along r-hat with F = -V'
    of 4 eps ((s/r)^12 - (s/r)^6) (kernel.cl:875-919, Q19)."""
    eps, sig, rc = 1.3, 1.05, 2.5
    s2 = sig * sig
    fr6 = (s2 / (rc * rc)) ** 3
    c4 = 48.0 * eps * fr6 * (fr6 - 0.5) / (rc * rc)
    f, d, npairs = _two_bead_contact(r, pair("lj", c1=sig, c2=eps, c3=rc, c4=c4), cutoff=rc)
    h = 1e-6
    F = lambda x: -(_lj_v(x + h, eps, sig) - _lj_v(x - h, eps, sig)) / (2 * h)
    mag = F(r) / r - F(rc) / rc
    assert npairs == 1
    assert np.allclose(f[0], mag * d, rtol=2e-6, atol=1e-9)
    assert np.array_equal(f[1], -f[0])


def test_lj_outside_cutoff_not_listed():
    f, d, npairs = _two_bead_contact(2.6, pair("lj", c1=1.0, c2=1.0, c3=2.5), cutoff=2.5)
    assert npairs == 0 and np.all(f == 0)


@pytest.mark.parametrize("r,expect_force", [(0.8, True), (1.2, False)])
def test_hs_contact(r, expect_force):
    """Penalty K (r_i + r_j - r) r-hat; a listed pair without overlap gives no
    force (listing radius max r + min skin = 1.5)."""
    f, d, npairs = _two_bead_contact(r, pair("hs", c1=1000.0), cutoff=1.0)
    assert npairs == 1
    if expect_force:
        assert np.allclose(f[0], 1000.0 * (1.0 - r) * d, rtol=1e-6)
    else:
        assert np.all(f == 0)


def test_st_contact():
    """Two-step harmonic (tokens t4..t7 = 2000 500 1.00 1.03): force
    t5 (t6 - r) + t4 (2 t6 - t7 - r) along r-hat while r < t7 (Q21)."""
    t4, t5, t6, t7 = 2000.0, 500.0, 1.00, 1.03
    p = pair("st", c1=t5, c2=t6, c3=t7 - t6, c4=t7 * t7, c5=t4, c6=2 * t6 - t7)
    for r in (0.9, 1.01):
        f, d, _ = _two_bead_contact(r, p, cutoff=1.15)
        assert np.allclose(f[0], (t5 * (t6 - r) + t4 * (2 * t6 - t7 - r)) * d, rtol=1e-5)
    f, d, _ = _two_bead_contact(1.05, p, cutoff=1.15)
    assert np.all(f == 0)


# ---- whole steps: invariants -------------------------------------------------------------------

def _random_network(seed, n_fib=30, n_bead=12, box=(20.0, 20.0, 10.0), kind="hs", bc=0):
    rng = np.random.default_rng(seed)
    L = np.array(box)
    if kind == "lj":
        # parallel fibres along x on a jittered 1.15 lattice: no near-coincident beads
        slots = np.array([(y, z) for y in np.arange(0.6, L[1] - 0.6, 1.15) for z in np.arange(0.6, L[2] - 0.6, 1.15)])
        yz = slots[rng.permutation(len(slots))[:n_fib]]
        st = np.column_stack([rng.uniform(0.5, L[0] - n_bead * 1.1, n_fib), yz])
        d = np.tile([1.1, 0.0, 0.0], (n_fib, 1))
        img = (st[:, None, :] + np.arange(n_bead)[None, :, None] * d[:, None, :]).reshape(-1, 3)
        img = img + rng.uniform(-0.05, 0.05, img.shape)
    else:
        st = rng.uniform(0, 1, (n_fib, 3)) * L
        z = rng.uniform(-1, 1, n_fib)
        ph = rng.uniform(0, 2 * np.pi, n_fib)
        d = np.stack([np.sqrt(1 - z * z) * np.cos(ph), np.sqrt(1 - z * z) * np.sin(ph), z], 1)
        img = (st[:, None, :] + np.arange(n_bead)[None, :, None] * d[:, None, :]).reshape(-1, 3)
    pos = np.mod(img, L) if bc == 0 else np.clip(img, 0, L)
    img = img if bc == 0 else pos
    s = settings(box=box, bc=bc, unit=2.5 if kind == "lj" else 1.0, cutoff=2.5 if kind == "lj" else 1.0,
                 dt=2e-4, wall_k=100.0)
    tags = np.repeat(np.arange(n_fib), n_bead)
    locs = np.tile(np.arange(n_bead), n_fib)
    b, q = beads_of(pos, img=img, s=s, tags=tags, locs=locs,
                    vel=rng.normal(0, 0.2, (len(pos), 3)))
    sp = springs_of([dict(type=0, p1=f * n_bead + j, p2=f * n_bead + j + 1, rest_length=1.0, stiffness=100.0)
                     for f in range(n_fib) for j in range(n_bead - 1)])
    an = angles_of([dict(type=0, b1=f * n_bead + j, b2=f * n_bead + j + 1, b3=f * n_bead + j + 2,
                         s1=f * (n_bead - 1) + j, s2=f * (n_bead - 1) + j + 1, angle=float(np.float32(np.pi)),
                         c1=10.0) for f in range(n_fib) for j in range(n_bead - 2)])
    p = pair("hs", c1=1000.0) if kind == "hs" else pair("lj", c1=1.0, c2=1.0, c3=2.5, c4=0.0)
    return b, q, sp, an, np.array([p], R.PAIR), s


@pytest.mark.parametrize("kind", ["hs", "lj"])
def test_forces_sum_to_zero_and_momentum_is_conserved(kind):
    b, q, sp, an, pr, s = _random_network(3, kind=kind)
    o = Oracle(b, q, sp, an, pr, s)
    p0 = vec(o.beads, "velocity_old").sum(0)
    for _ in range(20):
        o.run(5, 4)
        f = vec(o.beads, "force")
        assert np.abs(f.sum(0)).max() <= 1e-9 * np.abs(f).sum()
    assert np.allclose(vec(o.beads, "velocity_old").sum(0), p0, atol=1e-9)


@pytest.mark.parametrize("kind,bc", [("hs", 0), ("lj", 1)])
def test_threaded_oracle_matches_sequential(kind, bc):
    """The all-core CPU baseline (work-items on every thread, compare-and-swap
    adds) lists the same pairs and breaks the same springs as the sequential
    oracle, with positions equal up to the summation order of the atomic
    adds."""
    b, q, sp, an, pr, s = _random_network(5, n_fib=60, kind=kind, bc=bc)
    sp["type"] = 1                                  # HR_FL: failure decisions are compared too
    sp["c1"], sp["c2"] = 1000.0, 20.0
    seq = Oracle(b, q, sp, an, pr, s)
    par = Oracle(b, q, sp, an, pr, s, threaded=True)
    for _ in range(6):
        seq.run(10, 4)
        par.run(10, 4)
        assert np.array_equal(seq.pair_set(), par.pair_set())
        assert np.array_equal(seq.springs["status"], par.springs["status"])
        assert np.array_equal(seq.break_step, par.break_step)
        for f in ("position", "velocity_old", "force"):
            ref = vec(seq.beads, f)
            assert np.abs(vec(par.beads, f) - ref).max() <= 1e-9 * max(np.abs(ref).max(), 1.0), f


def _brute_pairs(b, q, pr, s, restrict_cells=True):
    """All (i<j) pairs under BuildContactList's predicate (kernel.cl:574-659),
    optionally limited to beads whose cells are equal or adjacent
    (with wrap in a periodic box)."""
    pos = vec(b, "position")
    n = len(pos)
    L = np.array([s["box_size"][c] for c in "xyz"])
    step = np.array([s["grid_step"][c] for c in "xyz"])
    ncell = np.array([s["cells_x"], s["cells_y"], s["cells_z"]])
    cell = np.floor(pos / step).astype(int)
    inside = np.all((cell >= 0) & (cell < ncell), axis=1)
    hbs = float(s["half_box_sq"])
    out = []
    for i in range(n):
        for j in range(i + 1, n):
            if not (inside[i] and inside[j]):
                continue
            if q["tag"][i] == q["tag"][j] and abs(int(q["loc_id"][i]) - int(q["loc_id"][j])) <= 1:
                continue
            if restrict_cells:
                dc = np.abs(cell[i] - cell[j])
                if s["bc_type"] == 0:
                    dc = np.minimum(dc, ncell - dc)
                if np.any(dc > 1):
                    continue
            r = pos[i] - pos[j]
            l2 = float(r @ r)
            if s["bc_type"] == 0 and l2 > hbs:
                r = np.where(r * r > hbs, r - np.sign(r) * L, r)
                l2 = float(r @ r)
            t = pr[0]["type"]
            if t in (1, 3):
                ok = l2 < s["cutoff_sq"]
            else:
                th = max(q["radius"][i], q["radius"][j]) + min(q["skin_radius"][i], q["skin_radius"][j])
                ok = l2 < th * th
            if ok:
                out.append((i, j))
    return np.array(out, np.int64).reshape(-1, 2)


@pytest.mark.parametrize("kind,bc", [("hs", 0), ("lj", 0), ("lj", 1)])
def test_contact_list_equals_brute_force(kind, bc):
    b, q, sp, an, pr, s = _random_network(5, n_fib=25, n_bead=12, kind=kind, bc=bc)
    o = Oracle(b, q, sp, an, pr, s)
    o.rebuild()
    got = o.pair_set()
    assert np.array_equal(got, _brute_pairs(b, q, pr, s))
    free = _brute_pairs(b, q, pr, s, restrict_cells=False)
    if kind == "lj":
        # listing radius (cutoff 2.5) = cell size: no close pair is missed
        assert np.array_equal(got, free)
    else:
        # hs lists up to max r + min skin = 1.5 > cell 1.0: the reference's
        # cell restriction drops the farther pairs (Q14)
        keys = lambda p: set(map(tuple, p))
        assert keys(got) < keys(free)
    assert len(got) > 10


def test_grid_capacity_quirk():
    """More than 64 beads in a cell: the reference lists only the first 64
    slots (kernel.cl:503,558); the grid head still counts all of them."""
    s = settings(box=(30.0, 30.0, 30.0), unit=10.0, cutoff=1.0)
    rng = np.random.default_rng(1)
    pos = 12.0 + rng.uniform(0, 0.5, (70, 3))
    b, q = beads_of(pos, s=s)
    o = Oracle(b, q, springs_of([]), angles_of([]), np.array([pair("hs", c1=1.0)], R.PAIR), s)
    o.rebuild()
    c = o.cell_of_bead[0]
    assert o.grid_head[c] == 70
    listed = np.unique(o.pair_set())
    assert listed.max() < 64 and len(o.pair_set()) == 64 * 63 // 2


def test_readme_fixture_internal_forces_cancel():
    """README model and input: with walls of stiffness 0 (the box_wall_re typo)
    and no gravity, the total force is damping * sum of predicted velocities."""
    m = OL.load_input(os.path.join(GOLD, "readme_input.txt"))
    m.beads["velocity_old"]["x"] = [0.1, -0.2, 0.05, 0.0, 0.3]
    o = Oracle.from_model(m)
    for _ in range(10):
        o.run(3, m.grid_update_freq)
        f = vec(o.beads, "force").sum(0)
        v = vec(o.beads, "velocity").sum(0)
        assert np.allclose(f, m.settings["damping"] * v, rtol=0, atol=1e-7 * np.abs(vec(o.beads, "force")).sum())
