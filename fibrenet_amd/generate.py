"""Synthetic fibre networks in FibreNet's formats.

A Network holds exactly the records FibreNet's loaders would produce from
its model and input files (SimulationManager.cpp:1372-1661 and the settings
of :644-819, :2070-2301, :827-970), so it can be uploaded directly
(api.bind_standard) or written out as text (write_files) and loaded through
the loaders; the two routes give identical records.

Workloads (fibres are straight, bead spacing 1.0, start points uniform in
the box, directions uniform on the sphere, tag = fibre index, loc_id = bead
index along the fibre):
  periodic_hs   - periodic box, HR springs k=100, HR angles k=10 (rest angle
                  pi), penalty contacts K=1000, dt 1e-3 (10,000 beads as
                  200 fibres x 50 in 100x100x20 by default)
  tensile_lj    - finite box, HR_FL springs (k=100, E=1000, failure stress
                  an integer uniform in [50,150]), HR angles k=10, LJ
                  contacts (eps 1, sigma 1, cutoff 2.5) between all three
                  types; beads with x<5 are clamped (mass 0) and beads with
                  x>995 pulled at +0.01 in x (mass 0, velocity set);
                  dt 5e-4 (2,000,000 beads as 20,000 fibres x 100 in
                  1000x1000x50 by default); fibres are placed no closer
                  than the Lennard-Jones minimum 2^(1/6) sigma to one
                  another, so no pair starts on the repulsive wall
                  (relax() then drains the attraction's energy)
Positions are written with 17 significant digits, which round-trips.
"""
import ctypes
import math
import os

import numpy as np

from . import records as R
from .api import lib

PI_TEXT = "3.14159265358979"


class Network:
    def __init__(self):
        self.beads = self.info = self.springs = self.angles = self.pairs = None
        self.settings = np.zeros((), R.SETTINGS)
        self.cell_nh = None
        self.grid_update_freq = 8
        self.run_steps = 1
        self.type_names = []
        self.pair_lines = []
        self.mod_lines = []
        self.meta = {}
        self.input_extra = []

    @property
    def n_beads(self):
        return len(self.beads)


def f32(x):
    return float(np.float32(x))


def grid_neighbours(bc_type, cx, cy, cz):
    """Half-shell table of SimulationManager.cpp:2193-2301, shape (C, 13)."""
    z, y, x = np.meshgrid(np.arange(cz), np.arange(cy), np.arange(cx), indexing="ij")
    x, y, z = x.ravel(), y.ravel(), z.ravel()
    offs = [(dx, dy, -1) for dx in (-1, 0, 1) for dy in (-1, 0, 1)] + [(1, dy, 0) for dy in (-1, 0, 1)] + [(0, -1, 0)]
    out = np.full((cx * cy * cz, 13), -1, np.int32)
    for k, (dx, dy, dz) in enumerate(offs):
        nx, ny, nz = x + dx, y + dy, z + dz
        if bc_type == 0:
            nx = np.where(nx < 0, cx - 1, np.where(nx >= cx, 0, nx))
            ny = np.where(ny < 0, cy - 1, np.where(ny >= cy, 0, ny))
            nz = np.where(nz < 0, cz - 1, np.where(nz >= cz, 0, nz))
            out[:, k] = nx + ny * cx + nz * cx * cy
        else:
            ok = (nx >= 0) & (nx < cx) & (ny >= 0) & (ny < cy) & (nz >= 0) & (nz < cz)
            out[:, k] = np.where(ok, nx + ny * cx + nz * cx * cy, -1)
    return out


def make_settings(box, bc_type, unit, cutoff, dt, lam, wall_k=0.0, damping=None, gravity=(0.0, 0.0, 0.0), types=1):
    """SimulationSettings as LoadGlobalForceInfo / LoadSolverInfo /
    LoadSimulationBoxInfo compute them (SimulationManager.cpp:644-819, 2070-2190);
    `damping` is the input-file value (stored negated; None: no damping line)."""
    s = np.zeros((), R.SETTINGS)
    s["types_num"] = types
    s["bc_type"] = bc_type
    s["wall_contact_stiffness"] = wall_k
    s["cutoff"] = cutoff
    s["cutoff_sq"] = cutoff * cutoff
    s["box_unit_cell"] = unit
    s["dt"] = dt
    s["half_dt"] = dt * f32(0.5)
    s["half_dt2"] = dt * dt * f32(0.5)
    s["lambda_dt"] = lam * dt
    s["damping"] = 0.0 if damping is None else damping * f32(-1.0)
    s["gravity"]["x"], s["gravity"]["y"], s["gravity"]["z"] = gravity
    L = [float(v) for v in box]
    s["box_size"]["x"], s["box_size"]["y"], s["box_size"]["z"] = L
    cells = [int(math.floor(v / unit)) for v in L]
    s["cells_x"], s["cells_y"], s["cells_z"] = cells
    s["cells_xy"] = cells[0] * cells[1]
    for c, v, n in zip("xyz", L, cells):
        s["grid_step"][c] = v / f32(n)
    h = min(min(L[0], L[1]), L[2]) * f32(0.5)
    s["half_box_sq"] = h * h
    return s


def lj_pair(a, b, sigma, eps, rc):
    p = np.zeros((), R.PAIR)
    p["type"], p["p1"], p["p2"] = 1, a, b
    p["c1"], p["c2"], p["c3"] = sigma, eps, rc
    r2 = rc * rc
    s2 = sigma * sigma
    fr2 = s2 / r2
    fr6 = fr2 * fr2 * fr2
    p["c4"] = 48.0 * eps * fr6 * (fr6 - 0.5) / r2
    return p


def hs_pair(a, b, k):
    p = np.zeros((), R.PAIR)
    p["type"], p["p1"], p["p2"], p["c1"] = 2, a, b, k
    return p


def st_pair(a, b, t4, t5, t6, t7):
    """A two-step harmonic ("st") entry from its input tokens
    `pair st A B t4 t5 t6 t7` (SimulationManager.cpp:922-933): c1 = t5,
    c2 = t6, c3 = t7 - t6, c4 = t7^2, c5 = t4, c6 = 2 t6 - t7; the force
    c1 (c2 - r) + c5 (c6 - r) acts while r^2 < c4 (kernel.cl:960-999)."""
    p = np.zeros((), R.PAIR)
    p["type"], p["p1"], p["p2"] = 3, a, b
    p["c1"], p["c2"], p["c3"] = t5, t6, t7 - t6
    p["c4"], p["c5"], p["c6"] = t7 * t7, t4, 2 * t6 - t7
    return p


# The README example's st tokens (README.md:590-595): t4 2000, t5 500, t6 1.00, t7 1.03.
README_ST = (2000.0, 500.0, 1.00, 1.03)


def pair_index(a, b, T):
    lo, hi = min(a, b), max(a, b)
    return T * lo + hi - lo * (lo + 1) // 2


def mass_coeffs(info, s, mass):
    """Per-bead mass coefficients of ReadBeadsData (SimulationManager.cpp:1433-1445)."""
    mass = np.broadcast_to(np.asarray(mass, np.float64), info.shape)
    info["mass"] = mass
    pos = mass > 0.0
    inv = np.where(pos, f32(1.0) / np.where(pos, mass, 1.0), 0.0)
    info["mass_inv"] = inv
    info["lambda_dt_mass_inv"] = np.where(pos, float(s["lambda_dt"]) * inv, 0.0)
    info["half_dt_mass_inv"] = np.where(pos, float(s["half_dt"]) * inv, 0.0)
    info["half_dt2_mass_inv"] = np.where(pos, float(s["half_dt2"]) * inv, 0.0)


def place_fibres(seed, n_fibres, n_beads, box, finite, spacing=1.0, d_min=0.0, max_tries=None):
    """Unwrapped bead coordinates (n_fibres, n_beads, 3) and the number of
    fibres sampled (fn_gen_fibres, run in the library's host code)."""
    out = np.zeros((n_fibres, n_beads, 3), np.float64)
    tries = ctypes.c_longlong()
    mt = max_tries if max_tries is not None else 10**9
    rc = lib().fn_gen_fibres(int(seed), int(n_fibres), int(n_beads), float(box[0]), float(box[1]), float(box[2]),
                             int(bool(finite)), float(spacing), float(d_min), int(mt),
                             out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(tries))
    if rc != 0:
        raise RuntimeError("fibre placement failed (rc %d after %d tries)" % (rc, tries.value))
    return out, tries.value


def _chain_bonds(n_fibres, n_beads, spring_type, k, l0, e_mod=None, sigma_f=None, angle_k=10.0, angle_type=0):
    """HR / HR_FL springs (j, j+1) and angles (j, j+1, j+2) along each fibre."""
    f = np.arange(n_fibres)[:, None]
    j = np.arange(n_beads - 1)[None, :]
    S = n_fibres * (n_beads - 1)
    sp = np.zeros(S, R.SPRING)
    sp["type"] = spring_type
    sp["tag"] = np.repeat(np.arange(n_fibres), n_beads - 1)
    sp["status"] = 1
    sp["p1"] = (f * n_beads + j).ravel()
    sp["p2"] = (f * n_beads + j + 1).ravel()
    sp["rest_length"] = l0
    sp["stiffness"] = k
    if spring_type == 1:
        sp["c1"] = e_mod
        sp["c2"] = sigma_f
    ja = np.arange(n_beads - 2)[None, :]
    A = n_fibres * max(n_beads - 2, 0)
    an = np.zeros(A, R.ANGLE)
    if A:
        an["type"] = angle_type
        an["tag"] = np.repeat(np.arange(n_fibres), n_beads - 2)
        an["status"] = 1
        base = f * n_beads + ja
        an["b1"], an["b2"], an["b3"] = base.ravel(), (base + 1).ravel(), (base + 2).ravel()
        sbase = f * (n_beads - 1) + ja
        an["s1"], an["s2"] = sbase.ravel(), (sbase + 1).ravel()
        an["angle"] = f32(PI_TEXT)
        an["c1"] = angle_k
    return sp, an


def periodic_hs(n_fibres=200, n_beads=50, box=(100.0, 100.0, 20.0), seed=0xF1B0, dt=1e-3, steps=10000,
                spring_k=100.0, angle_k=10.0, contact_k=1000.0, radius=0.5, mass=1.0, grid_update_freq=8,
                cell_table=True):
    net = Network()
    s = make_settings(box, 0, 1.0, 1.0, dt, 0.5, wall_k=100.0)
    xyz, tries = place_fibres(seed, n_fibres, n_beads, box, finite=False)
    img = xyz.reshape(-1, 3)
    L = np.array(box, np.float64)
    pos = np.mod(img, L)
    pos = np.where(pos >= L, 0.0, pos)   # keep [0, L) after rounding
    N = len(img)
    net.beads = np.zeros(N, R.BEAD)
    R.set_vec(net.beads, "position", pos)
    R.set_vec(net.beads, "image_position", img)
    net.info = np.zeros(N, R.BEAD_INFO)
    net.info["type"] = 0
    net.info["tag"] = np.repeat(np.arange(n_fibres), n_beads)
    net.info["loc_id"] = np.tile(np.arange(n_beads), n_fibres)
    net.info["radius"] = radius
    net.info["radius_sq"] = radius * radius
    net.info["skin_radius"] = radius + 0.5
    net.info["skin_radius_sq"] = (radius + 0.5) ** 2
    mass_coeffs(net.info, s, mass)
    net.springs, net.angles = _chain_bonds(n_fibres, n_beads, 0, spring_k, 1.0, angle_k=angle_k)
    net.pairs = np.array([hs_pair(0, 0, contact_k)], R.PAIR)
    net.settings = s
    if cell_table:   # the GridCellNh buffer of the per-kernel path; an ensemble does not use it
        net.cell_nh = grid_neighbours(0, int(s["cells_x"]), int(s["cells_y"]), int(s["cells_z"]))
    net.grid_update_freq = grid_update_freq
    net.run_steps = steps
    net.type_names = ["C1"]
    net.pair_lines = ["pair hs C1 C1 %r" % contact_k]
    net.input_extra = ["box_type pbc", "box_wall_rep 100", "box_unit_cell 1.0", "cutoff 1.0"]
    net.meta = dict(workload="periodic_hs", n_fibres=n_fibres, n_beads=n_beads, box=list(box), seed=seed,
                    fibre_tries=tries, dt=dt)
    return net


def tensile_lj(n_fibres=20000, n_beads=100, box=(1000.0, 1000.0, 50.0), seed=0xF1B1, dt=5e-4, steps=20000,
               spring_k=100.0, e_mod=1000, angle_k=10.0, eps=1.0, sigma=1.0, rc=2.5, clamp_x=5.0, pull_v=0.01,
               d_min=2.0 ** (1.0 / 6.0), grid_update_freq=10, radius=0.5, mass=1.0):
    net = Network()
    s = make_settings(box, 1, 2.5, rc, dt, 0.5, wall_k=100.0, types=3)
    xyz, tries = place_fibres(seed, n_fibres, n_beads, box, finite=True, d_min=d_min)
    pos = xyz.reshape(-1, 3)
    N = len(pos)
    net.beads = np.zeros(N, R.BEAD)
    R.set_vec(net.beads, "position", pos)
    R.set_vec(net.beads, "image_position", pos)
    lo = pos[:, 0] < clamp_x
    hi = pos[:, 0] > box[0] - clamp_x
    typ = np.where(lo, 1, np.where(hi, 2, 0)).astype(np.int32)
    net.info = np.zeros(N, R.BEAD_INFO)
    net.info["type"] = typ
    net.info["tag"] = np.repeat(np.arange(n_fibres), n_beads)
    net.info["loc_id"] = np.tile(np.arange(n_beads), n_fibres)
    net.info["radius"] = radius
    net.info["radius_sq"] = radius * radius
    net.info["skin_radius"] = radius + 0.5
    net.info["skin_radius_sq"] = (radius + 0.5) ** 2
    mass_coeffs(net.info, s, np.where(typ > 0, 0.0, mass))
    v = np.zeros((N, 3))
    v[hi, 0] = pull_v
    R.set_vec(net.beads, "velocity", v)
    R.set_vec(net.beads, "velocity_old", v)
    rs = np.random.Generator(np.random.PCG64(seed))
    sigma_f = rs.integers(50, 151, size=n_fibres * (n_beads - 1)).astype(np.float64)
    net.springs, net.angles = _chain_bonds(n_fibres, n_beads, 1, spring_k, 1.0, e_mod=float(e_mod),
                                           sigma_f=sigma_f, angle_k=angle_k)
    T = 3
    net.pairs = np.zeros(T * (T + 1) // 2, R.PAIR)
    for a in range(T):
        for b in range(a, T):
            net.pairs[pair_index(a, b, T)] = lj_pair(a, b, sigma, eps, rc)
    net.settings = s
    net.cell_nh = grid_neighbours(1, int(s["cells_x"]), int(s["cells_y"]), int(s["cells_z"]))
    net.grid_update_freq = grid_update_freq
    net.run_steps = steps
    net.type_names = ["C1", "C2", "C3"]
    net.pair_lines = ["pair lj %s %s %r %r %r" % (net.type_names[a], net.type_names[b], sigma, eps, rc)
                      for a in range(T) for b in range(a, T)]
    # the model file holds the nominal mass; the clamp and pull come from modifiers
    net.model_mass = mass
    net.mod_lines = ["modif_bead type C2 mass 0.0", "modif_bead type C3 mass 0.0",
                     "modif_bead type C3 velocity %r 0.0 0.0" % pull_v]
    net.input_extra = ["box_type finite", "box_wall_rep 100", "box_unit_cell 2.5", "cutoff %r" % rc,
                       "grid_update_freq %d" % grid_update_freq]
    net.meta = dict(workload="tensile_lj", n_fibres=n_fibres, n_beads=n_beads, box=list(box), seed=seed,
                    fibre_tries=tries, d_min=d_min, dt=dt, clamped=int(lo.sum()), pulled=int(hi.sum()))
    return net


def st_network(n_fibres=150, n_beads=30, box=(40.0, 40.0, 20.0), seed=0xF1B5, dt=1e-4, steps=1000,
               tokens=README_ST, cutoff=1.15, unit=1.25, damping=0.03, spring_k=2000.0, angle_k=50.0,
               grid_update_freq=8, d_min=0.9):
    """Fibres with the README example's contact law: two-step harmonic
    ("st") contacts between one bead type with the README's tokens, cutoff
    1.15 and grid cell 1.25 (README.md:565-606), damping 0.03, dt 1e-4, in a
    finite box with walls; HR springs and HR angles along straight fibres
    placed as in periodic_hs (they cross, so contacts form at every
    distance below the cutoff)."""
    net = Network()
    s = make_settings(box, 1, unit, cutoff, dt, 0.5, wall_k=100.0, damping=damping)
    xyz, tries = place_fibres(seed, n_fibres, n_beads, box, finite=True, d_min=d_min)
    pos = xyz.reshape(-1, 3)
    N = len(pos)
    net.beads = np.zeros(N, R.BEAD)
    R.set_vec(net.beads, "position", pos)
    R.set_vec(net.beads, "image_position", pos)
    net.info = np.zeros(N, R.BEAD_INFO)
    net.info["tag"] = np.repeat(np.arange(n_fibres), n_beads)
    net.info["loc_id"] = np.tile(np.arange(n_beads), n_fibres)
    net.info["radius"] = 0.5
    net.info["radius_sq"] = 0.25
    net.info["skin_radius"] = 1.0
    net.info["skin_radius_sq"] = 1.0
    mass_coeffs(net.info, s, 1.0)
    net.springs, net.angles = _chain_bonds(n_fibres, n_beads, 0, spring_k, 1.0, angle_k=angle_k)
    net.pairs = np.array([st_pair(0, 0, *tokens)], R.PAIR)
    net.settings = s
    net.cell_nh = grid_neighbours(1, int(s["cells_x"]), int(s["cells_y"]), int(s["cells_z"]))
    net.grid_update_freq = grid_update_freq
    net.run_steps = steps
    net.type_names = ["C1"]
    net.pair_lines = ["pair st C1 C1 %r %r %r %r" % tuple(tokens)]
    net.input_extra = ["box_type finite", "box_wall_rep 100", "box_unit_cell %r" % unit, "cutoff %r" % cutoff,
                       "damping %r" % damping, "grid_update_freq %d" % grid_update_freq]
    net.meta = dict(workload="st_network", n_fibres=n_fibres, n_beads=n_beads, box=list(box), seed=seed,
                    fibre_tries=tries, dt=dt)
    return net


def relax(net, steps=4000, substeps=200, precision="f64", device=0):
    """The network after SimulationManager::MinimizeSystemEnergy with reset
    (SimulationManager.cpp:111-143): ResetBeadEnergy, then steps // substeps
    runs of `substeps` steps, each followed by ResetBeadEnergy, on the HIP
    engine.  Returns a copy of `net` holding the relaxed beads and bond
    statuses.  ResetBeadEnergy also zeroes the velocities of massless beads
    (kernel.cl:468-483), which would stop a pull; as boundary conditions,
    the massless beads keep the velocities they had before the
    minimisation (they do not move during it)."""
    import copy
    from .api import Engine, bind_standard
    out = copy.copy(net)
    eng = Engine(precision, device)
    try:
        ids = bind_standard(eng, net)
        spr, ang = len(net.springs) > 0, len(net.angles) > 0
        eng.launch("ResetBeadEnergy")
        for _ in range(steps // substeps):
            eng.run_steps(substeps, net.grid_update_freq, spr, ang)
            eng.launch("ResetBeadEnergy")
        beads = eng.read_buffer(ids["bead"], R.BEAD).copy()
        out.springs = eng.read_buffer(ids["spring"], R.SPRING).copy() if spr else net.springs
        out.angles = eng.read_buffer(ids["angle"], R.ANGLE).copy() if ang else net.angles
    finally:
        eng.close()
    fixed = np.asarray(net.info["mass"]) == 0.0
    for f in ("velocity", "velocity_old"):
        beads[f][fixed] = net.beads[f][fixed]
    out.beads = beads
    out.meta = dict(net.meta, relax_steps=steps, relax_substeps=substeps, relax_precision=precision)
    return out


def tensile_fit(n_fibres=100, n_beads=50, box=(100.0, 100.0, 20.0), seed=0xF1B3, dt=1e-3, steps=2000, jitter=0.1,
                clamp_x=5.0, pull_v=0.01, hr_fl=False, spring_k=100.0, angle_k=10.0, contact_k=1000.0, e_mod=1000.0,
                sigma_f=(50.0, 150.0), damping=None, gravity=(0.0, 0.0, 0.0), unit=2.0, grid_update_freq=10,
                d_min=0.0, contact="hs", periodic=False, st_tokens=(200.0, 50.0, 1.0, 1.03)):
    """A differentiable tensile case: bent fibres (straight chains with a
    random transverse jitter, so that angles and springs start away from
    rest and every stiffness shapes the response) in a finite box with
    walls; HR_FL springs (failure stress uniform in `sigma_f`) or HR
    springs, HR angles, penalty contacts between three bead types; type 1
    clamped (mass 0) at x < clamp_x, type 2 pulled at +pull_v (mass 0) at
    x > box - clamp_x.  Fibres are placed as in periodic_hs (d_min = 0:
    overlaps allowed) unless d_min asks for rejection sampling.  contact
    "lj": Lennard-Jones contacts (eps 1, sigma 1, cutoff 2.5) instead of
    penalty ones, "st": two-step harmonic contacts of the tokens st_tokens
    (cutoff 1.15); periodic: a periodic box (images unwrapped, positions
    wrapped; the clamp and pull groups keep their masses and velocities)."""
    net = Network()
    cut = {"lj": 2.5, "st": 1.15}.get(contact, 1.5)
    s = make_settings(box, 0 if periodic else 1, max(unit, cut), cut, dt, 0.5, wall_k=100.0, damping=damping,
                      gravity=gravity, types=3)
    xyz, tries = place_fibres(seed, n_fibres, n_beads, (box[0], box[1], box[2]), finite=not periodic, d_min=d_min)
    rs = np.random.Generator(np.random.PCG64(seed))
    img = xyz.reshape(-1, 3) + rs.uniform(-jitter, jitter, size=(n_fibres * n_beads, 3))
    L = np.array(box)
    if periodic:
        pos = np.mod(img, L)
        pos = np.where(pos >= L, 0.0, pos)
    else:
        img = pos = np.clip(img, 0.05, L - 0.05)
    N = len(pos)
    net.beads = np.zeros(N, R.BEAD)
    R.set_vec(net.beads, "position", pos)
    R.set_vec(net.beads, "image_position", img)
    lo = pos[:, 0] < clamp_x
    hi = pos[:, 0] > box[0] - clamp_x
    typ = np.where(lo, 1, np.where(hi, 2, 0)).astype(np.int32)
    net.info = np.zeros(N, R.BEAD_INFO)
    net.info["type"] = typ
    net.info["tag"] = np.repeat(np.arange(n_fibres), n_beads)
    net.info["loc_id"] = np.tile(np.arange(n_beads), n_fibres)
    net.info["radius"] = 0.5
    net.info["radius_sq"] = 0.25
    net.info["skin_radius"] = 1.0
    net.info["skin_radius_sq"] = 1.0
    mass_coeffs(net.info, s, np.where(typ > 0, 0.0, 1.0))
    v = np.zeros((N, 3))
    v[hi, 0] = pull_v
    R.set_vec(net.beads, "velocity", v)
    R.set_vec(net.beads, "velocity_old", v)
    if hr_fl:
        sf = rs.uniform(sigma_f[0], sigma_f[1], size=n_fibres * (n_beads - 1))
        net.springs, net.angles = _chain_bonds(n_fibres, n_beads, 1, spring_k, 1.0, e_mod=e_mod, sigma_f=sf,
                                               angle_k=angle_k)
    else:
        net.springs, net.angles = _chain_bonds(n_fibres, n_beads, 0, spring_k, 1.0, angle_k=angle_k)
    T = 3
    net.pairs = np.zeros(T * (T + 1) // 2, R.PAIR)
    for a in range(T):
        for b in range(a, T):
            net.pairs[pair_index(a, b, T)] = (lj_pair(a, b, 1.0, 1.0, 2.5) if contact == "lj" else
                                              st_pair(a, b, *st_tokens) if contact == "st"
                                              else hs_pair(a, b, contact_k))
    net.settings = s
    net.cell_nh = grid_neighbours(0 if periodic else 1, int(s["cells_x"]), int(s["cells_y"]), int(s["cells_z"]))
    net.grid_update_freq = grid_update_freq
    net.run_steps = steps
    net.type_names = ["C1", "C2", "C3"]
    net.meta = dict(workload="tensile_fit", n_fibres=n_fibres, n_beads=n_beads, box=list(box), seed=seed,
                    fibre_tries=tries, dt=dt, jitter=jitter, clamped=int(lo.sum()), pulled=int(hi.sum()))
    return net


def tensile_batch(n_networks=128, seed=0xF1B4, fibres=(50, 300), lengths=(30, 60), hidden=(50.0, 200.0),
                  density=0.025, height=20.0, jitter=0.05, dt=1e-3, steps=1000):
    """Independent tensile networks of uneven size for a batched fit: per
    network a fibre count and a fibre length uniform in the given ranges, a
    finite box of the given height whose square base holds the beads at
    `density`, HR_FL springs (failure stress uniform in [50, 150]), HR
    angles, penalty contacts, clamp and pull as in tensile_fit.  Each
    network's meta carries its hidden spring stiffness, log-uniform in
    `hidden`.  Network i uses the seed seed + 3000 + i."""
    rs = np.random.Generator(np.random.PCG64(seed))
    nets = []
    for i in range(n_networks):
        nf = int(rs.integers(fibres[0], fibres[1] + 1))
        nb = int(rs.integers(lengths[0], lengths[1] + 1))
        side = float(max(2.0 * nb, np.sqrt(nf * nb / (height * density))))
        k = float(np.exp(rs.uniform(np.log(hidden[0]), np.log(hidden[1]))))
        net = tensile_fit(n_fibres=nf, n_beads=nb, box=(side, side, height), seed=seed + 3000 + i, dt=dt,
                          steps=steps, jitter=jitter, hr_fl=True)
        net.meta.update(network=i, hidden_spring_k=k, side=side)
        nets.append(net)
    return nets


def _g(x):
    return "%.17g" % x


def write_files(net, directory, name="model"):
    """Writes <name>_model.txt and <name>_input.txt (tab-delimited as in
    README.md:34-120) and returns the input file path."""
    os.makedirs(directory, exist_ok=True)
    mpath = os.path.join(directory, name + "_model.txt")
    ipath = os.path.join(directory, name + "_input.txt")
    b, q = net.beads, net.info
    mass = getattr(net, "model_mass", None)
    lines = []
    for i in range(len(b)):
        m = q["mass"][i] if mass is None else mass
        vel = b["velocity"][i] if not net.mod_lines else np.zeros((), R.VEC3)
        row = [net.type_names[q["type"][i]], str(q["tag"][i]), str(q["loc_id"][i]), _g(q["radius"][i]), _g(m),
               _g(b["position"]["x"][i]), _g(b["position"]["y"][i]), _g(b["position"]["z"][i]),
               _g(vel["x"]), _g(vel["y"]), _g(vel["z"]),
               _g(b["force"]["x"][i]), _g(b["force"]["y"][i]), _g(b["force"]["z"][i]),
               _g(b["image_position"]["x"][i]), _g(b["image_position"]["y"][i]), _g(b["image_position"]["z"][i])]
        lines.append("BEAD\t" + "\t".join(row))
    for sp in net.springs:
        if sp["type"] == 0:
            lines.append("SPRING\tHR\t%d\t%d\t%d\t%s\t%s" % (sp["tag"], sp["p1"], sp["p2"], _g(sp["rest_length"]),
                                                           _g(sp["stiffness"])))
        else:
            lines.append("SPRING\tHR_FL\t%d\t%d\t%d\t%s\t%s\t%d\t%d" % (
                sp["tag"], sp["p1"], sp["p2"], _g(sp["rest_length"]), _g(sp["stiffness"]), int(sp["c1"]),
                int(sp["c2"])))
    for a in net.angles:
        lines.append("ANGLE\t%s\t%d\t%d\t%d\t%d\t%d\t%d\t%s\t%s" % ("HR" if a["type"] == 0 else "HR_FL", a["tag"],
                                                                  a["b1"], a["b2"], a["b3"], a["s1"], a["s2"],
                                                                  PI_TEXT, _g(a["c1"])))
    with open(mpath, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    s = net.settings
    L = s["box_size"]
    inp = ["model_file\t%s" % os.path.basename(mpath), "output_file\t%s_out" % name,
           "output_freq\t%d" % net.run_steps, "output_data\tposition", "output_type\tall",
           "run\t%d" % net.run_steps, "lambda\t0.5", "step\t%s" % _g(s["dt"]),
           "box_size\t%s\t%s\t%s" % (_g(L["x"]), _g(L["y"]), _g(L["z"]))]
    inp += ["type\t%s" % t for t in net.type_names]
    inp += net.pair_lines + net.input_extra + net.mod_lines
    with open(ipath, "w") as fh:
        fh.write("\n".join(inp) + "\n")
    return ipath


def ensemble(n_networks=512, seed=0xF1B0, fibres=(20, 400), lengths=(20, 80), side=(60.0, 200.0),
             volume_fraction=0.05, stiffness=(30.0, 300.0), dt=1e-3):
    """Independent periodic networks of uneven size: per network a fibre count
    and a fibre length uniform in the given ranges, a cubic box whose side
    gives the volume fraction (clamped to `side`), a spring stiffness
    log-uniform in `stiffness`; angles and penalty contacts as in
    periodic_hs.  Network i uses the seed seed + 2000 + i."""
    rs = np.random.Generator(np.random.PCG64(seed))
    nets = []
    for i in range(n_networks):
        nf = int(rs.integers(fibres[0], fibres[1] + 1))
        nb = int(rs.integers(lengths[0], lengths[1] + 1))
        n = nf * nb
        L = float(np.clip((n * (np.pi / 6.0) / volume_fraction) ** (1.0 / 3.0), side[0], side[1]))
        k = float(np.exp(rs.uniform(np.log(stiffness[0]), np.log(stiffness[1]))))
        net = periodic_hs(n_fibres=nf, n_beads=nb, box=(L, L, L), seed=seed + 2000 + i, dt=dt, spring_k=k,
                          cell_table=False)
        net.meta.update(network=i, spring_k=k, side=L)
        nets.append(net)
    return nets

