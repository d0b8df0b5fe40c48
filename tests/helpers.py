"""Shared builders for the parity tests: small networks of the two workloads
and a driver running the HIP engine and the oracle side by side."""
import numpy as np

from fibrenet_amd import Engine, bind_standard, generate as G, records as R, run_loop_granular
from oracle.oracle import Oracle


def small_periodic(seed=11, n_fibres=60, n_beads=20, box=(30.0, 30.0, 12.0)):
    return G.periodic_hs(n_fibres=n_fibres, n_beads=n_beads, box=box, seed=seed)


def small_tensile(seed=12, n_fibres=50, n_beads=20, box=(60.0, 60.0, 30.0), weak=False):
    net = G.tensile_lj(n_fibres=n_fibres, n_beads=n_beads, box=box, seed=seed, d_min=0.6)
    if weak:   # low failure stress so that springs break within a few hundred steps
        rs = np.random.Generator(np.random.PCG64(seed))
        net.springs["c2"] = rs.integers(1, 4, size=len(net.springs)).astype(np.float64)
        net.pull_v = 0.5
        hi = net.info["type"] == 2
        v = np.zeros((len(net.beads), 3))
        v[hi, 0] = 0.5
        R.set_vec(net.beads, "velocity", v)
        R.set_vec(net.beads, "velocity_old", v)
    return net


def dense_lj(seed=14, n_side=6, n_beads=12, spacing=1.1, lattice=1.15):
    """A dense block of the tensile network's fibres: n_side x n_side
    parallel fibres along x on a jittered lattice of 1.15 (no near-coincident
    beads), all free, in a finite 17.5 x 10 x 10 box.  About 30 listed
    partners per bead, so a wave's lists run to thousands of entries (several
    contact passes, the compaction's per-bead writes) and its neighbour search
    to many candidate passes."""
    net = G.tensile_lj(n_fibres=n_side * n_side, n_beads=n_beads, box=(17.5, 10.0, 10.0), seed=seed, d_min=0.6)
    rs = np.random.Generator(np.random.PCG64(seed))
    yz = np.array([(1.5 + a * lattice, 1.5 + b * lattice) for a in range(n_side) for b in range(n_side)])
    x0 = 2.0 + rs.uniform(0.0, 1.0, len(yz))
    img = np.stack([x0[:, None] + spacing * np.arange(n_beads)[None, :],
                    np.repeat(yz[:, :1], n_beads, 1), np.repeat(yz[:, 1:], n_beads, 1)], axis=2).reshape(-1, 3)
    img = img + rs.uniform(-0.05, 0.05, img.shape)
    R.set_vec(net.beads, "position", img)
    R.set_vec(net.beads, "image_position", img)
    R.set_vec(net.beads, "velocity", np.zeros_like(img))
    R.set_vec(net.beads, "velocity_old", np.zeros_like(img))
    net.info["type"] = 0
    G.mass_coeffs(net.info, np.array(net.settings, R.SETTINGS).reshape(()), np.ones(len(img)))
    return net


def crosslinked(seed=13, n_links=60):
    """The tensile network plus HR springs between beads of different fibres
    (both index orders) and its angles listed in reverse: bonds outside the
    chain layout, so the fast path sums bond terms through incidence slots."""
    net = small_tensile(seed=seed, weak=True)
    img = R.vec(net.beads, "image_position")
    fib = np.asarray(net.info["tag"])
    d = np.linalg.norm(img[:, None, :] - img[None, :, :], axis=2)
    i, j = np.nonzero((d < 2.0) & (fib[:, None] != fib[None, :]))
    keep = i < j
    i, j = i[keep][:n_links], j[keep][:n_links]
    extra = np.zeros(len(i), R.SPRING)
    extra["type"] = 0
    extra["status"] = 1
    extra["p1"] = np.where(np.arange(len(i)) % 2 == 0, i, j)
    extra["p2"] = np.where(np.arange(len(i)) % 2 == 0, j, i)
    extra["rest_length"] = 0.9 * d[i, j]
    extra["stiffness"] = 50.0
    net.springs = np.concatenate([net.springs, extra])
    net.angles = net.angles[::-1].copy()
    return net


def engine_for(net, precision):
    eng = Engine(precision)
    ids = bind_standard(eng, net)
    return eng, ids


def oracle_for(net):
    return Oracle(net.beads, net.info, net.springs, net.angles, net.pairs, net.settings, net.cell_nh)


def read_beads(eng, ids):
    return eng.read_buffer(ids["bead"], R.BEAD)


def run_engine(eng, net, n, fast=True):
    spr, ang = len(net.springs) > 0, len(net.angles) > 0
    if fast:
        eng.run_steps(n, net.grid_update_freq, spr, ang)
    else:
        run_loop_granular(eng, n, net.grid_update_freq, spr, ang)


def rel_err(a, b, floor=1.0):
    """max |a-b| / max(floor, max|b|)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(floor, float(np.max(np.abs(b))))) if a.size else 0.0


FIELDS = ("position", "image_position", "velocity", "force", "velocity_old", "force_old")


def field_errors(beads, ref):
    return {f: rel_err(R.vec(beads, f), R.vec(ref, f)) for f in FIELDS}


def adjoint_net(seed=21, n_fibres=40, n_beads=12, box=(24.0, 24.0, 8.0), jitter=0.15, pull_v=0.5, hr_fl=False,
                contact="hs", periodic=False):
    """A small differentiable tensile case (generate.tensile_fit at 480
    beads): bent fibres in a finite box with walls, damping and gravity;
    HR springs (or HR_FL with failure stress in [20, 60]), HR angles,
    penalty (or Lennard-Jones) contacts between three bead types; type 1
    clamped at x < 3, type 2 pulled at +x; or the same in a periodic box."""
    return G.tensile_fit(n_fibres=n_fibres, n_beads=n_beads, box=box, seed=seed, dt=2e-3, steps=300, jitter=jitter,
                         clamp_x=3.0, pull_v=pull_v, hr_fl=hr_fl, sigma_f=(20.0, 60.0), damping=0.5,
                         gravity=(0.0, 0.0, -0.2), d_min=0.9, contact=contact, periodic=periodic)


# Adjoint test networks by kind: penalty contacts in a finite box (with or
# without failures), Lennard-Jones contacts, a periodic box.
def adjoint_xlink(seed=23, n_links=40):
    """adjoint_net plus HR springs between beads of different fibres (both
    index orders) and its angles listed in reverse: bonds outside the chain
    layout, so the adjoint sums bond terms through incidence slots."""
    net = adjoint_net(seed=seed)
    img = R.vec(net.beads, "image_position")
    fib = np.asarray(net.info["tag"])
    d = np.linalg.norm(img[:, None, :] - img[None, :, :], axis=2)
    i, j = np.nonzero((d < 1.6) & (fib[:, None] != fib[None, :]))
    keep = i < j
    i, j = i[keep][:n_links], j[keep][:n_links]
    extra = np.zeros(len(i), R.SPRING)
    extra["type"] = 0
    extra["status"] = 1
    extra["p1"] = np.where(np.arange(len(i)) % 2 == 0, i, j)
    extra["p2"] = np.where(np.arange(len(i)) % 2 == 0, j, i)
    extra["rest_length"] = 0.9 * d[i, j]
    extra["stiffness"] = 20.0
    net.springs = np.concatenate([net.springs, extra])
    net.angles = net.angles[::-1].copy()
    return net


def adjoint_types4(seed=51):
    """adjoint_net with a fourth bead type (the free beads of even fibres) and
    penalty contacts of a different stiffness for each of the 10 type pairs:
    a pair table past the eight entries the one-launch chain adjoint step
    keeps in LDS, so the adjoint runs its bond kernels and per-bead pair
    rows."""
    net = adjoint_net(seed=seed)
    free = net.info["type"] == 0
    net.info["type"] = np.where(free & (net.info["tag"] % 2 == 0), 3, net.info["type"])
    T = 4
    net.settings = net.settings.copy()
    net.settings["types_num"] = T
    net.pairs = np.zeros(T * (T + 1) // 2, R.PAIR)
    for a in range(T):
        for b in range(a, T):
            k = G.pair_index(a, b, T)
            net.pairs[k] = G.hs_pair(a, b, 600.0 + 100.0 * k)
    net.type_names = ["C1", "C2", "C3", "C4"]
    return net


ADJ_NETS = {"hs": lambda: adjoint_net(), "hs_fl": lambda: adjoint_net(hr_fl=True),
            "st": lambda: adjoint_net(seed=43, contact="st"), "types4": lambda: adjoint_types4(),
            "lj": lambda: adjoint_net(seed=41, contact="lj"),
            "periodic": lambda: adjoint_net(seed=41, box=(20.0, 20.0, 8.0), periodic=True),
            "xlink": lambda: adjoint_xlink()}


def dual_run(net, u_pos=None, u_vel=None, round_lengths=True, scales=(1.0, 1.0, 1.0)):
    from oracle.dual_reference import DualRun
    N = len(net.beads)
    up = np.zeros((N, 3)) if u_pos is None else u_pos
    uv = np.zeros((N, 3)) if u_vel is None else u_vel
    return DualRun(net.beads, net.info, net.springs, net.angles, net.pairs, net.settings, up, uv,
                   round_lengths=round_lengths, scales=scales)
