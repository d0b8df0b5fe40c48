"""torch.autograd.Function over the engine's adjoint.

A small harness: the forward run and the backward sweep both execute in the
library's HIP kernels (fn_adjoint_forward / fn_adjoint_backward); torch only
carries the three scalar inputs and the gradients.  The inputs are force
scales relative to a base network (stiffness of every spring, stiffness of
every angle, and every contact pair type's force), so that fitting
stiffnesses is fitting these scales; the output is the clamp's reaction
force (x-force summed over a bead group) at the sample steps.
"""
import numpy as np
import torch

from . import api, records as R
from .api import SCALE_OUTPUTS


def scaled_pairs(pairs, s):
    """Pair table with every contact force multiplied by s: penalty c1, the
    lj energy (c2 and the cutoff shift c4), both st terms (c1, c5)."""
    p = pairs.copy()
    t = p["type"]
    p["c1"] = np.where((t == 2) | (t == 3), p["c1"] * s, p["c1"])
    p["c2"] = np.where(t == 1, p["c2"] * s, p["c2"])
    p["c4"] = np.where(t == 1, p["c4"] * s, p["c4"])
    p["c5"] = np.where(t == 3, p["c5"] * s, p["c5"])
    return p


class TensileRun:
    """One network on one engine; each forward re-uploads its initial state
    with the given scales and runs n_steps recording checkpoints."""

    def __init__(self, net, precision="f32", group=None, sample_steps=(), n_steps=None, grid_update_freq=None,
                 checkpoint_every=0, engine=None):
        self.net = net
        self.eng = engine if engine is not None else api.Engine(precision)
        self.ids = api.bind_standard(self.eng, net)
        self.group = np.asarray(group if group is not None else net.info["type"] == 1, np.int32)
        self.steps = np.asarray(sample_steps, np.int32)
        self.n = int(n_steps if n_steps is not None else net.run_steps)
        self.guf = int(grid_update_freq if grid_update_freq is not None else net.grid_update_freq)
        self.every = int(checkpoint_every)
        self.spr, self.ang = len(net.springs) > 0, len(net.angles) > 0

    def _upload(self, scales):
        ss, sa, sc = (float(x) for x in scales)
        e, net, ids = self.eng, self.net, self.ids
        e.write_buffer(ids["bead"], net.beads)
        if self.spr:
            sp = net.springs.copy()
            sp["stiffness"] = sp["stiffness"] * ss
            e.write_buffer(ids["spring"], sp)
        if self.ang:
            an = net.angles.copy()
            an["c1"] = an["c1"] * sa
            e.write_buffer(ids["angle"], an)
        e.write_buffer(ids["pair"], scaled_pairs(net.pairs, sc))

    def forward(self, scales):
        self._upload(scales)
        return self.eng.adjoint_forward(self.n, self.group, self.steps, self.guf, self.spr, self.ang, self.every)

    def backward(self, dloss, outputs=SCALE_OUTPUTS):
        """The adjoint's derivatives named in `outputs` (by default the
        force-scale ones)."""
        net = self.net
        return self.eng.adjoint_backward(dloss, len(net.springs), len(net.angles), len(net.pairs), outputs)

    def final_beads(self):
        return self.eng.read_buffer(self.ids["bead"], R.BEAD)


class _Tensile(torch.autograd.Function):
    @staticmethod
    def forward(ctx, run, spring_scale, angle_scale, contact_scale):
        scales = [float(x) for x in (spring_scale, angle_scale, contact_scale)]
        ctx.run, ctx.scales = run, scales
        return torch.as_tensor(run.forward(scales), dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad):
        g = ctx.run.backward(grad.detach().to("cpu", torch.float64).numpy())
        ss, sa, sc = ctx.scales
        # the engine's derivatives are per unit relative scale (sigma = 1 + eps):
        # d/d(scale) = sum / scale
        out = [g["spring"].sum() / ss, g["angle"].sum() / sa, g["pair"].sum() / sc]
        return (None,) + tuple(torch.tensor(v, dtype=torch.float64) for v in out)


def tensile_samples(run, spring_scale, angle_scale, contact_scale):
    """Clamp reaction-force samples of `run` as a differentiable function of
    the three scales (0-d float64 tensors)."""
    return _Tensile.apply(run, spring_scale, angle_scale, contact_scale)


class BatchedTensileRun:
    """Independent networks in one batched ensemble, each with its own three
    force scales; each forward reloads every network's initial records with
    its scales (fn_batch_load) and runs the adjoint forward for all of them
    at once; the backward gives every network's scale derivatives."""

    def __init__(self, nets, precision="f32", sample_steps=(), n_steps=None, grid_update_freq=None,
                 checkpoint_every=0):
        self.nets = nets
        self.ens = api.Ensemble(nets, precision)
        self.group = np.concatenate([np.asarray(n.info["type"]) == 1 for n in nets]).astype(np.int32)
        self.steps = np.asarray(sample_steps, np.int32)
        self.n = int(n_steps if n_steps is not None else nets[0].run_steps)
        self.guf = int(grid_update_freq if grid_update_freq is not None else nets[0].grid_update_freq)
        self.every = int(checkpoint_every)
        self.off = {k: self.ens.offsets(k) for k in ("spring", "angle", "pair")}

    def _scaled(self, scales):
        out = []
        for net, (ss, sa, sc) in zip(self.nets, np.asarray(scales, np.float64).reshape(len(self.nets), 3)):
            m = type("Scaled", (), {})()
            m.beads, m.info, m.settings = net.beads, net.info, net.settings
            m.springs = net.springs.copy()
            m.springs["stiffness"] = m.springs["stiffness"] * ss
            m.angles = net.angles.copy()
            m.angles["c1"] = m.angles["c1"] * sa
            m.pairs = scaled_pairs(net.pairs, sc)
            out.append(m)
        return out

    def forward(self, scales):
        """Samples (networks x sample steps) at per-network scales (networks x 3)."""
        self.ens.reload(self._scaled(scales))
        return self.ens.adjoint_forward(self.n, self.group, self.steps, self.guf, True, True, self.every)

    def backward(self, dloss):
        """Per-network derivatives of the loss with respect to each scale at
        sigma = 1 of the scaled run (networks x 3)."""
        nn = len(self.nets)
        g = self.ens.adjoint_backward(np.asarray(dloss, np.float64).reshape(nn, -1), int(self.off["spring"][-1]),
                                      int(self.off["angle"][-1]), int(self.off["pair"][-1]), SCALE_OUTPUTS)
        out = np.zeros((nn, 3))
        for k, key in enumerate(("spring", "angle", "pair")):
            o = self.off[key]
            out[:, k] = np.add.reduceat(np.concatenate([g[key], [0.0]]), o[:-1])[:nn] * (np.diff(o) > 0)
        return out
