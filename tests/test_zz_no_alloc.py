"""fn_group_force_noalloc, the entry point that allocates nothing (GPU): it
returns fn_group_force's result bit for bit, through the C ABI, on two calls
in a row (the second with a smaller engine and group), and refuses a group
larger than its fixed workspace before touching the device."""
import ctypes

import numpy as np
import pytest

from fibrenet_amd import api, generate as G
from helpers import adjoint_net, engine_for

pytestmark = pytest.mark.gpu


def _call(name, eng, group, n_networks=1):
    g = np.ascontiguousarray(group, np.int32)
    out = np.full(n_networks, np.nan)
    rc = getattr(api.lib(), name)(eng.h, g.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p))
    return rc, out


def test_noalloc_group_force_equals_group_force(need_gpu):
    big = G.periodic_hs(n_fibres=16000, n_beads=50, box=(300.0, 300.0, 60.0), seed=71)   # 800,000 beads
    eng, _ = engine_for(big, "f32")
    eng.run_steps(30, big.grid_update_freq)
    group = (np.arange(len(big.beads)) % 7 == 3).astype(np.int32)                     # 114,286 flagged
    rc, a = _call("fn_group_force_noalloc", eng, group)
    assert rc == 0
    rc, b = _call("fn_group_force", eng, group)
    assert rc == 0 and a.tobytes() == b.tobytes() and np.isfinite(a).all()
    eng.close()
    small = adjoint_net(hr_fl=True)
    e2, _ = engine_for(small, "f64")
    e2.run_steps(50, 10)
    g2 = (small.info["type"] == 1).astype(np.int32)
    rc, a2 = _call("fn_group_force_noalloc", e2, g2)
    assert rc == 0
    rc, b2 = _call("fn_group_force", e2, g2)
    assert rc == 0 and a2.tobytes() == b2.tobytes() and a2[0] != 0.0
    rc, z = _call("fn_group_force_noalloc", e2, np.zeros_like(g2))
    assert rc == 0 and z[0] == 0.0


def test_noalloc_group_force_on_an_ensemble(need_gpu):
    from fibrenet_amd import Ensemble
    nets = [adjoint_net(seed=61 + k, n_fibres=20 + 5 * k) for k in range(3)]
    ens = Ensemble(nets, "f32")
    ens.run_steps(40, 10, True, True)
    group = np.concatenate([(n.info["type"] == 1).astype(np.int32) for n in nets])
    rc, a = _call("fn_group_force_noalloc", ens, group, 3)
    assert rc == 0
    rc, b = _call("fn_group_force", ens, group, 3)
    assert rc == 0 and a.tobytes() == b.tobytes()


def test_noalloc_group_force_refuses_a_group_past_its_workspace(need_gpu):
    net = G.periodic_hs(n_fibres=3000, n_beads=50, box=(150.0, 150.0, 60.0), seed=72)    # 150,000 beads
    eng, _ = engine_for(net, "f32")
    rc, out = _call("fn_group_force_noalloc", eng, np.ones(len(net.beads), np.int32))
    assert rc == -6 and np.isnan(out).all()
    assert b"fixed workspace" in api.lib().fn_last_error(eng.h)
    rc, out = _call("fn_group_force_noalloc", eng, np.ones(len(net.beads), np.int32) * (np.arange(len(net.beads)) < 1000))
    assert rc == 0 and np.isfinite(out).all()
