"""Degenerate states the engine refuses instead of searching them: a grid
cell holding more beads than the searches take (the reference keeps 64 per
cell and drops the rest, kernel.cl:503; the engine stops at
kMaxCellBeads = 4096 and reports the run unsupported)."""
import numpy as np
import pytest

from fibrenet_amd import FibreNetError, generate as G, records as R
from helpers import engine_for

pytestmark = pytest.mark.gpu

FN_ERR_UNSUPPORTED = -5


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_collapsed_cell_is_refused(need_gpu, prec):
    net = G.periodic_hs(n_fibres=100, n_beads=50, box=(30.0, 30.0, 12.0), seed=41)   # 5,000 beads
    p = np.tile(np.array([[15.2, 15.2, 6.2]]), (len(net.beads), 1))                  # all in one cell
    R.set_vec(net.beads, "position", p)
    R.set_vec(net.beads, "image_position", p)
    eng, ids = engine_for(net, prec)
    with pytest.raises(FibreNetError) as err:
        eng.run_steps(20, net.grid_update_freq)
    assert err.value.code == FN_ERR_UNSUPPORTED
    assert "4096" in str(err.value)
    # the engine stays usable: the same network spread out again runs
    net2 = G.periodic_hs(n_fibres=100, n_beads=50, box=(30.0, 30.0, 12.0), seed=41)
    eng.write_buffer(ids["bead"], net2.beads)
    eng.run_steps(20, net2.grid_update_freq)
    assert np.isfinite(R.vec(eng.read_buffer(ids["bead"], R.BEAD), "position")).all()
