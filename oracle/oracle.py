"""ctypes driver of the C oracle (fibrenet_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the product.
"""
import ctypes
import os
import subprocess

import numpy as np

from . import records as R
from .loader import grid_neighbours

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
# The same restatement run on all host threads (atomic adds, dynamic 64-item
# chunks): the all-core CPU baseline.  Its sums are order-nondeterministic.
LIB_MT_PATH = os.path.join(_HERE, "build", "liboracle_mt.so")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load(path):
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    assert L.or_sizeof(0) == R.BEAD.itemsize
    assert L.or_sizeof(1) == R.BEAD_INFO.itemsize
    assert L.or_sizeof(2) == R.SPRING.itemsize
    assert L.or_sizeof(3) == R.ANGLE.itemsize
    assert L.or_sizeof(4) == R.PAIR.itemsize
    assert L.or_sizeof(5) == R.SETTINGS.itemsize
    assert L.or_sizeof(6) == ctypes.sizeof(_State)
    return L


_LIBS = {}


def lib(threaded=False):
    path = LIB_MT_PATH if threaded else LIB_PATH
    if path not in _LIBS:
        _LIBS[path] = _load(path)
    return _LIBS[path]


class _State(ctypes.Structure):
    _fields_ = [("beads", ctypes.c_void_p), ("info", ctypes.c_void_p), ("springs", ctypes.c_void_p),
                ("angles", ctypes.c_void_p), ("pairs", ctypes.c_void_p),
                ("settings", ctypes.c_ubyte * R.SETTINGS.itemsize),
                ("n_beads", ctypes.c_int), ("n_springs", ctypes.c_int), ("n_angles", ctypes.c_int),
                ("n_cells", ctypes.c_int),
                ("cell_nh", ctypes.c_void_p), ("grid_head", ctypes.c_void_p), ("grid_tail", ctypes.c_void_p),
                ("nh_head", ctypes.c_void_p), ("nh_tail", ctypes.c_void_p), ("pair_ids", ctypes.c_void_p),
                ("cell_of_bead", ctypes.c_void_p), ("break_step", ctypes.c_void_p), ("step", ctypes.c_int)]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None and a.size else None


class Oracle:
    """One fp64 run of the reference step over numpy-held records:
    sequential (the checker), or with `threaded` every kernel's work-items
    on all host threads (OMP_NUM_THREADS) with atomic adds (the CPU
    baseline; summation order, and so the last bits, vary from run to run)."""

    def __init__(self, beads, info, springs, angles, pairs, settings, cell_nh=None, threaded=False):
        self._lib = lib(threaded)
        self.beads = np.ascontiguousarray(beads.copy())
        self.info = np.ascontiguousarray(info.copy())
        self.springs = np.ascontiguousarray(springs.copy())
        self.angles = np.ascontiguousarray(angles.copy())
        self.pairs = np.ascontiguousarray(pairs.copy())
        self.settings = np.array(settings, R.SETTINGS).reshape(())
        s = self.settings
        n = len(self.beads)
        C = int(s["cells_x"]) * int(s["cells_y"]) * int(s["cells_z"])
        if cell_nh is None:
            cell_nh = grid_neighbours(int(s["bc_type"]), int(s["cells_x"]), int(s["cells_y"]),
                                      int(s["cells_z"]))
        self.cell_nh = np.ascontiguousarray(cell_nh, np.int32)
        self.grid_head = np.zeros(C, np.int32)
        self.grid_tail = np.zeros(C * 64, np.int32)
        self.nh_head = np.zeros(n, np.int32)
        self.nh_tail = np.zeros(n * 256, np.int32)
        self.pair_ids = np.zeros(n * 256, np.int32)
        self.cell_of_bead = np.full(n, -1, np.int32)
        self.break_step = np.full(len(self.springs), -1, np.int32)
        st = _State()
        st.beads, st.info = _p(self.beads), _p(self.info)
        st.springs, st.angles, st.pairs = _p(self.springs), _p(self.angles), _p(self.pairs)
        st.settings = (ctypes.c_ubyte * R.SETTINGS.itemsize).from_buffer_copy(self.settings.tobytes())
        assert bytes(st.settings) == self.settings.tobytes()
        st.n_beads, st.n_springs, st.n_angles, st.n_cells = n, len(self.springs), len(self.angles), C
        st.cell_nh, st.grid_head, st.grid_tail = _p(self.cell_nh), _p(self.grid_head), _p(self.grid_tail)
        st.nh_head, st.nh_tail, st.pair_ids = _p(self.nh_head), _p(self.nh_tail), _p(self.pair_ids)
        st.cell_of_bead, st.break_step = _p(self.cell_of_bead), _p(self.break_step)
        st.step = 0
        self._st = st

    @classmethod
    def from_model(cls, m):
        return cls(m.beads, m.info, m.springs, m.angles, m.pairs, m.settings, m.cell_nh)

    @property
    def step(self):
        return self._st.step

    def run(self, n_steps, grid_update_freq=8):
        rc = self._lib.or_run_loop(ctypes.byref(self._st), int(n_steps), int(grid_update_freq))
        if rc:
            raise ValueError("or_run_loop failed")

    def minimize(self, steps, sub_steps, reset, grid_update_freq=8):
        rc = self._lib.or_minimize(ctypes.byref(self._st), int(steps), int(sub_steps), int(bool(reset)),
                               int(grid_update_freq))
        if rc:
            raise ValueError("or_minimize failed")

    def reset_energy(self):
        self._lib.or_reset_bead_energy(_p(self.beads), len(self.beads))

    # -- single kernels, for stage-wise checks --------------------------------------
    def advance_position(self):
        self._lib.or_advance_position(_p(self.beads), _p(self.info), _p(self.settings.reshape(1)),
                                  len(self.beads))

    def resolve_springs(self, step=0):
        self._lib.or_resolve_springs(_p(self.beads), _p(self.springs), len(self.springs), int(step),
                                 _p(self.break_step))

    def resolve_angles(self):
        self._lib.or_resolve_angles(_p(self.angles), _p(self.springs), _p(self.beads), len(self.angles))

    def rebuild(self):
        s1 = self.settings.reshape(1)
        self.grid_head[:] = 0
        self.nh_head[:] = 0
        self._lib.or_fill_grid(_p(self.beads), _p(self.grid_head), _p(self.grid_tail), _p(s1), len(self.beads),
                           _p(self.cell_of_bead))
        self._lib.or_build_contact_list(_p(self.grid_head), _p(self.grid_tail), _p(self.cell_nh),
                                    _p(self.nh_head), _p(self.nh_tail), _p(self.beads), _p(self.info),
                                    _p(self.pair_ids), _p(self.pairs), _p(s1), len(self.grid_head))

    def resolve_contacts(self):
        self._lib.or_resolve_contacts(_p(self.nh_head), _p(self.nh_tail), _p(self.beads), _p(self.info),
                                  _p(self.pair_ids), _p(self.pairs), _p(self.settings.reshape(1)),
                                  len(self.beads))

    def advance_velocity(self):
        self._lib.or_advance_velocity(_p(self.beads), _p(self.info), len(self.beads))

    # -- canonical views -------------------------------------------------------------
    def pair_set(self):
        """Sorted unique (i<j) pairs of the last contact list, as an (P,2) array.
        Raises if a pair is listed twice."""
        n = len(self.beads)
        cnt = np.minimum(self.nh_head, 256)
        owners = np.repeat(np.arange(n, dtype=np.int64), cnt)
        tail = self.nh_tail.reshape(n, 256)
        mask = np.arange(256)[None, :] < cnt[:, None]
        partners = tail[mask].astype(np.int64)
        lo = np.minimum(owners, partners)
        hi = np.maximum(owners, partners)
        key = lo * n + hi
        uk = np.unique(key)
        if len(uk) != len(key):
            raise AssertionError("a pair is listed twice")
        return np.stack([uk // n, uk % n], axis=1)
