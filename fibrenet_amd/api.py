"""ctypes bindings of libfibrenet_amd.so (include/fibrenet.h).

The library is the product: every step runs in its HIP kernels.  Loading
fails loudly when the library has not been built; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

from . import records as R

_HERE = os.path.dirname(os.path.abspath(__file__))
# The in-tree build; FIBRENET_AMD_LIB names another build of the same library.
LIB_PATH = os.environ.get("FIBRENET_AMD_LIB") or os.path.join(_HERE, "libfibrenet_amd.so")

FN_OK = 0
PREC = {"f64": 0, "f32": 1}
ARG_BUFFER, ARG_INT, ARG_SETTINGS = 1, 2, 3
KERNEL_NAMES = ["AdvancePosition", "ResetGrid", "ResetContactList", "FillGrid", "BuildContactList",
                "ResolveContacts", "AdvanceVelocity", "ResetBeadEnergy", "ResolveSprings", "ResolveAngles"]

# Every symbol include/fibrenet.h declares: (name, restype, argtypes).
_c = ctypes
_VP, _CP, _I, _LL = _c.c_void_p, _c.c_char_p, _c.c_int, _c.c_longlong
SIGNATURES = [
    ("fn_version", _CP, []),
    ("fn_create", _I, [_I, _I, _c.POINTER(_VP)]),
    ("fn_destroy", _I, [_VP]),
    ("fn_last_error", _CP, [_VP]),
    ("fn_set_log_callback", _I, [_VP, _VP, _VP]),
    ("fn_register", _I, [_VP, _c.POINTER(_CP), _I]),
    ("fn_buffer_create", _I, [_VP, _I, _I, _I, _VP, _c.POINTER(_I)]),
    ("fn_buffer_write", _I, [_VP, _I, _VP, _I, _I]),
    ("fn_buffer_read", _I, [_VP, _I, _VP]),
    ("fn_buffer_info", _I, [_VP, _I, _c.POINTER(_I), _c.POINTER(_I), _c.POINTER(_I), _c.POINTER(_I)]),
    ("fn_set_arg", _I, [_VP, _CP, _I, _I, _VP, _I]),
    ("fn_set_range", _I, [_VP, _CP, _I]),
    ("fn_launch", _I, [_VP, _CP]),
    ("fn_finish", _I, [_VP]),
    ("fn_run_steps", _I, [_VP, _I, _I, _I, _I]),
    ("fn_export_cells", _I, [_VP, _VP]),
    ("fn_export_pairs", _I, [_VP, _VP, _LL, _c.POINTER(_LL)]),
    ("fn_export_breaks", _I, [_VP, _VP]),
    ("fn_get_stats", _I, [_VP, _VP]),
    ("fn_set_option", _I, [_VP, _CP, _I]),
    ("fn_get_option", _I, [_VP, _CP, _c.POINTER(_I)]),
    ("fn_set_timing", _I, [_VP, _I]),
    ("fn_get_timing", _I, [_VP, _VP]),
    ("fn_copy_probe", _I, [_VP, _c.c_longlong, _VP]),
    ("fn_batch_create", _I, [_I, _I, _I, _VP, _VP]),
    ("fn_batch_read", _I, [_VP, _I, _I, _VP]),
    ("fn_batch_load", _I, [_VP, _I, _VP]),
    ("fn_autotune", _I, [_VP, _I, _I, _I, _I, _VP, _I, _c.POINTER(_I)]),
    ("fn_adjoint_forward", _I, [_VP, _VP, _VP, _VP, _I, _VP]),
    ("fn_adjoint_backward", _I, [_VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    ("fn_model_load", _I, [_CP, _c.POINTER(_VP)]),
    ("fn_model_free", _I, [_VP]),
    ("fn_model_error", _CP, []),
    ("fn_model_counts", _I, [_VP] + [_c.POINTER(_I)] * 5),
    ("fn_model_params", _I, [_VP] + [_c.POINTER(_I)] * 6),
    ("fn_model_copy", _I, [_VP, _I, _VP]),
    ("fn_sim_create", _I, [_I, _I, _CP, _c.POINTER(_VP)]),
    ("fn_sim_destroy", _I, [_VP]),
    ("fn_sim_engine", _VP, [_VP]),
    ("fn_sim_buffer_id", _I, [_VP, _I]),
    ("fn_sim_run_loop", _I, [_VP, _I, _I]),
    ("fn_sim_minimize", _I, [_VP, _I]),
    ("fn_sim_run", _I, [_VP, _I]),
    ("fn_sim_dump", _I, [_VP, _I]),
    ("fn_sim_write_restart", _I, [_VP, _CP, _I]),
    ("fn_sim_set_log_callback", _I, [_VP, _VP, _VP]),
    ("fn_sim_error", _CP, [_VP]),
    ("fn_sim_output_file", _CP, [_VP]),
    ("fn_model_write_restart", _I, [_VP, _VP, _I, _CP]),
    ("fn_model_write_xyz", _I, [_VP, _VP, _I, _I, _CP]),
    ("fn_checkpoint_save", _I, [_VP, _CP]),
    ("fn_checkpoint_load", _I, [_VP, _CP]),
    ("fn_group_force", _I, [_VP, _VP, _VP]),
    ("fn_group_force_noalloc", _I, [_VP, _VP, _VP]),
    ("fn_gen_fibres", _I, [_c.c_ulonglong, _I, _I, _c.c_double, _c.c_double, _c.c_double, _I, _c.c_double,
                           _c.c_double, _LL, _VP, _c.POINTER(_LL)]),
]


class FnStats(ctypes.Structure):
    _fields_ = [("steps", _LL), ("rebuilds", _LL), ("list_entries", _LL), ("list_capacity", _LL),
                ("candidate_entries", _LL), ("candidate_builds", _I),
                ("hashed_beads", _I), ("max_cell_occupancy", _I), ("cells_over_capacity", _I),
                ("max_list_length", _I), ("broken_springs", _I), ("img_aliases_pos", _I)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FnTiming(ctypes.Structure):
    _fields_ = [("steps", _I), ("rebuilds", _I), ("step_ms", _c.c_double), ("rebuild_ms", _c.c_double),
                ("spring_ms", _c.c_double), ("angle_ms", _c.c_double), ("bead_ms", _c.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FnNetwork(ctypes.Structure):
    """fn_network: one network of a batched ensemble (host records + settings)."""
    _fields_ = [("beads", _VP), ("info", _VP), ("n_beads", _I), ("springs", _VP), ("n_springs", _I),
                ("angles", _VP), ("n_angles", _I), ("pairs", _VP), ("n_pairs", _I), ("settings", _VP)]


class FnTuneTrial(ctypes.Structure):
    """fn_tune_trial: one timed chunk of the autotuner."""
    _fields_ = [("knob", _I), ("value", _I), ("ms_per_step", _c.c_double)]


TUNE_KNOBS = ("chunk", "chain", "staged", "verlet", "block", "skin")


class FnAdjointSpec(ctypes.Structure):
    """fn_adjoint_spec: the run an adjoint forward records."""
    _fields_ = [("n_steps", _I), ("grid_update_freq", _I), ("run_springs", _I), ("run_angles", _I),
                ("checkpoint_every", _I)]


class FibreNetError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("fibrenet error %d: %s" % (code, msg))
        self.code = code


_LIB = None


def lib():
    """The loaded library; raises if it has not been built (no fallback)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libfibrenet_amd.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


XYZ_FIELDS = {"position": 0, "velocity": 1, "force": 2}


class Model:
    """Host records parsed from a FibreNet input file (fn_model_*)."""

    def __init__(self, input_path):
        L = lib()
        h = ctypes.c_void_p()
        rc = L.fn_model_load(os.fsencode(input_path), ctypes.byref(h))
        if rc != FN_OK:
            raise FibreNetError(rc, L.fn_model_error().decode())
        self._h = h
        try:
            n = [ctypes.c_int() for _ in range(5)]
            L.fn_model_counts(h, *[ctypes.byref(x) for x in n])
            self.n_beads, self.n_springs, self.n_angles, self.n_pairs, self.n_cells = [x.value for x in n]
            p = [ctypes.c_int() for _ in range(6)]
            L.fn_model_params(h, *[ctypes.byref(x) for x in p])
            (self.run_steps, self.dump_frequency, self.grid_update_freq, self.minimize_steps,
             self.minimize_substeps, self.minimize_reset) = [x.value for x in p]
            self.minimize_reset = bool(self.minimize_reset)
            self.settings = np.zeros((), R.SETTINGS)
            L.fn_model_copy(h, 0, _ptr(self.settings))
            self.beads = self._copy(h, 1, R.BEAD, self.n_beads)
            self.info = self._copy(h, 2, R.BEAD_INFO, self.n_beads)
            self.springs = self._copy(h, 3, R.SPRING, self.n_springs)
            self.angles = self._copy(h, 4, R.ANGLE, self.n_angles)
            self.pairs = self._copy(h, 5, R.PAIR, self.n_pairs)
            self.cell_nh = self._copy(h, 8, R.GRID_CELL_NH, self.n_cells)["n"]
        except Exception:
            self.close()
            raise

    def close(self):
        if getattr(self, "_h", None):
            lib().fn_model_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def write_restart(self, beads, progress, path):
        """DumpRestartData's text of `beads` (this model's bead count of
        records) into `path` (fn_model_write_restart)."""
        b = np.ascontiguousarray(beads, R.BEAD)
        assert len(b) == self.n_beads
        rc = lib().fn_model_write_restart(self._h, _ptr(b), int(progress), os.fsencode(path))
        if rc != FN_OK:
            raise FibreNetError(rc, "fn_model_write_restart failed")

    def write_xyz(self, beads, progress, field, path):
        """One DumpData frame of `field` ("position", "velocity", "force")
        appended to `path` (fn_model_write_xyz)."""
        b = np.ascontiguousarray(beads, R.BEAD)
        assert len(b) == self.n_beads
        rc = lib().fn_model_write_xyz(self._h, _ptr(b), int(progress), XYZ_FIELDS[field], os.fsencode(path))
        if rc != FN_OK:
            raise FibreNetError(rc, "fn_model_write_xyz failed")

    @staticmethod
    def _copy(h, kind, dtype, n):
        a = np.zeros(n, dtype)
        if n:
            lib().fn_model_copy(h, kind, _ptr(a))
        return a


# What fn_adjoint_backward can return, in its argument order.
ADJOINT_OUTPUTS = ("spring", "angle", "pair", "position", "velocity_old")
# The force-scale derivatives alone (what a stiffness fit needs).
SCALE_OUTPUTS = ("spring", "angle", "pair")


class Engine:
    """One engine instance (fn_engine): buffers by id, kernels by name."""

    def __init__(self, precision="f32", device=0, handle=None):
        L = lib()
        self._owned = handle is None
        if handle is None:
            h = ctypes.c_void_p()
            rc = L.fn_create(int(device), PREC[precision], ctypes.byref(h))
            if rc != FN_OK:
                raise FibreNetError(rc, "fn_create failed (no HIP device?)")
            handle = h.value
        self.h = ctypes.c_void_p(handle)
        self.precision = precision

    n_networks = 1          # an Ensemble holds several
    batched = False

    def close(self):
        if self._owned and self.h:
            lib().fn_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != FN_OK:
            raise FibreNetError(rc, lib().fn_last_error(self.h).decode())

    # -- OpenclManager surface -----------------------------------------------------------
    def register(self, names=KERNEL_NAMES):
        arr = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
        self._chk(lib().fn_register(self.h, arr, len(names)))

    def create_buffer(self, kind, data=None, n=None, elem_size=None):
        kind = R.KIND[kind] if isinstance(kind, str) else kind
        es = elem_size or R.SIZE_OF_KIND[kind]
        if data is not None:
            data = np.ascontiguousarray(data)
            n = len(data)
        out = ctypes.c_int(-1)
        self._chk(lib().fn_buffer_create(self.h, kind, es, int(n), _ptr(data) if data is not None else None,
                                         ctypes.byref(out)))
        return out.value

    def buffer_info(self, bid):
        v = [ctypes.c_int() for _ in range(4)]
        self._chk(lib().fn_buffer_info(self.h, bid, *[ctypes.byref(x) for x in v]))
        return dict(zip(("size", "elements", "length", "struct_size"), [x.value for x in v]))

    def write_buffer(self, bid, data):
        """OpenclManager::WriteBuffer: the records padded to the buffer's
        length (unpadded records of the buffer's element count are padded
        here with zero records, as CreateBuffer pads)."""
        info = self.buffer_info(bid)
        data = np.ascontiguousarray(data)
        if len(data) == info["elements"] and info["length"] > len(data):
            padded = np.zeros(info["length"], data.dtype)
            padded[:len(data)] = data
            data = padded
        self._chk(lib().fn_buffer_write(self.h, bid, _ptr(data), len(data), data.dtype.itemsize))
        return info

    def read_buffer(self, bid, dtype):
        info = self.buffer_info(bid)
        out = np.zeros(info["length"], dtype)
        self._chk(lib().fn_buffer_read(self.h, bid, _ptr(out)))
        return out[:info["elements"]]

    def set_arg(self, kernel, idx, value):
        if isinstance(value, np.ndarray) and value.dtype == R.SETTINGS:
            v = np.ascontiguousarray(value.reshape(()))
            self._chk(lib().fn_set_arg(self.h, kernel.encode(), idx, ARG_SETTINGS, _ptr(v), R.SETTINGS.itemsize))
        elif isinstance(value, tuple) and value[0] == "buffer":
            v = ctypes.c_int(value[1])
            self._chk(lib().fn_set_arg(self.h, kernel.encode(), idx, ARG_BUFFER, ctypes.byref(v), 4))
        else:
            v = ctypes.c_int(int(value))
            self._chk(lib().fn_set_arg(self.h, kernel.encode(), idx, ARG_INT, ctypes.byref(v), 4))

    def set_range(self, kernel, n):
        self._chk(lib().fn_set_range(self.h, kernel.encode(), int(n)))

    def launch(self, kernel):
        self._chk(lib().fn_launch(self.h, kernel.encode()))

    def finish(self):
        self._chk(lib().fn_finish(self.h))

    def run_steps(self, n, grid_update_freq=8, springs=True, angles=True):
        self._chk(lib().fn_run_steps(self.h, int(n), int(grid_update_freq), int(springs), int(angles)))

    # -- exports ----------------------------------------------------------------------------
    def export_cells(self, n_beads):
        out = np.zeros(n_beads, np.int32)
        self._chk(lib().fn_export_cells(self.h, _ptr(out)))
        return out

    def export_pairs(self):
        n = ctypes.c_longlong()
        self._chk(lib().fn_export_pairs(self.h, None, 0, ctypes.byref(n)))
        out = np.zeros((max(n.value, 1), 2), np.int32)
        self._chk(lib().fn_export_pairs(self.h, _ptr(out), n.value, ctypes.byref(n)))
        return out[:n.value]

    def export_breaks(self, n_springs):
        out = np.zeros(max(n_springs, 1), np.int32)
        self._chk(lib().fn_export_breaks(self.h, _ptr(out)))
        return out[:n_springs]

    def stats(self):
        s = FnStats()
        self._chk(lib().fn_get_stats(self.h, ctypes.byref(s)))
        return s.as_dict()

    def set_option(self, name, value):
        self._chk(lib().fn_set_option(self.h, name.encode(), int(value)))

    def get_option(self, name):
        v = ctypes.c_int()
        self._chk(lib().fn_get_option(self.h, name.encode(), ctypes.byref(v)))
        return v.value

    def copy_probe(self, nbytes):
        """Streaming device copy of nbytes; returns its duration in ms."""
        ms = ctypes.c_double()
        self._chk(lib().fn_copy_probe(self.h, int(nbytes), ctypes.byref(ms)))
        return ms.value

    def autotune(self, n, grid_update_freq=8, springs=True, angles=True):
        """Runs n steps as run_steps does while timing the result-neutral
        settings (fn_autotune); returns the timed trials and the settings
        chosen (which stay set)."""
        arr = (FnTuneTrial * 64)()
        cnt = ctypes.c_int()
        self._chk(lib().fn_autotune(self.h, int(n), int(grid_update_freq), int(bool(springs)), int(bool(angles)),
                                    arr, 64, ctypes.byref(cnt)))
        trials = [{"knob": TUNE_KNOBS[t.knob], "value": t.value, "ms_per_step": t.ms_per_step}
                  for t in arr[:min(cnt.value, 64)]]
        chosen = {k: self.get_option(k) for k in TUNE_KNOBS}
        return trials, chosen

    # -- adjoint ------------------------------------------------------------------------------
    def adjoint_forward(self, n, group, sample_steps, grid_update_freq=8, springs=True, angles=True,
                        checkpoint_every=0):
        """Runs n steps as run_steps does, recording checkpoints; returns the
        x-force summed over the beads flagged in `group` at each of
        `sample_steps` (strictly increasing, in [0, n))."""
        g = np.ascontiguousarray(group, np.int32)
        st = np.ascontiguousarray(sample_steps, np.int32)
        nn = self.n_networks
        out = np.zeros(max(len(st) * nn, 1), np.float64)
        spec = FnAdjointSpec(int(n), int(grid_update_freq), int(bool(springs)), int(bool(angles)),
                             int(checkpoint_every))
        self._chk(lib().fn_adjoint_forward(self.h, ctypes.byref(spec), _ptr(g), _ptr(st), len(st), _ptr(out)))
        self._adj_beads = len(g)
        out = out[:len(st) * nn]
        return out.reshape(nn, len(st)) if self.batched else out

    def adjoint_backward(self, dloss, n_springs, n_angles, n_pairs, outputs=ADJOINT_OUTPUTS):
        """Derivatives of a loss with dL/d(sample) = dloss through the last
        adjoint_forward run: per-bond and per-pair-type force-scale
        derivatives ("spring", "angle", "pair") and those of the initial
        positions and velocity_old ("position", "velocity_old").  Only the
        derivatives named in `outputs` are copied to the host and returned."""
        bad = set(outputs) - set(ADJOINT_OUTPUTS)
        if bad:
            raise ValueError("unknown adjoint outputs %s" % sorted(bad))
        d = np.ascontiguousarray(dloss, np.float64).ravel()
        n = getattr(self, "_adj_beads", 0)
        shape = {"spring": (max(n_springs, 1),), "angle": (max(n_angles, 1),), "pair": (max(n_pairs, 1),),
                 "position": (n, 3), "velocity_old": (n, 3)}
        out = {k: np.zeros(shape[k]) for k in ADJOINT_OUTPUTS if k in outputs}
        ptrs = [_ptr(out[k]) if k in out else None for k in ADJOINT_OUTPUTS]
        self._chk(lib().fn_adjoint_backward(self.h, _ptr(d) if d.size else None, *ptrs))
        for k, m in (("spring", n_springs), ("angle", n_angles), ("pair", n_pairs)):
            if k in out:
                out[k] = out[k][:m]
        return out

    def set_timing(self, on):
        self._chk(lib().fn_set_timing(self.h, int(bool(on))))

    def checkpoint_save(self, path):
        """Bit-exact binary checkpoint of the dynamic state (fn_checkpoint_save)."""
        self._chk(lib().fn_checkpoint_save(self.h, os.fsencode(path)))

    def checkpoint_load(self, path):
        """Restores a checkpoint of the same network (fn_checkpoint_load)."""
        self._chk(lib().fn_checkpoint_load(self.h, os.fsencode(path)))

    def group_force(self, group):
        """x-force of the current step summed over the flagged beads, per
        network (fn_group_force)."""
        g = np.ascontiguousarray(group, np.int32)
        out = np.zeros(self.n_networks, np.float64)
        self._chk(lib().fn_group_force(self.h, _ptr(g), _ptr(out)))
        return out if self.batched else float(out[0])

    def timing(self):
        t = FnTiming()
        self._chk(lib().fn_get_timing(self.h, ctypes.byref(t)))
        return t.as_dict()


def bind_standard(eng, net):
    """Uploads a network (generate.Network or Model) and binds every kernel
    argument as SimulationManager::ReadKernelData does
    (SimulationManager.cpp:1229-1304).  Returns the buffer ids."""
    s = np.array(net.settings, R.SETTINGS).reshape(())
    C = int(s["cells_x"]) * int(s["cells_y"]) * int(s["cells_z"])
    N = len(net.beads)
    ids = {}
    eng.register()
    ids["head"] = eng.create_buffer("int", n=C)
    ids["tail"] = eng.create_buffer("grid_cell", n=C)
    cnh = np.zeros(len(net.cell_nh), R.GRID_CELL_NH)
    cnh["n"] = net.cell_nh
    ids["cellnh"] = eng.create_buffer("grid_cell_nh", cnh)
    ids["pair"] = eng.create_buffer("pair", net.pairs)
    ids["bead"] = eng.create_buffer("bead", net.beads)
    ids["info"] = eng.create_buffer("bead_info", net.info)
    ids["nh_head"] = eng.create_buffer("int", n=N)
    ids["nh_tail"] = eng.create_buffer("particle_nh", n=N)
    ids["coeff"] = eng.create_buffer("particle_nh", n=N)
    if len(net.springs):
        ids["spring"] = eng.create_buffer("spring", net.springs)
    if len(net.angles):
        ids["angle"] = eng.create_buffer("angle", net.angles)
    B = lambda k: ("buffer", ids[k])
    L = lambda k: eng.buffer_info(ids[k])["length"]
    eng.set_arg("AdvancePosition", 0, B("bead")); eng.set_arg("AdvancePosition", 1, B("info"))
    eng.set_arg("AdvancePosition", 2, s); eng.set_arg("AdvancePosition", 3, N)
    eng.set_range("AdvancePosition", L("bead"))
    if "spring" in ids:
        eng.set_arg("ResolveSprings", 0, B("bead")); eng.set_arg("ResolveSprings", 1, B("spring"))
        eng.set_arg("ResolveSprings", 2, len(net.springs)); eng.set_range("ResolveSprings", L("spring"))
    if "angle" in ids:
        eng.set_arg("ResolveAngles", 0, B("angle")); eng.set_arg("ResolveAngles", 1, B("spring"))
        eng.set_arg("ResolveAngles", 2, B("bead")); eng.set_arg("ResolveAngles", 3, len(net.angles))
        eng.set_range("ResolveAngles", L("angle"))
    eng.set_arg("ResetGrid", 0, B("head")); eng.set_arg("ResetGrid", 1, C); eng.set_range("ResetGrid", L("head"))
    eng.set_arg("ResetContactList", 0, B("nh_head")); eng.set_arg("ResetContactList", 1, N)
    eng.set_range("ResetContactList", L("nh_head"))
    for i, v in enumerate([B("bead"), B("head"), B("tail"), s, N]):
        eng.set_arg("FillGrid", i, v)
    eng.set_range("FillGrid", L("bead"))
    for i, v in enumerate([B("head"), B("tail"), B("cellnh"), B("nh_head"), B("nh_tail"), B("bead"), B("info"),
                           B("coeff"), B("pair"), s, C]):
        eng.set_arg("BuildContactList", i, v)
    eng.set_range("BuildContactList", L("head"))
    for i, v in enumerate([B("nh_head"), B("nh_tail"), B("bead"), B("info"), B("coeff"), B("pair"), s, N]):
        eng.set_arg("ResolveContacts", i, v)
    eng.set_range("ResolveContacts", L("nh_head"))
    for i, v in enumerate([B("bead"), B("info"), N, L("bead")]):
        eng.set_arg("AdvanceVelocity", i, v)
    eng.set_range("AdvanceVelocity", L("bead"))
    eng.set_arg("ResetBeadEnergy", 0, B("bead")); eng.set_arg("ResetBeadEnergy", 1, N)
    eng.set_range("ResetBeadEnergy", L("bead"))
    return ids


def run_loop_granular(eng, n_steps, grid_update_freq, springs=True, angles=True):
    """SimulationManager::RunSimulationLoop through per-kernel launches with a
    finish every step (SimulationManager.cpp:151-181)."""
    cnt = 0
    for _ in range(n_steps):
        if cnt == grid_update_freq:
            cnt = 0
        eng.launch("AdvancePosition")
        if springs:
            eng.launch("ResolveSprings")
        if angles:
            eng.launch("ResolveAngles")
        if cnt == 0:
            for k in ("ResetGrid", "ResetContactList", "FillGrid", "BuildContactList"):
                eng.launch(k)
        eng.launch("ResolveContacts")
        eng.launch("AdvanceVelocity")
        eng.finish()
        cnt += 1


class Simulation:
    """fn_sim: SimulationManager over the HIP engine, loaded from an input file."""

    def __init__(self, input_path, precision="f32", device=0):
        L = lib()
        h = ctypes.c_void_p()
        rc = L.fn_sim_create(int(device), PREC[precision], os.fsencode(input_path), ctypes.byref(h))
        if rc != FN_OK:
            raise FibreNetError(rc, L.fn_model_error().decode())
        self.h = h
        self.engine = Engine(precision, handle=L.fn_sim_engine(h))

    def close(self):
        if self.h:
            lib().fn_sim_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def buffer_id(self, kind):
        return lib().fn_sim_buffer_id(self.h, R.KIND[kind])

    def run_loop(self, n, fast=True):
        self._chk(lib().fn_sim_run_loop(self.h, int(n), int(bool(fast))))

    def minimize(self, fast=True):
        self._chk(lib().fn_sim_minimize(self.h, int(bool(fast))))

    def _chk(self, rc):
        if rc != FN_OK:
            raise FibreNetError(rc, lib().fn_sim_error(self.h).decode())

    def run(self, fast=True):
        """SimulationManager::Run: minimisation, output_freq chunks with
        their XYZ frames and restart stages, the final restart frame."""
        self._chk(lib().fn_sim_run(self.h, int(bool(fast))))

    def dump(self, progress):
        self._chk(lib().fn_sim_dump(self.h, int(progress)))

    def write_restart(self, progress, path=None):
        self._chk(lib().fn_sim_write_restart(self.h, os.fsencode(path) if path else None, int(progress)))

    @property
    def output_file(self):
        return lib().fn_sim_output_file(self.h).decode()

    def set_log(self, fn):
        """Routes the simulation's (and its engine's) messages to fn(str);
        None restores stderr."""
        if fn is None:
            self._log_cb = None
            self._chk(lib().fn_sim_set_log_callback(self.h, None, None))
            return
        cb = ctypes.CFUNCTYPE(None, ctypes.c_char_p, ctypes.c_void_p)(lambda m, u: fn(m.decode()))
        self._log_cb = cb                      # keeps the trampoline alive
        self._chk(lib().fn_sim_set_log_callback(self.h, ctypes.cast(cb, ctypes.c_void_p), None))

    def beads(self):
        return self.engine.read_buffer(self.buffer_id("bead"), R.BEAD)

    def springs(self):
        return self.engine.read_buffer(self.buffer_id("spring"), R.SPRING)


class Ensemble(Engine):
    """A batched ensemble (fn_batch_create): independent networks - each a
    generate.Network or Model with its own records and settings - advanced
    together by run_steps, one launch per step for all of them."""

    _KINDS = {"bead": (R.KIND["bead"], R.BEAD), "spring": (R.KIND["spring"], R.SPRING),
              "angle": (R.KIND["angle"], R.ANGLE)}

    batched = True

    @staticmethod
    def _networks(nets):
        keep, counts = [], []
        arr = (FnNetwork * len(nets))()
        for k, net in enumerate(nets):
            b = np.ascontiguousarray(net.beads, R.BEAD)
            q = np.ascontiguousarray(net.info, R.BEAD_INFO)
            sp = np.ascontiguousarray(net.springs, R.SPRING)
            an = np.ascontiguousarray(net.angles, R.ANGLE)
            pr = np.ascontiguousarray(net.pairs, R.PAIR)
            st = np.array(net.settings, R.SETTINGS).reshape(())
            keep += [b, q, sp, an, pr, st]
            ptr = lambda a: a.ctypes.data if a.size else None
            arr[k] = FnNetwork(ptr(b), ptr(q), len(b), ptr(sp), len(sp), ptr(an), len(an), ptr(pr), len(pr),
                               st.ctypes.data)
            counts.append({"bead": len(b), "spring": len(sp), "angle": len(an), "pair": len(pr)})
        return arr, keep, counts

    def __init__(self, nets, precision="f32", device=0):
        arr, self._keep, self.counts = self._networks(nets)
        self.n_networks = len(nets)
        h = ctypes.c_void_p()
        L = lib()
        rc = L.fn_batch_create(int(device), PREC[precision], len(nets), arr, ctypes.byref(h))
        if rc != FN_OK:
            msg = L.fn_last_error(h).decode() if h.value else "fn_batch_create failed (no HIP device?)"
            if h.value:
                L.fn_destroy(h)
            raise FibreNetError(rc, msg)
        super().__init__(precision, device, handle=h.value)
        self._owned = True

    def reload(self, nets):
        """New records for every network, of the loaded sizes (fn_batch_load)."""
        arr, keep, _ = self._networks(nets)
        self._chk(lib().fn_batch_load(self.h, len(nets), arr))
        del keep

    def offsets(self, kind):
        """First global index of each network's records of `kind`, and the total."""
        return np.concatenate([[0], np.cumsum([c[kind] for c in self.counts])]).astype(np.int64)

    def read(self, k, kind="bead"):
        code, rec = self._KINDS[kind]
        out = np.zeros(self.counts[k][kind], rec)
        if len(out):
            self._chk(lib().fn_batch_read(self.h, int(k), code, out.ctypes.data))
        return out

