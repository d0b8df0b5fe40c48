"""Restatement of FibreNet's input-file and model-file loaders.

TEST INFRASTRUCTURE ONLY.  Used by tests/ to check the product's C++ loader
(fibrenet_amd/csrc/loader.cpp) record-for-record, and to feed the C oracle.

Follows SimulationManager.cpp: ReadInputFile (1153-1171) in its order
LoadGlobalForceInfo (644-673), LoadSolverInfo (681-819), LoadSimulationBoxInfo
(2070-2325), LoadTypeAndPairInfo (827-970), LoadExportInfo (979-1091),
LoadModelData (1315-1364) = ReadBeadsData (1372-1493), ReadSpringData
(1501-1583), ReadAngleData (1591-1661), ReadRestartFile (1670-1823); then
LoadBeadModifications (189-415) and LoadSpringModifications (423-636).
Tokenising follows TextParser.h:7-38.

Numbers are parsed with the C library's strtod / strtol / strtof through
ctypes, which is what std::stod / std::stoi / std::stof call, so prefix
parsing ("1e3" -> 1 under stoi) and float rounding of stof match exactly.
Fields the reference leaves uninitialised (Vector3.w, Spring.damping, unused
Pair coefficients, wall stiffness without box_wall_rep, ...) are zero here.
"""
import ctypes
import math
import os
import re

import numpy as np

from . import records as R

_libc = ctypes.CDLL(None, use_errno=True)
_libc.strtod.restype = ctypes.c_double
_libc.strtod.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)]
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)]
_libc.strtoll.restype = ctypes.c_longlong
_libc.strtoll.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int]

_WS = re.compile(rb"[ \t\n\v\f\r]+")


class LoadError(Exception):
    """A condition under which the reference's loader returns false (or throws)."""


def _conv(fn, tok, *extra):
    b = tok.encode() if isinstance(tok, str) else tok
    buf = ctypes.create_string_buffer(b)
    end = ctypes.c_char_p()
    ctypes.set_errno(0)
    val = fn(buf, ctypes.byref(end), *extra)
    err = ctypes.get_errno()
    consumed = ctypes.cast(end, ctypes.c_void_p).value - ctypes.addressof(buf)
    if consumed == 0:
        raise LoadError("invalid numeric token %r" % tok)     # std::invalid_argument
    if err == 34:                                             # ERANGE -> std::out_of_range
        raise LoadError("numeric token out of range %r" % tok)
    return val


def stod(tok):
    return float(_conv(_libc.strtod, tok))


def stof(tok):
    return float(np.float32(_conv(_libc.strtof, tok)))


def stoi(tok):
    v = _conv(_libc.strtoll, tok, 10)
    if v < -2**31 or v > 2**31 - 1:
        raise LoadError("stoi out of range %r" % tok)
    return int(v)


def f32(x):
    """A C++ float literal / float cast (exact for the literals used)."""
    return float(np.float32(x))


def read_lines(path):
    """TextParser::ReadFileContent: std::getline on '\\n'."""
    with open(path, "rb") as fh:
        data = fh.read()
    lines = data.split(b"\n")
    if lines and lines[-1] == b"":
        lines.pop()
    return lines


def tokens_of(line):
    return [t.decode("latin-1") for t in _WS.split(line) if t]


def scan(lines, key, start):
    """TextParser::ScanText: (1-based line, tokens) of the first line at or
    after index `start` whose first token is `key`; (-1, []) if none."""
    if start < 0:
        return -1, []
    for ii in range(start, len(lines)):
        toks = tokens_of(lines[ii])
        if toks and toks[0] == key:
            return ii + 1, toks
    return -1, []


class Model:
    """Everything the loaders produce for the step."""

    def __init__(self):
        self.settings = np.zeros(1, R.SETTINGS)[0]
        self.type_id = {}
        self.pairs = None
        self.beads = None
        self.info = None
        self.springs = np.zeros(0, R.SPRING)
        self.angles = np.zeros(0, R.ANGLE)
        self.cell_nh = None
        self.grid_update_freq = 8
        self.run_steps = -1
        self.dump_frequency = -1
        self.minimize_steps = 0
        self.minimize_substeps = 0
        self.minimize_reset = False
        self.skin_radius_ratio = 0.5
        self.restart = False


def _global_force(m, lines):
    s = m.settings
    ln, t = scan(lines, "gravity", 0)
    if ln > 0:
        if len(t) != 4:
            raise LoadError("gravity")
        s["gravity"]["x"], s["gravity"]["y"], s["gravity"]["z"] = stod(t[1]), stod(t[2]), stod(t[3])
    ln, t = scan(lines, "damping", 0)
    if ln > 0:
        if len(t) != 2:
            raise LoadError("damping")
        s["damping"] = stod(t[1]) * f32(-1.0)


def _solver(m, lines):
    s = m.settings
    ln, t = scan(lines, "step", 0)
    if ln <= 0 or len(t) != 2:
        raise LoadError("step")
    dt = stod(t[1])
    if not dt > 0.0:
        raise LoadError("step value")
    s["dt"] = dt
    s["half_dt"] = dt * f32(0.5)
    s["half_dt2"] = dt * dt * f32(0.5)
    ln, t = scan(lines, "lambda", 0)
    if ln <= 0:
        raise LoadError("lambda missing (the reference returns false)")
    if len(t) != 2:
        raise LoadError("lambda")
    lam = stod(t[1])
    if not lam > 0.0:
        raise LoadError("lambda value")
    s["lambda_dt"] = lam * dt
    ln, t = scan(lines, "run", 0)
    if ln <= 0 or len(t) != 2:
        raise LoadError("run")
    m.run_steps = stoi(t[1])
    if m.run_steps < 0:
        raise LoadError("run value")
    ln, t = scan(lines, "minimize_steps", 0)
    if ln > 0:
        if len(t) != 2:
            raise LoadError("minimize_steps")
        m.minimize_steps = stoi(t[1])
        if m.minimize_steps < 0:
            raise LoadError("minimize_steps value")
    ln, t = scan(lines, "minimize_substeps", 0)
    if ln > 0:
        if len(t) != 2:
            raise LoadError("minimize_substeps")
        m.minimize_substeps = stoi(t[1])
        if m.minimize_substeps < 0 or m.minimize_substeps > m.minimize_steps:
            raise LoadError("minimize_substeps value")
    else:
        m.minimize_substeps = m.minimize_steps if m.minimize_steps > 0 else 0
    ln, t = scan(lines, "minimize_reset", 0)
    if ln > 0:
        if len(t) != 2 or t[1] not in ("yes", "no"):
            raise LoadError("minimize_reset")
        m.minimize_reset = t[1] == "yes"
    ln, t = scan(lines, "cutoff", 0)
    if ln <= 0 or len(t) != 2:
        raise LoadError("cutoff")
    c = stod(t[1])
    s["cutoff"] = c
    s["cutoff_sq"] = c * c
    if c <= 0.0:
        raise LoadError("cutoff value")


def _box(m, lines):
    s = m.settings
    ln, t = scan(lines, "box_unit_cell", 0)
    if ln <= 0 or len(t) != 2:
        raise LoadError("box_unit_cell")
    unit = stod(t[1])
    if unit <= 0.0:
        raise LoadError("box_unit_cell value")
    s["box_unit_cell"] = unit
    ln, t = scan(lines, "box_size", 0)
    if ln <= 0 or len(t) != 4:
        raise LoadError("box_size")
    L = [stod(t[1]), stod(t[2]), stod(t[3])]
    s["box_size"]["x"], s["box_size"]["y"], s["box_size"]["z"] = L
    if min(L) < 0.0:
        raise LoadError("box_size negative")
    if any(v / unit < f32(3.0) for v in L):
        raise LoadError("box smaller than three unit cells")
    ln, t = scan(lines, "box_type", 0)
    if ln <= 0 or len(t) != 2 or t[1] not in ("pbc", "finite"):
        raise LoadError("box_type")
    s["bc_type"] = 0 if t[1] == "pbc" else 1
    ln, t = scan(lines, "grid_update_freq", 0)
    if ln > 0:
        if len(t) != 2:
            raise LoadError("grid_update_freq")
        m.grid_update_freq = stoi(t[1])
        if m.grid_update_freq <= 0:
            raise LoadError("grid_update_freq value")
    ln, t = scan(lines, "box_wall_rep", 0)
    if ln > 0:
        if len(t) != 2:
            raise LoadError("box_wall_rep")
        s["wall_contact_stiffness"] = stod(t[1])
        if s["wall_contact_stiffness"] < 0.0:
            raise LoadError("box_wall_rep value")
    else:
        s["bc_type"] = 1
    if min(L) <= 0.0:
        raise LoadError("grid size")
    cx, cy, cz = (int(math.floor(v / unit)) for v in L)
    s["cells_x"], s["cells_y"], s["cells_z"] = cx, cy, cz
    s["cells_xy"] = cx * cy
    s["grid_step"]["x"] = L[0] / float(np.float32(cx))
    s["grid_step"]["y"] = L[1] / float(np.float32(cy))
    s["grid_step"]["z"] = L[2] / float(np.float32(cz))
    h = min(min(L[0], L[1]), L[2]) * f32(0.5)
    s["half_box_sq"] = h * h
    m.cell_nh = grid_neighbours(int(s["bc_type"]), cx, cy, cz)


def grid_neighbours(bc_type, cx, cy, cz):
    """SimulationManager.cpp:2193-2301 as a vectorised table (C, 13)."""
    x, y, z = np.meshgrid(np.arange(cx), np.arange(cy), np.arange(cz), indexing="ij")
    x = x.transpose(2, 1, 0).ravel()
    y = y.transpose(2, 1, 0).ravel()
    z = z.transpose(2, 1, 0).ravel()
    offs = [(dx, dy, -1) for dx in (-1, 0, 1) for dy in (-1, 0, 1)]
    offs += [(1, dy, 0) for dy in (-1, 0, 1)] + [(0, -1, 0)]
    out = np.full((cx * cy * cz, 13), -1, np.int32)
    for k, (dx, dy, dz) in enumerate(offs):
        nx, ny, nz = x + dx, y + dy, z + dz
        if bc_type == 0:
            nx = np.where(nx < 0, cx - 1, np.where(nx >= cx, 0, nx))
            ny = np.where(ny < 0, cy - 1, np.where(ny >= cy, 0, ny))
            nz = np.where(nz < 0, cz - 1, np.where(nz >= cz, 0, nz))
            out[:, k] = nx + ny * cx + nz * cx * cy
        elif bc_type == 1:
            ok = (nx >= 0) & (nx < cx) & (ny >= 0) & (ny < cy) & (nz >= 0) & (nz < cz)
            out[:, k] = np.where(ok, nx + ny * cx + nz * cx * cy, -1)
    return out


def lj_shift(sigma, eps, rc):
    """SimulationManager.cpp:899-907: the lj force magnitude/r at the cutoff."""
    r2 = rc * rc
    s2 = sigma * sigma
    fr2 = s2 / r2
    fr6 = fr2 * fr2 * fr2
    return 48.0 * eps * fr6 * (fr6 - 0.5) / r2


def _types_pairs(m, lines):
    s = m.settings
    ln = 0
    while True:
        ln, t = scan(lines, "type", ln)
        if ln == -1:
            break
        if len(t) != 2 or t[1] in m.type_id:
            raise LoadError("type definition")
        m.type_id[t[1]] = len(m.type_id)
    T = len(m.type_id)
    if T <= 0:
        raise LoadError("no types")
    s["types_num"] = T
    pairs = np.zeros(T * (T - 1) // 2 + T, R.PAIR)
    pairs["p1"] = -1
    pairs["p2"] = -1
    ln = 0
    while True:
        ln, t = scan(lines, "pair", ln)
        if ln == -1:
            break
        if len(t) <= 4:
            raise LoadError("pair incomplete")
        if t[2] not in m.type_id or t[3] not in m.type_id:
            raise LoadError("pair type name")
        a, b = m.type_id[t[2]], m.type_id[t[3]]
        if a > b:
            a, b = b, a
        idx = T * a + b - (a * (a + 1)) // 2
        p = np.zeros(1, R.PAIR)[0]
        if t[1] == "lj":
            if len(t) == 7:
                p["type"], p["p1"], p["p2"] = 1, a, b
                p["c1"], p["c2"], p["c3"] = stod(t[4]), stod(t[5]), stod(t[6])
                p["c4"] = lj_shift(p["c1"], p["c2"], p["c3"])
        elif t[1] == "hs":
            if len(t) == 5:
                p["type"], p["p1"], p["p2"] = 2, a, b
                p["c1"] = stod(t[4])
        elif t[1] == "st":
            if len(t) == 8:
                p["type"], p["p1"], p["p2"] = 3, a, b
                p["c1"] = stod(t[5])
                p["c2"] = stod(t[6])
                p["c3"] = stod(t[7]) - p["c2"]
                p["c4"] = stod(t[7]) * stod(t[7])
                p["c5"] = stod(t[4])
                p["c6"] = 2 * stod(t[6]) - stod(t[7])
        else:
            raise LoadError("unknown pair type")
        pairs[idx] = p
    if np.any((pairs["p1"] == -1) & (pairs["p2"] == -1)):
        raise LoadError("undefined type pair")
    m.pairs = pairs


def _export(m, lines):
    ln, t = scan(lines, "output_file", 0)
    if ln <= 0 or len(t) != 2:
        raise LoadError("output_file")
    ln, t = scan(lines, "output_freq", 0)
    if ln <= 0 or len(t) != 2:
        raise LoadError("output_freq")
    m.dump_frequency = stoi(t[1])
    if m.dump_frequency <= 0 or m.dump_frequency > m.run_steps:
        raise LoadError("output_freq value")
    types = []
    ln = 0
    while True:
        ln, t = scan(lines, "output_type", ln)
        if ln == -1:
            break
        if len(t) != 2:
            raise LoadError("output_type")
        if t[1] == "all":
            types = sorted(m.type_id.values())
        elif t[1] not in m.type_id:
            raise LoadError("output_type name")
        elif m.type_id[t[1]] not in types:
            types.append(m.type_id[t[1]])
    if not types:
        raise LoadError("no output_type")
    any_data = False
    ln = 0
    while True:
        ln, t = scan(lines, "output_data", ln)
        if ln == -1:
            break
        if len(t) != 2 or t[1] not in ("position", "velocity", "force", "all"):
            raise LoadError("output_data")
        any_data = True
    if not any_data:
        raise LoadError("no output_data")


def _mass_coeffs(info_row, s, mass):
    info_row["mass"] = mass
    if mass > 0.0:
        inv = f32(1.0) / mass
        info_row["mass_inv"] = inv
        info_row["lambda_dt_mass_inv"] = s["lambda_dt"] * inv
        info_row["half_dt_mass_inv"] = s["half_dt"] * inv
        info_row["half_dt2_mass_inv"] = s["half_dt2"] * inv
    else:
        for k in ("mass_inv", "lambda_dt_mass_inv", "half_dt_mass_inv", "half_dt2_mass_inv"):
            info_row[k] = 0.0


def _beads(m, lines):
    s = m.settings
    L = s["box_size"]
    rows = []
    ln = 0
    while True:
        ln, t = scan(lines, "BEAD", ln)
        if ln == -1:
            break
        if len(t) != 18:
            raise LoadError("BEAD token count at line %d" % ln)
        rows.append(t)
    n = len(rows)
    if n == 0:
        raise LoadError("no beads")
    beads = np.zeros(n, R.BEAD)
    info = np.zeros(n, R.BEAD_INFO)
    for i, t in enumerate(rows):
        b = beads[i]
        px, py, pz = stod(t[6]), stod(t[7]), stod(t[8])
        if px > L["x"] or py > L["y"] or pz > L["z"] or px < 0.0 or py < 0.0 or pz < 0.0:
            raise LoadError("bead outside the box")
        b["position"]["x"], b["position"]["y"], b["position"]["z"] = px, py, pz
        vx, vy, vz = stod(t[9]), stod(t[10]), stod(t[11])
        b["velocity"]["x"], b["velocity"]["y"], b["velocity"]["z"] = vx, vy, vz
        b["velocity_old"]["x"] = b["velocity_old"]["y"] = b["velocity_old"]["z"] = vx
        fx, fy, fz = stod(t[12]), stod(t[13]), stod(t[14])
        b["force"]["x"], b["force"]["y"], b["force"]["z"] = fx, fy, fz
        b["force_old"]["x"], b["force_old"]["y"], b["force_old"]["z"] = fx, fy, fz
        b["image_position"]["x"], b["image_position"]["y"], b["image_position"]["z"] = (
            stod(t[15]), stod(t[16]), stod(t[17]))
        if t[1] not in m.type_id:
            raise LoadError("bead type")
        r = info[i]
        r["type"] = m.type_id[t[1]]
        r["tag"] = stoi(t[2])
        r["loc_id"] = stoi(t[3])
        r["radius"] = stod(t[4])
        r["radius_sq"] = r["radius"] * r["radius"]
        r["skin_radius"] = r["radius"] + m.skin_radius_ratio
        r["skin_radius_sq"] = r["skin_radius"] * r["skin_radius"]
        _mass_coeffs(r, s, stod(t[5]))
    m.beads, m.info = beads, info


def _records(rows, dtype):
    """Zero-filled record array (padding bytes included) holding `rows`."""
    out = np.zeros(len(rows), dtype)
    for i, r in enumerate(rows):
        for f in dtype.names:
            out[f][i] = r[f]
    return out


def _springs(m, lines):
    rows = []
    ln = 0
    while True:
        ln, t = scan(lines, "SPRING", ln)
        if ln == -1:
            break
        if len(t) < 2:
            raise LoadError("SPRING line")
        if t[1] == "HR_FL" and len(t) != 9:
            raise LoadError("SPRING HR_FL token count")
        if t[1] == "HR" and len(t) < 7:
            raise LoadError("SPRING HR token count")
        sp = np.zeros(1, R.SPRING)[0]
        if t[1] == "HR":
            sp["type"] = 0
        elif t[1] == "HR_FL":
            sp["type"] = 1
            sp["c1"] = stoi(t[7])
            sp["c2"] = stoi(t[8])
        else:
            raise LoadError("SPRING type")
        sp["tag"] = stoi(t[2])
        sp["p1"] = stoi(t[3])
        sp["p2"] = stoi(t[4])
        sp["rest_length"] = stod(t[5])
        if sp["rest_length"] < 0.0:
            raise LoadError("rest length")
        sp["stiffness"] = stod(t[6])
        if sp["stiffness"] < 0.0:
            raise LoadError("stiffness")
        sp["status"] = 1
        rows.append(sp)
    m.springs = _records(rows, R.SPRING)


def _angles(m, lines):
    rows = []
    ln = 0
    while True:
        ln, t = scan(lines, "ANGLE", ln)
        if ln == -1:
            break
        if len(t) != 10:
            raise LoadError("ANGLE token count")
        a = np.zeros(1, R.ANGLE)[0]
        if t[1] == "HR":
            a["type"] = 0
        elif t[1] == "HR_FL":
            a["type"] = 1
        else:
            raise LoadError("ANGLE type")
        a["tag"] = stoi(t[2])
        a["b1"], a["b2"], a["b3"] = stoi(t[3]), stoi(t[4]), stoi(t[5])
        a["s1"], a["s2"] = stoi(t[6]), stoi(t[7])
        a["angle"] = stof(t[8])
        a["c1"] = stod(t[9])
        a["status"] = 1
        rows.append(a)
    m.angles = _records(rows, R.ANGLE)


def _restart(m, lines, base_dir):
    name = None
    ln = 0
    while True:
        ln, t = scan(lines, "restart_file", ln)
        if ln == -1:
            break
        if len(t) != 2:
            raise LoadError("restart_file")
        name = t[1]
    flags = set()
    ln = 0
    while True:
        ln, t = scan(lines, "restart_data", ln)
        if ln == -1:
            break
        if len(t) != 2 or t[1] not in ("position", "velocity", "force", "all"):
            raise LoadError("restart_data")
        flags |= {"position", "velocity", "force"} if t[1] == "all" else {t[1]}
    if not name:
        return
    path = name if os.path.isabs(name) else os.path.join(base_dir, name)
    try:
        rl = read_lines(path)
    except OSError:
        raise LoadError("restart file missing")
    if not rl:
        raise LoadError("restart file empty")
    n = len(m.beads)
    k = 0
    for line in rl:
        t = tokens_of(line)
        if len(t) != 13:
            continue
        if k >= n:
            raise LoadError("restart has more beads")
        if t[0] not in m.type_id or m.info[k]["type"] != m.type_id[t[0]]:
            raise LoadError("restart bead type")
        b = m.beads[k]
        if "position" in flags:
            b["position"]["x"], b["position"]["y"], b["position"]["z"] = stod(t[1]), stod(t[2]), stod(t[3])
            b["image_position"]["x"], b["image_position"]["y"], b["image_position"]["z"] = (
                stod(t[4]), stod(t[5]), stod(t[6]))
        if "velocity" in flags:
            for f in ("velocity", "velocity_old"):
                b[f]["x"], b[f]["y"], b[f]["z"] = stod(t[7]), stod(t[8]), stod(t[9])
        if "force" in flags:
            b["force"]["x"], b["force"]["y"], b["force"]["z"] = stod(t[10]), stod(t[11]), stod(t[12])
            b["force_old"]["x"] = b["force_old"]["y"] = b["force_old"]["z"] = 0.0
        k += 1
    if k != n:
        raise LoadError("restart bead count")
    m.restart = True


def _select(kind, key, m, n, table):
    if kind == "type":
        return table["type"][:n] == key
    if kind == "tag":
        return table["tag"][:n] == key
    sel = np.zeros(n, bool)
    if 0 <= key < n:
        sel[key] = True
    return sel


def _bead_mods(m, lines):
    s = m.settings
    n = len(m.beads)
    ln = 0
    while True:
        ln, t = scan(lines, "modif_bead", ln)
        if ln == -1:
            break
        if len(t) <= 4 or t[1] not in ("id", "tag", "type"):
            raise LoadError("modif_bead")
        if t[1] == "type":
            if t[2] not in m.type_id:
                raise LoadError("modif_bead type name")
            key = m.type_id[t[2]]
        else:
            key = stoi(t[2])
        sel = _select(t[1], key, m, n, m.info)
        if t[3] == "mass":
            if len(t) != 5:
                raise LoadError("modif_bead mass")
            mass = stod(t[4])
            if not mass >= 0.0:
                raise LoadError("modif_bead mass value")
            for i in np.nonzero(sel)[0]:
                _mass_coeffs(m.info[i], s, mass)
        elif t[3] == "velocity":
            if len(t) != 7:
                raise LoadError("modif_bead velocity")
            v = (stod(t[4]), stod(t[5]), stod(t[6]))
            for f in ("velocity", "velocity_old"):
                for c, val in zip("xyz", v):
                    m.beads[f][c][sel] = val
                m.beads[f]["w"][sel] = 0.0
        elif t[3] == "skin_radius":
            if len(t) != 5:
                raise LoadError("modif_bead skin_radius")
            sk = stod(t[4])
            m.info["skin_radius"][sel] = sk
            m.info["skin_radius_sq"][sel] = sk * sk
        else:
            raise LoadError("modif_bead property")


def _spring_mods(m, lines):
    if len(m.springs) == 0:
        raise LoadError("spring buffer not initialized (the reference requires springs)")
    S = len(m.springs)
    ln = 0
    while True:
        ln, t = scan(lines, "modif_spring", ln)
        if ln == -1:
            break
        if len(t) <= 4 or t[1] not in ("id", "tag"):
            raise LoadError("modif_spring")
        key = stoi(t[2])
        sel = _select(t[1], key, m, S, m.springs)
        prop = {"c1": "c1", "c2": "c2", "rest_length": "rest_length", "stiffness": "stiffness",
                "damping": "damping"}.get(t[3])
        if prop is None:
            raise LoadError("modif_spring property")
        if len(t) != 5:
            if prop in ("stiffness", "damping"):
                continue        # the reference logs and carries on (SimulationManager.cpp:586,616)
            raise LoadError("modif_spring token count")
        v = stod(t[4])
        if not v >= 0.0:
            raise LoadError("modif_spring value")
        m.springs[prop][sel] = v


def load_input(path):
    """SimulationManager::ReadInputFile for `path`; model and restart paths are
    resolved relative to the input file's directory when not absolute."""
    lines = read_lines(path)
    base = os.path.dirname(os.path.abspath(path))
    m = Model()
    _global_force(m, lines)
    _solver(m, lines)
    _box(m, lines)
    _types_pairs(m, lines)
    _export(m, lines)
    ln, t = scan(lines, "model_file", 0)
    if ln <= 0 or len(t) != 2:
        raise LoadError("model_file")
    mpath = t[1] if os.path.isabs(t[1]) else os.path.join(base, t[1])
    try:
        mlines = read_lines(mpath)
    except OSError:
        raise LoadError("model file missing")
    if not mlines:
        raise LoadError("model file empty")
    _beads(m, mlines)
    _springs(m, mlines)
    _angles(m, mlines)
    _restart(m, lines, base)
    _bead_mods(m, lines)
    _spring_mods(m, lines)
    return m
