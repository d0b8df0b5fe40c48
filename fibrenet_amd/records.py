"""numpy dtypes of the wire records of include/fibrenet_types.h.

The layouts are FibreNet's host structs (Bead.h, BeadInfo.h,
SpringStructure.h, AngleBond.h, Pair.h, SimulationSettings.h), which the C ABI
takes and returns unchanged.
"""
import numpy as np

VEC3 = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("w", "<f8")])
BEAD = np.dtype([("position", VEC3), ("image_position", VEC3), ("velocity", VEC3), ("force", VEC3),
                 ("velocity_old", VEC3), ("force_old", VEC3)])
BEAD_INFO = np.dtype([("type", "<i4"), ("tag", "<i4"), ("loc_id", "<i4"), ("mass", "<f8"), ("mass_inv", "<f8"),
                      ("lambda_dt_mass_inv", "<f8"), ("half_dt_mass_inv", "<f8"), ("half_dt2_mass_inv", "<f8"),
                      ("radius", "<f8"), ("radius_sq", "<f8"), ("skin_radius", "<f8"),
                      ("skin_radius_sq", "<f8")], align=True)
SPRING = np.dtype([("type", "<i4"), ("tag", "<i4"), ("status", "<i4"), ("p1", "<i4"), ("p2", "<i4"),
                   ("rest_length", "<f8"), ("stiffness", "<f8"), ("damping", "<f8"), ("c1", "<f8"),
                   ("c2", "<f8")], align=True)
ANGLE = np.dtype([("type", "<i4"), ("tag", "<i4"), ("status", "<i4"), ("b1", "<i4"), ("b2", "<i4"), ("b3", "<i4"),
                  ("s1", "<i4"), ("s2", "<i4"), ("angle", "<f8"), ("current_angle", "<f8"), ("c1", "<f8"),
                  ("c2", "<f8"), ("c3", "<f8")], align=True)
PAIR = np.dtype([("type", "<i4"), ("p1", "<i4"), ("p2", "<i4"), ("c1", "<f8"), ("c2", "<f8"), ("c3", "<f8"),
                 ("c4", "<f8"), ("c5", "<f8"), ("c6", "<f8")], align=True)
SETTINGS = np.dtype([("types_num", "<i4"), ("bc_type", "<i4"), ("wall_contact_stiffness", "<f8"), ("cutoff", "<f8"),
                     ("cutoff_sq", "<f8"), ("box_unit_cell", "<f8"), ("dt", "<f8"), ("half_dt", "<f8"),
                     ("half_dt2", "<f8"), ("lambda_dt", "<f8"), ("gamma", "<f8"), ("sigma", "<f8"),
                     ("damping", "<f8"), ("cells_x", "<i4"), ("cells_y", "<i4"), ("cells_z", "<i4"),
                     ("cells_xy", "<i4"), ("half_box_sq", "<f8"), ("grid_step", VEC3), ("box_size", VEC3),
                     ("gravity", VEC3)], align=True)
GRID_CELL_NH = np.dtype([("n", "<i4", (13,))])

KIND = {"bead": 1, "bead_info": 2, "spring": 3, "angle": 4, "pair": 5, "int": 6, "grid_cell": 7,
        "grid_cell_nh": 8, "particle_nh": 9}
DTYPE_OF_KIND = {1: BEAD, 2: BEAD_INFO, 3: SPRING, 4: ANGLE, 5: PAIR, 8: GRID_CELL_NH}
SIZE_OF_KIND = {1: 192, 2: 88, 3: 64, 4: 72, 5: 64, 6: 4, 7: 256, 8: 52, 9: 1024}

for _k, _d in DTYPE_OF_KIND.items():
    assert _d.itemsize == SIZE_OF_KIND[_k]
assert SETTINGS.itemsize == 216


def vec(arr, field):
    """(n, 3) float64 copy of one Vector3 field of a BEAD array."""
    v = arr[field]
    return np.stack([v["x"], v["y"], v["z"]], axis=1)


def set_vec(arr, field, xyz):
    xyz = np.asarray(xyz, np.float64)
    arr[field]["x"], arr[field]["y"], arr[field]["z"] = xyz[..., 0], xyz[..., 1], xyz[..., 2]
