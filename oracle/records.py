"""numpy dtypes of FibreNet's host records (the wire format).

TEST INFRASTRUCTURE ONLY (oracle side).  Field order and natural alignment
restate kernel.cl:13-156 and the host headers Bead.h, BeadInfo.h,
SpringStructure.h, AngleBond.h, Pair.h, SimulationSettings.h.
"""
import numpy as np

VEC3 = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("w", "<f8")])

BEAD = np.dtype([("position", VEC3), ("image_position", VEC3), ("velocity", VEC3),
                 ("force", VEC3), ("velocity_old", VEC3), ("force_old", VEC3)])

BEAD_INFO = np.dtype([("type", "<i4"), ("tag", "<i4"), ("loc_id", "<i4"),
                      ("mass", "<f8"), ("mass_inv", "<f8"), ("lambda_dt_mass_inv", "<f8"),
                      ("half_dt_mass_inv", "<f8"), ("half_dt2_mass_inv", "<f8"),
                      ("radius", "<f8"), ("radius_sq", "<f8"), ("skin_radius", "<f8"),
                      ("skin_radius_sq", "<f8")], align=True)

SPRING = np.dtype([("type", "<i4"), ("tag", "<i4"), ("status", "<i4"), ("p1", "<i4"), ("p2", "<i4"),
                   ("rest_length", "<f8"), ("stiffness", "<f8"), ("damping", "<f8"),
                   ("c1", "<f8"), ("c2", "<f8")], align=True)

ANGLE = np.dtype([("type", "<i4"), ("tag", "<i4"), ("status", "<i4"), ("b1", "<i4"), ("b2", "<i4"),
                  ("b3", "<i4"), ("s1", "<i4"), ("s2", "<i4"), ("angle", "<f8"),
                  ("current_angle", "<f8"), ("c1", "<f8"), ("c2", "<f8"), ("c3", "<f8")], align=True)

PAIR = np.dtype([("type", "<i4"), ("p1", "<i4"), ("p2", "<i4"), ("c1", "<f8"), ("c2", "<f8"),
                 ("c3", "<f8"), ("c4", "<f8"), ("c5", "<f8"), ("c6", "<f8")], align=True)

SETTINGS = np.dtype([("types_num", "<i4"), ("bc_type", "<i4"), ("wall_contact_stiffness", "<f8"),
                     ("cutoff", "<f8"), ("cutoff_sq", "<f8"), ("box_unit_cell", "<f8"), ("dt", "<f8"),
                     ("half_dt", "<f8"), ("half_dt2", "<f8"), ("lambda_dt", "<f8"), ("gamma", "<f8"),
                     ("sigma", "<f8"), ("damping", "<f8"), ("cells_x", "<i4"), ("cells_y", "<i4"),
                     ("cells_z", "<i4"), ("cells_xy", "<i4"), ("half_box_sq", "<f8"),
                     ("grid_step", VEC3), ("box_size", VEC3), ("gravity", VEC3)], align=True)

SIZES = {"bead": 192, "bead_info": 88, "spring": 64, "angle": 72, "pair": 64, "settings": 216}
assert BEAD.itemsize == 192 and BEAD_INFO.itemsize == 88 and SPRING.itemsize == 64
assert ANGLE.itemsize == 72 and PAIR.itemsize == 64 and SETTINGS.itemsize == 216


def xyz(arr, field):
    """(n,3) float64 view-copy of a Vector3 field of a BEAD array."""
    v = arr[field]
    return np.stack([v["x"], v["y"], v["z"]], axis=1)
