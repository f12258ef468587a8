"""Field numbering of the C-ABI (mhd_field in include/beta1_mhd.h).

Names follow the reference's registered fields (Z4_Lunar/main.cpp:61-121 and
src/3_field/4_MPCNS_Geometry.cpp:6-73).  Locations follow StaggerLocation
(include/3_field/Field_Type.h:4-14).
"""
CELL, NODE, FACE_XI, FACE_ET, FACE_ZE, EDGE_XI, EDGE_ET, EDGE_ZE = range(8)

# (name, location, components)
FIELDS = [
    ("U_H", CELL, 5), ("PV_H", CELL, 5), ("U_plus", CELL, 3),
    ("B_xi", FACE_XI, 1), ("B_eta", FACE_ET, 1), ("B_zeta", FACE_ZE, 1),
    ("Badd_xi", FACE_XI, 1), ("Badd_eta", FACE_ET, 1), ("Badd_zeta", FACE_ZE, 1),
    ("B_cell", CELL, 3), ("Bind_cell", CELL, 3),
    ("J_xi", EDGE_XI, 1), ("J_eta", EDGE_ET, 1), ("J_zeta", EDGE_ZE, 1), ("J_cell", CELL, 3),
    ("E_xi", EDGE_XI, 1), ("E_eta", EDGE_ET, 1), ("E_zeta", EDGE_ZE, 1),
    ("Eface_xi", FACE_XI, 3), ("Eface_eta", FACE_ET, 3), ("Eface_zeta", FACE_ZE, 3),
    ("F_xi", FACE_XI, 5), ("F_eta", FACE_ET, 5), ("F_zeta", FACE_ZE, 5),
    ("RHS_H", CELL, 5), ("RHS_B_xi", FACE_XI, 1), ("RHS_B_eta", FACE_ET, 1), ("RHS_B_zeta", FACE_ZE, 1),
    ("divB", CELL, 1), ("Jac", CELL, 1), ("JDxi", FACE_XI, 3), ("JDet", FACE_ET, 3), ("JDze", FACE_ZE, 3),
    ("beta_xi", FACE_XI, 1), ("beta_eta", FACE_ET, 1), ("beta_zeta", FACE_ZE, 1),
    ("alpha_xi", EDGE_XI, 1), ("alpha_eta", EDGE_ET, 1), ("alpha_zeta", EDGE_ZE, 1),
    ("Bcell_from_Bface_w", CELL, 18), ("Jcell_from_Jedge_w", CELL, 36),
    # two-fluid (Mercury) cases
    ("U_Na", CELL, 5), ("PV_Na", CELL, 5), ("RHS_Na", CELL, 5), ("Photo_rate", CELL, 1), ("drag_rate", CELL, 1),
    # oracle-only arrays of the two-fluid step (Hall and resistive EMFs, resistive RHS)
    ("Ehall_xi", EDGE_XI, 1), ("Ehall_eta", EDGE_ET, 1), ("Ehall_zeta", EDGE_ZE, 1),
    ("Eres_xi", EDGE_XI, 1), ("Eres_eta", EDGE_ET, 1), ("Eres_zeta", EDGE_ZE, 1),
    ("RHS_b_res_xi", FACE_XI, 1), ("RHS_b_res_eta", FACE_ET, 1), ("RHS_b_res_zeta", FACE_ZE, 1),
]
FIELD_ID = {name: i for i, (name, _, _) in enumerate(FIELDS)}
LOCATION = {name: loc for name, loc, _ in FIELDS}
NCOMP = {name: nc for name, _, nc in FIELDS}

# fields held by the device context / host case (the rest live only in the
# reference's solver as intermediate arrays)
STORED = ["U_H", "B_xi", "B_eta", "B_zeta", "Badd_xi", "Badd_eta", "Badd_zeta", "B_cell", "J_cell",
          "E_xi", "E_eta", "E_zeta", "Eface_xi", "Eface_eta", "Eface_zeta", "Jac", "JDxi", "JDet", "JDze",
          "beta_xi", "beta_eta", "beta_zeta", "alpha_xi", "alpha_eta", "alpha_zeta",
          "Bcell_from_Bface_w", "Jcell_from_Jedge_w", "U_Na", "Photo_rate", "drag_rate"]


def delta(loc, ax):
    """1 if cell-like along axis ``ax``, 0 if node-like."""
    if loc == CELL:
        return 1
    if loc == NODE:
        return 0
    if FACE_XI <= loc <= FACE_ZE:
        return 0 if ax == loc - FACE_XI else 1
    return 1 if ax == loc - EDGE_XI else 0


def shape(name, n, g):
    """AoS FieldBlock shape (i, j, k, m) with g ghost layers for n=(Ni,Nj,Nk)."""
    loc = LOCATION[name]
    return tuple(n[a] + 1 - delta(loc, a) + 2 * g for a in range(3)) + (NCOMP[name],)


def interior(name, n, g):
    """Slices selecting the interior DOFs of a field block array."""
    loc = LOCATION[name]
    return tuple(slice(g, g + n[a] + 1 - delta(loc, a)) for a in range(3))
