"""Host-side objects over the C-ABI.

``MhdContext`` owns a device context with one or more cases; ``LunarSolver``
mirrors the reference driver's solver object (Z4_Lunar/include/LunarSolver.h:
55-67: construct, ``Advance()``, ``StepOnce()``) on top of it; ``HostCase``
builds a case on the CPU only (no device needed).
"""
import ctypes as C
from dataclasses import dataclass, field as dfield

import numpy as np

from .capi import MhdDipoleDesc, MhdError, MhdOptions, MhdOTDesc, MhdWindDesc, lib
from .fields import FIELD_ID, shape

PREC = {"fp64": 0, "mixed": 1}


@dataclass
class OTDesc:
    """Orszag-Tang vortex in a periodic box of n^3 cells."""
    n: int = 64
    gamma: float = 5.0 / 3.0
    cfl: float = 0.4
    seed: int = 20261203
    noise: float = 1e-3

    def c(self):
        return MhdOTDesc(self.n, self.gamma, self.cfl, self.seed, self.noise)


@dataclass
class WindDesc:
    """Solar-wind shell case (6-block cubed sphere)."""
    nang: int = 64
    nrad: int = 128
    r_in: float = 1.0
    r_out: float = 8.0
    gamma: float = 5.0 / 3.0
    cfl: float = 0.4
    n_cc: float = 5.0
    v_kms: float = 400.0
    vdir: tuple = (-1.0, 0.0, 0.0)
    T_eV: float = 10.0
    B_nT: tuple = dfield(default_factory=lambda: (5.0 * 0.7071067811865476, 5.0 * 0.7071067811865476, 0.0))
    m_amu: float = 1.0

    def c(self):
        d = MhdWindDesc()
        d.nang, d.nrad, d.r_in, d.r_out = self.nang, self.nrad, self.r_in, self.r_out
        d.gamma, d.cfl, d.n_cc, d.v_kms = self.gamma, self.cfl, self.n_cc, self.v_kms
        for a in range(3):
            d.vdir[a] = self.vdir[a]
            d.B_nT[a] = self.B_nT[a]
        d.T_eV, d.m_amu = self.T_eV, self.m_amu
        return d

    @property
    def cells(self):
        return 6 * self.nang * self.nang * self.nrad


@dataclass
class MercuryDesc(WindDesc):
    """Mercury-configured shell: solar wind at 0.3-0.47 AU plus an offset
    dipole (190 nT equatorial surface field, centre +0.2 R_M along z).
    two_fluid: H+ and Na+ with the reference's Mercury sources (photo-
    ionization, drag, induction and ambipolar forces); otherwise the shared
    single-fluid step with the dipole background."""
    r_out: float = 10.0
    n_cc: float = 40.0
    T_eV: float = 15.0
    B_nT: tuple = (-15.2, 8.4, -8.5)
    B_eq_nT: float = 190.0
    offset_z: float = 0.2
    softening: float = 0.05
    two_fluid: bool = True
    ambipolar: bool = True
    hall: bool = True
    hall_r: tuple = (1.1, 1.2)
    # local artificial resistivity of Z5_Mercury/9999setup/3_MHD_Control.txt:33-38
    local_eta_max: float = 0.1
    local_center: tuple = (1.53, 0.0, 0.25)
    local_r_decay: float = 0.05
    local_r_cutoff: float = 0.05
    # interior resistivity (explicit subcycles): eta_max_mercury of 0_General_Setting.txt
    resist: bool = True
    resist_subcycles: int = 1
    resist_eta_max: float = 1.25e7

    def dipole(self):
        d = MhdDipoleDesc()
        d.B_eq_nT, d.offset_z, d.softening = self.B_eq_nT, self.offset_z, self.softening
        d.single_fluid, d.no_ambipolar, d.no_hall = int(not self.two_fluid), int(not self.ambipolar), int(not self.hall)
        d.hall_r_min, d.hall_r_max = self.hall_r
        d.local_eta_max, d.local_r_decay, d.local_r_cutoff = self.local_eta_max, self.local_r_decay, self.local_r_cutoff
        for a in range(3):
            d.local_center[a] = self.local_center[a]
        d.no_resistance, d.resist_subcycles, d.resist_eta_max = int(not self.resist), self.resist_subcycles, \
            self.resist_eta_max
        return d


@dataclass
class CaseDirDesc:
    """A Lunar case on the grid of a reference CASE directory (Plot3D grid,
    boundary and inner files); the wind comes from `wind` (its grid fields
    are ignored)."""
    path: str
    wind: WindDesc = dfield(default_factory=WindDesc)


def _check(rc, ctx=None, what=""):
    if rc != 0:
        msg = ""
        if ctx is not None:
            m = lib().mhd_last_error(ctx)
            msg = m.decode() if m else ""
        raise MhdError(f"{what} failed ({rc}): {msg}", rc)


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class HostCase:
    """A case built on the CPU: grid, topology, geometry, weights, ops, state."""

    def __init__(self, desc, ghost=2):
        self.h = C.c_void_p()
        if isinstance(desc, OTDesc):
            _check(lib().mhd_host_case_ot(C.byref(desc.c()), ghost, C.byref(self.h)), what="mhd_host_case_ot")
        elif isinstance(desc, MercuryDesc):
            _check(lib().mhd_host_case_mercury(C.byref(desc.c()), C.byref(desc.dipole()), ghost, C.byref(self.h)),
                   what="mhd_host_case_mercury")
        elif isinstance(desc, CaseDirDesc):
            rc = lib().mhd_host_case_from_dir(desc.path.encode(), C.byref(desc.wind.c()), ghost, C.byref(self.h))
            if rc:
                m = lib().mhd_host_last_error(None)
                raise MhdError(f"mhd_host_case_from_dir({desc.path}) failed ({rc}): {m.decode() if m else ''}", rc)
        else:
            _check(lib().mhd_host_case_lunar(C.byref(desc.c()), ghost, C.byref(self.h)), what="mhd_host_case_lunar")
        self.desc = desc
        self._info()

    def _info(self):
        nb, ni, nj, nk, g = (C.c_int() for _ in range(5))
        lib().mhd_host_case_info(self.h, C.byref(nb), C.byref(ni), C.byref(nj), C.byref(nk), C.byref(g))
        self.nblocks, self.n, self.g = nb.value, (ni.value, nj.value, nk.value), g.value
        self.ns = []  # cells of each block (self.n: block 0)
        for b in range(self.nblocks):
            _check(lib().mhd_host_block_info(self.h, b, C.byref(ni), C.byref(nj), C.byref(nk)), what="block_info")
            self.ns.append((ni.value, nj.value, nk.value))

    def __del__(self):
        if getattr(self, "h", None):
            lib().mhd_host_case_destroy(self.h)
            self.h = None

    @classmethod
    def from_parts(cls, consts, blocks, ghost=2):
        """A host case from a driver's own arrays (mhd_host_case_begin ..
        mhd_host_case_end): `consts` as constants() reports them, `blocks` as
        case_parts() returns them."""
        self = cls.__new__(cls)
        self.h = C.c_void_p()
        k = np.ascontiguousarray(consts, dtype=np.float64)
        _check(lib().mhd_host_case_begin(len(blocks), ghost, _dptr(k), C.byref(self.h)), what="mhd_host_case_begin")

        def err(what):
            m = lib().mhd_host_last_error(self.h)
            return f"{what}: {m.decode() if m else ''}"
        for ib, (ni, nj, nk, xyz, keep, ptrs, faces, parts) in enumerate(_parts_args(blocks, ghost)):
            rc = lib().mhd_host_add_block(self.h, ib, ni, nj, nk, _dptr(xyz[0]), _dptr(xyz[1]), _dptr(xyz[2]),
                                          xyz[0].size, ptrs)
            if rc:
                raise MhdError(err(f"mhd_host_add_block({ib})"), rc)
            rc = lib().mhd_host_set_topology(self.h, ib, faces)
            if rc:
                raise MhdError(err(f"mhd_host_set_topology({ib})"), rc)
            for f, npart, rows in parts:
                rc = lib().mhd_host_set_face_parts(self.h, ib, f, npart, rows)
                if rc:
                    raise MhdError(err(f"mhd_host_set_face_parts({ib}, {f})"), rc)
        rc = lib().mhd_host_case_end(self.h)
        if rc:
            raise MhdError(err("mhd_host_case_end"), rc)
        self.desc = None
        self._info()
        return self

    def field(self, name, block):
        a = np.empty(shape(name, self.ns[block], self.g), dtype=np.float64)
        _check(lib().mhd_host_download(self.h, block, FIELD_ID[name], _dptr(a), a.size), what=f"download {name}")
        return a

    def set_field(self, name, block, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        _check(lib().mhd_host_upload(self.h, block, FIELD_ID[name], _dptr(a), a.size), what=f"upload {name}")

    def faces(self, block):
        out = (C.c_int * 84)()
        _check(lib().mhd_host_faces(self.h, block, out), what="faces")
        return np.array(out, dtype=np.int32).reshape(6, 14)

    def face_parts(self, block, face):
        """Patches of a face several interface patches share: rows of
        (lo[2], hi[2], nb_block, nb_face, periodic, perm[3], sign[3],
        offset[3]); empty for other faces."""
        n = C.c_int()
        cap = 64
        out = (C.c_int * (16 * cap))()
        _check(lib().mhd_host_face_parts(self.h, block, face, out, cap, C.byref(n)), what="face parts")
        return np.array(out[:16 * n.value], dtype=np.int32).reshape(n.value, 16)

    def nodes(self, block):
        gn = self.g + 1
        s = tuple(self.ns[block][a] + 1 + 2 * gn for a in range(3))
        out = []
        for comp in range(3):
            a = np.empty(s, dtype=np.float64)
            _check(lib().mhd_host_nodes(self.h, block, comp, _dptr(a), a.size), what="nodes")
            out.append(a)
        return out

    def constants_two_fluid(self):
        out = np.zeros(31)
        lib().mhd_host_constants_two_fluid(self.h, _dptr(out))
        return out

    def constants(self):
        out = np.zeros(22)
        lib().mhd_host_constants(self.h, _dptr(out))
        return out

    def write_dir(self, path):
        """Write the case's grid as a reference CASE directory."""
        _check(lib().mhd_host_case_write_dir(self.h, path.encode()), what="write_dir")

    def write_restart(self, path, step=0, time=0.0):
        _check(lib().mhd_host_write_restart(self.h, path.encode(), step, time), what="write_restart")

    def read_restart(self, path):
        st, t = C.c_longlong(), C.c_double()
        _check(lib().mhd_host_read_restart(self.h, path.encode(), C.byref(st), C.byref(t)), what="read_restart")
        return st.value, t.value

    def write_tecplot(self, path, time=0.0):
        """Tecplot binary (.plt) of the case in the reference's Mixed mode."""
        _check(lib().mhd_host_write_tecplot(self.h, path.encode(), time), what="write_tecplot")

    def apply_sync(self, sync):
        _check(lib().mhd_host_apply_sync(self.h, sync), what="apply_sync")

    def op_count(self, sync):
        n = C.c_size_t()
        lib().mhd_host_op_count(self.h, sync, C.byref(n))
        return n.value

    def ops_independent(self, sync):
        """True when op list `sync` writes no element twice and reads none it writes."""
        v = C.c_int()
        _check(lib().mhd_host_ops_independent(self.h, sync, C.byref(v)), what="ops_independent")
        return bool(v.value)

    def emf_shell_ok(self):
        """True when the Eface sync of this case is what the one-pass face/edge
        EMF of a whole step assumes (shell enumeration, halo sources and
        destinations, wall zeros)."""
        v = C.c_int()
        _check(lib().mhd_host_emf_shell_check(self.h, C.byref(v)), what="emf_shell_check")
        return bool(v.value)


def _parts_args(blocks, ghost):
    """ctypes arguments of mhd_(host_)add_block / set_topology for each block
    of `blocks`: dicts with "n" (ni, nj, nk), "nodes" (x, y, z over the
    interior nodes, shape (ni+1, nj+1, nk+1)), "faces" (6 x 14 ints), an
    optional "parts" {face: rows of 16 ints} for faces of kind 2 (several
    interface patches) and an optional "geometry" {field name: AoS array}."""
    geo_names = ["Jac", "JDxi", "JDet", "JDze", "beta_xi", "beta_eta", "beta_zeta", "alpha_xi", "alpha_eta",
                 "alpha_zeta", "Bcell_from_Bface_w", "Jcell_from_Jedge_w"]
    out = []
    for blk in blocks:
        ni, nj, nk = blk["n"]
        xyz = [np.ascontiguousarray(a, dtype=np.float64) for a in blk["nodes"]]
        for a in xyz:
            if a.shape != (ni + 1, nj + 1, nk + 1):
                raise ValueError("nodes must be the (ni+1, nj+1, nk+1) interior nodes")
        geo = blk.get("geometry") or {}
        keep = [np.ascontiguousarray(geo[n], dtype=np.float64) if n in geo else None for n in geo_names]
        ptrs = (C.POINTER(C.c_double) * 12)(*[_dptr(a) if a is not None else None for a in keep])
        faces = (C.c_int * 84)(*[int(v) for v in np.asarray(blk["faces"]).ravel()])
        parts = []
        for f, rows in sorted((blk.get("parts") or {}).items()):
            flat = [int(v) for v in np.asarray(rows).ravel()]
            if len(flat) % 16:
                raise ValueError("face patches are rows of 16 ints")
            parts.append((int(f), len(flat) // 16, (C.c_int * max(len(flat), 1))(*flat)))
        out.append((ni, nj, nk, xyz, keep, ptrs, faces, parts))
    return out


def case_parts(hc, geometry=True):
    """The arrays a driver would hand over for the grid of host case `hc`:
    constants, and per block its extent, interior nodes, faces and
    (optionally) the geometry fields."""
    gn = hc.g + 1
    names = ["Jac", "JDxi", "JDet", "JDze", "beta_xi", "beta_eta", "beta_zeta", "alpha_xi", "alpha_eta",
             "alpha_zeta", "Bcell_from_Bface_w", "Jcell_from_Jedge_w"]
    blocks = []
    for b in range(hc.nblocks):
        n = hc.ns[b]
        nodes = [a[gn:gn + n[0] + 1, gn:gn + n[1] + 1, gn:gn + n[2] + 1] for a in hc.nodes(b)]
        blk = {"n": n, "nodes": nodes, "faces": hc.faces(b)}
        parts = {f: hc.face_parts(b, f) for f in range(6) if blk["faces"][f][0] == 2}
        if parts:
            blk["parts"] = parts
        if geometry:
            blk["geometry"] = {nm: hc.field(nm, b) for nm in names}
        blocks.append(blk)
    return hc.constants(), blocks


class MhdContext:
    """Device context holding one or more cases (all advanced together)."""

    def __init__(self, precision="fp64", recon=0, ghost=2, device=0, hbm_budget_gb=0.0, autotune=False,
                 xcd_order=True, overlap=True):
        opt = MhdOptions(device, PREC[precision], recon, ghost, hbm_budget_gb, 1 if autotune else 0,
                         1 if xcd_order else 0, 1 if overlap else 0)
        self.ctx = C.c_void_p()
        _check(lib().mhd_create(C.byref(opt), C.byref(self.ctx)), what="mhd_create")
        self.cases = []
        self.ghost = ghost

    def close(self):
        if getattr(self, "ctx", None):
            lib().mhd_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    def _add(self, desc, cid):
        if isinstance(desc, OTDesc):
            return lib().mhd_add_case_orszag_tang(self.ctx, C.byref(desc.c()), C.byref(cid))
        if isinstance(desc, MercuryDesc):
            return lib().mhd_add_case_mercury(self.ctx, C.byref(desc.c()), C.byref(desc.dipole()), C.byref(cid))
        if isinstance(desc, CaseDirDesc):
            return lib().mhd_add_case_from_dir(self.ctx, desc.path.encode(), C.byref(desc.wind.c()), C.byref(cid))
        return lib().mhd_add_case_lunar(self.ctx, C.byref(desc.c()), C.byref(cid))

    def add(self, desc):
        cid = C.c_int()
        _check(self._add(desc, cid), self.ctx, "add case")
        self.cases.append(desc)
        return cid.value

    def finalize(self):
        _check(lib().mhd_finalize(self.ctx), self.ctx, "mhd_finalize")

    def add_parts(self, consts, blocks):
        """A case from a driver's own arrays through mhd_case_begin,
        mhd_add_block, mhd_set_topology, mhd_set_face_parts and mhd_case_end."""
        cid = C.c_int()
        k = np.ascontiguousarray(consts, dtype=np.float64)
        _check(lib().mhd_case_begin(self.ctx, len(blocks), _dptr(k), C.byref(cid)), self.ctx, "mhd_case_begin")
        for ib, (ni, nj, nk, xyz, keep, ptrs, faces, parts) in enumerate(_parts_args(blocks, self.ghost)):
            _check(lib().mhd_add_block(self.ctx, cid, ib, ni, nj, nk, _dptr(xyz[0]), _dptr(xyz[1]), _dptr(xyz[2]),
                                       xyz[0].size, ptrs), self.ctx, f"mhd_add_block({ib})")
            _check(lib().mhd_set_topology(self.ctx, cid, ib, faces), self.ctx, f"mhd_set_topology({ib})")
            for f, npart, rows in parts:
                _check(lib().mhd_set_face_parts(self.ctx, cid, ib, f, npart, rows), self.ctx,
                       f"mhd_set_face_parts({ib}, {f})")
        _check(lib().mhd_case_end(self.ctx, cid), self.ctx, "mhd_case_end")
        self.cases.append(None)
        return cid.value

    def stage(self, name):
        """One stage of the step: "dt", "fluid", "face_emf", "edge_emf",
        "curl", "update", "derived", or "sync:<group>"."""
        if name.startswith("sync:"):
            _check(lib().mhd_sync(self.ctx, name[5:].encode()), self.ctx, name)
            return
        f = {"dt": lib().mhd_step_dt, "fluid": lib().mhd_fluid_rhs, "face_emf": lib().mhd_face_emf,
             "edge_emf": lib().mhd_edge_emf, "curl": lib().mhd_curl_update, "update": lib().mhd_apply_update,
             "derived": lib().mhd_derived}[name]
        _check(f(self.ctx), self.ctx, name)

    def try_add(self, desc):
        """Admit a case into the running context; None when the HBM budget is full."""
        cid = C.c_int()
        rc = self._add(desc, cid)
        if rc == -5:
            return None
        _check(rc, self.ctx, "admit case")
        self.cases.append(desc)
        return cid.value

    def adopt(self, host_case):
        """Move a ``HostCase`` into the context; None (host case intact) when
        the HBM budget is full."""
        cid = C.c_int()
        rc = lib().mhd_adopt_host_case(self.ctx, host_case.h, C.byref(cid))
        if rc == -5:
            return None
        _check(rc, self.ctx, "adopt case")
        self.cases.append(host_case.desc)
        return cid.value

    def retire(self, cid):
        _check(lib().mhd_retire(self.ctx, cid), self.ctx, "mhd_retire")

    def case_steps(self, cid):
        v = C.c_longlong()
        _check(lib().mhd_case_steps(self.ctx, cid, C.byref(v)), self.ctx, "case_steps")
        return v.value

    def case_error(self, cid):
        """True when a step of case ``cid`` met the dt abort condition."""
        v = C.c_int()
        _check(lib().mhd_case_error(self.ctx, cid, C.byref(v)), self.ctx, "case_error")
        return bool(v.value)

    def tuning(self):
        """kernel -> (workgroup rows, tile order, "per-case", "geo-shared" (cases of one
        grid back to back per tile) or "geo-loop" (one tile updates every case of its grid))"""
        out = {}
        for k, name in enumerate(self.profile_read()):
            rows, inner, sh = C.c_int(), C.c_int(), C.c_int()
            _check(lib().mhd_tuning(self.ctx, k, C.byref(rows), C.byref(inner), C.byref(sh)), self.ctx, "tuning")
            out[name] = (rows.value, "i-inner" if inner.value else "j-inner",
                         ("per-case", "geo-shared", "geo-loop")[sh.value])
        return out

    def set_tuning(self, kernel, rows, i_inner=False, geo_shared=False):
        """Pin a kernel's launch shape (kernel by name, as in tuning());
        geo_shared 0/False, 1/True or 2 (member-looping tiles, k_derived)."""
        k = list(self.profile_read()).index(kernel)
        _check(lib().mhd_set_tuning(self.ctx, k, rows, int(i_inner), int(geo_shared)), self.ctx, "set_tuning")

    def autotune(self):
        """Re-tune launch shapes on the cases resident now."""
        _check(lib().mhd_autotune(self.ctx), self.ctx, "autotune")

    def step(self, n=1, graph=False):
        f = lib().mhd_step_graph if graph else lib().mhd_step
        _check(f(self.ctx, n), self.ctx, "mhd_step")

    def sync(self):
        _check(lib().mhd_synchronize(self.ctx), self.ctx, "sync")

    def refresh(self):
        """Re-sync U/B ghosts and recompute derived fields after uploads."""
        _check(lib().mhd_refresh(self.ctx), self.ctx, "refresh")

    def write_restart(self, cid, path, step=0, time=0.0):
        _check(lib().mhd_write_restart(self.ctx, cid, path.encode(), step, time), self.ctx, "write_restart")

    def read_restart(self, cid, path):
        st, t = C.c_longlong(), C.c_double()
        _check(lib().mhd_read_restart(self.ctx, cid, path.encode(), C.byref(st), C.byref(t)), self.ctx,
               "read_restart")
        return st.value, t.value

    def write_tecplot(self, cid, path, time=0.0):
        """Tecplot binary (.plt) of case ``cid`` in the reference's Mixed mode."""
        _check(lib().mhd_write_tecplot(self.ctx, cid, path.encode(), time), self.ctx, "write_tecplot")

    def info(self, cid):
        nb, ni, nj, nk, g = (C.c_int() for _ in range(5))
        _check(lib().mhd_case_info(self.ctx, cid, C.byref(nb), C.byref(ni), C.byref(nj), C.byref(nk), C.byref(g)),
               self.ctx, "info")
        return nb.value, (ni.value, nj.value, nk.value), g.value

    def block_n(self, cid, block):
        ni, nj, nk = C.c_int(), C.c_int(), C.c_int()
        _check(lib().mhd_block_info(self.ctx, cid, block, C.byref(ni), C.byref(nj), C.byref(nk)), self.ctx,
               "block_info")
        return ni.value, nj.value, nk.value

    def field(self, cid, name, block):
        a = np.empty(shape(name, self.block_n(cid, block), self.ghost), dtype=np.float64)
        _check(lib().mhd_download(self.ctx, cid, block, FIELD_ID[name], _dptr(a), a.size), self.ctx,
               f"download {name}")
        return a

    def set_field(self, cid, name, block, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        _check(lib().mhd_upload(self.ctx, cid, block, FIELD_ID[name], _dptr(a), a.size), self.ctx, f"upload {name}")

    def compute_dt(self, cid):
        v = C.c_double()
        _check(lib().mhd_compute_dt(self.ctx, cid, C.byref(v)), self.ctx, "compute_dt")
        return v.value

    def last_dt(self, cid):
        v = C.c_double()
        _check(lib().mhd_last_dt(self.ctx, cid, C.byref(v)), self.ctx, "last_dt")
        return v.value

    def diagnostics(self, cid):
        """Min/max of rho_H, p_H, rho_Na, p_Na, Bx, By, Bz, |B|^2/2 and max |div B|."""
        out = np.zeros(17)
        _check(lib().mhd_diagnostics(self.ctx, cid, _dptr(out)), self.ctx, "diagnostics")
        names = ["rhoH", "pH", "rhoNa", "pNa", "Bx", "By", "Bz", "Pmag"]
        d = {n: (out[q], out[8 + q]) for q, n in enumerate(names)}
        d["divB_max"] = out[16]
        return d

    def diagnostics_noalloc(self, cids):
        """Raw 17 values per case (mhd_diagnostics order) of up to 64 cases,
        through the entry point that allocates no device memory."""
        ids = (C.c_int * len(cids))(*cids)
        out = np.zeros(17 * max(len(cids), 1))
        _check(lib().mhd_diagnostics_noalloc(self.ctx, ids, len(cids), _dptr(out)), self.ctx,
               "diagnostics_noalloc")
        return out.reshape(-1, 17)

    def diagnostics_raw(self, cid):
        out = np.zeros(17)
        _check(lib().mhd_diagnostics(self.ctx, cid, _dptr(out)), self.ctx, "diagnostics")
        return out

    def divb_max(self, cid):
        v = C.c_double()
        _check(lib().mhd_divb_max(self.ctx, cid, C.byref(v)), self.ctx, "divb")
        return v.value

    def profile(self, on):
        lib().mhd_profile(self.ctx, 1 if on else 0)

    def profile_read(self):
        out = {}
        k = 0
        while True:
            ms, n, name = C.c_double(), C.c_int(), C.c_char_p()
            if lib().mhd_profile_read(self.ctx, k, C.byref(ms), C.byref(n), C.byref(name)) != 0:
                break
            out[name.value.decode()] = (ms.value, n.value)
            k += 1
        return out

    def memory_bytes(self):
        v = C.c_double()
        lib().mhd_memory(self.ctx, C.byref(v))
        return v.value


class LunarSolver:
    """Solver object with the reference driver's life cycle.

    ``LunarSolver(desc)`` sets up grid, topology, geometry, halo patterns,
    boundary registry and the initial state (reference constructor,
    Z4_Lunar/src/LunarSolver.cpp:12-235); ``StepOnce()`` is one explicit step
    (Z4_Lunar/src/LunarSolver_Driver.cpp:21-92); ``Advance(nsteps)`` loops it
    (the reference loops until its stop switch, :6-19).
    """

    def __init__(self, desc, precision="fp64", recon=0, ghost=2, device=0):
        self.ctx = MhdContext(precision, recon, ghost, device)
        self.case = self.ctx.add(desc)
        self.ctx.finalize()
        self.step = 0
        self.time = 0.0

    def StepOnce(self):
        self.ctx.step(1)
        self.step += 1
        self.time += self.ctx.last_dt(self.case)
        return False

    def Advance(self, nsteps, graph=True):
        self.ctx.step(nsteps, graph=graph)
        self.step += nsteps

    def field(self, name, block):
        return self.ctx.field(self.case, name, block)
