"""ctypes view of the C ABI declared in include/aether_fdtd.h.

The engine library is ``aether_amd/libaether_fdtd.so`` (built by
``aether_amd/csrc/Makefile``).  :func:`engine` loads it and fails loudly when it
is missing: there is no fallback path.  The same struct definitions and the
:class:`Grid` wrapper also serve the test-side oracle library, whose functions
carry the ``aeoracle_`` prefix instead of ``aefdtd_``.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ENGINE_SO = os.path.join(HERE, "libaether_fdtd.so")

OK, EINVAL, ENOMEM, EDEVICE, ESTATE, ECANCELLED, ERANGE = 0, 1, 2, 3, 4, 5, 6
MODE_NORM, MODE_CUSTOM = 0, 5
FP64, FP32, FP16 = 0, 1, 2
NORMAL_X, NORMAL_XM, NORMAL_Y, NORMAL_YM, NORMAL_Z, NORMAL_ZM, NORMAL_UNDEF = range(7)
SENSOR_POYNTING, SENSOR_FIELDBLOCK, SENSOR_FIELDMAP, SENSOR_FIELDMAP2 = range(4)
KIND_NORMAL, KIND_SINGLE, KIND_CUSTOM = 0, 1, 2
EX, EY, EZ, HX, HY, HZ = range(6)

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class CMaterial(C.Structure):
    _fields_ = [("eps_inf", C.c_double),
                ("n_debye", C.c_int), ("debye", _dp),
                ("n_drude", C.c_int), ("drude", _dp),
                ("n_lorentz", C.c_int), ("lorentz", _dp),
                ("n_critpoint", C.c_int), ("critpoint", _dp)]


class CPulse(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("lambda_", "omega", "k", "kx", "ky", "kz", "n_ind", "dz", "C1",
                 "Ex0", "Ey0", "Ez0", "Hx0", "Hy0", "Hz0")]


class CSourceReq(C.Structure):
    _fields_ = [("type", C.c_int), ("orientation", C.c_int),
                ("x1", C.c_int), ("x2", C.c_int), ("y1", C.c_int), ("y2", C.c_int),
                ("z1", C.c_int), ("z2", C.c_int),
                ("lambda_min", C.c_double), ("lambda_max", C.c_double)]


class CMode(C.Structure):
    _fields_ = [("Nt", C.c_int), ("Dx", C.c_double), ("Dy", C.c_double), ("Dz", C.c_double),
                ("time_mod", C.c_double), ("tapering", C.c_int),
                ("pml", C.c_int * 6), ("pad", C.c_int * 6),
                ("kappa", C.c_double * 6), ("sigma", C.c_double * 6), ("alpha", C.c_double * 6),
                ("Nl", C.c_int), ("lambda_min", C.c_double), ("lambda_max", C.c_double),
                ("polarization", C.c_int), ("polar_angle", C.c_double),
                ("n_disable", C.c_int), ("disable_fields", _ip),
                ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("matsgrid", C.POINTER(C.c_uint32)),
                ("n_materials", C.c_int), ("materials", C.POINTER(CMaterial)),
                ("precision", C.c_int),
                ("n_sources", C.c_int), ("sources", C.POINTER(CSourceReq)),
                ("time_type", C.c_int), ("cc_step", C.c_int), ("cc_coeff", C.c_double),
                ("cc_lmin", C.c_double), ("cc_lmax", C.c_double), ("cc_quant", C.c_int),
                ("cc_layout", C.c_char * 4), ("cc_seed", C.c_uint)]


class CSensorReq(C.Structure):
    _fields_ = [("box", C.c_int), ("orientation", C.c_int),
                ("x1", C.c_int), ("x2", C.c_int), ("y1", C.c_int), ("y2", C.c_int),
                ("z1", C.c_int), ("z2", C.c_int), ("disable", C.c_int * 6),
                ("Nl", C.c_int), ("lambda_min", C.c_double), ("lambda_max", C.c_double),
                ("kind", C.c_int)]


class CNormalResult(C.Structure):
    _fields_ = [("lambda_", _dp), ("R", _dp), ("T", _dp),
                ("refx", _dp), ("refy", _dp), ("transx", _dp), ("transy", _dp),
                ("index_sup", C.c_double), ("index_sub", C.c_double), ("Nt_run", C.c_int)]


OP_BLOCK, OP_SPHERE, OP_CYLINDER, OP_CONE, OP_ELLIPSOID, OP_LAYER, OP_COATING = range(7)


class CStructureOp(C.Structure):
    _fields_ = [("type", C.c_int), ("mat_index", C.c_int), ("p", C.c_double * 7)]


class CStructure(C.Structure):
    _fields_ = [("n_ops", C.c_int), ("ops", C.POINTER(CStructureOp)),
                ("default_material", C.c_int),
                ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
                ("flip", C.c_int * 3), ("loop", C.c_int * 3)]


def _sig(lib, p):
    """Declares argument/return types of every `p`-prefixed function."""
    vp = C.c_void_p
    d = C.c_double
    i = C.c_int
    table = {
        "create": (i, [C.POINTER(vp), i, i, i, i, d, d, d, d, i, _ip, _ip, i]),
        "destroy": (i, [vp]),
        "last_error": (C.c_char_p, [vp]),
        "layout": (i, [vp, _ip]),
        "set_pml": (i, [vp, i, d, d, d]),
        "set_tapering": (i, [vp, i]),
        "set_matsgrid": (i, [vp, C.POINTER(C.c_uint32), i, i, i]),
        "set_material": (i, [vp, C.c_uint, C.POINTER(CMaterial)]),
        "disable_fields": (i, [vp, _ip, i]),
        "bootstrap": (i, [vp]),
        "update_E": (i, [vp]),
        "update_H": (i, [vp]),
        "reset_fields": (i, [vp]),
        "read_field": (i, [vp, i, _dp]),
        "write_field": (i, [vp, i, _dp]),
        "set_plane_injection": (i, [vp, i, i, _dp, _dp, _dp, _dp]),
        "set_tfsf": (i, [vp, vp, i, i, i, i, i, i, i, i]),
        "add_spect_poynting": (i, [vp, i, i, i, i, i, i, i, i, _dp, _ip]),
        "feed_sensors": (i, [vp]),
        "add_field_block": (i, [vp, i, i, i, i, i, i, d, _ip]),
        "add_field_map": (i, [vp, i, i, i, i, i, i, i, d, i, i, _ip]),
        "field_sensor_read": (i, [vp, i, _dp, _ip]),
        "field_sensor_mats": (i, [vp, i, C.POINTER(C.c_uint32)]),
        "sensor_poynting": (i, [vp, i, _dp]),
        "sensor_poynting_noalloc": (i, [vp, i, _dp]),
        "sensor_spectra": (i, [vp, i, _dp, _ip]),
        "add_plane_sum": (i, [vp, i, C.POINTER(CPulse), d, i, _dp, _ip]),
        "read_plane_sum": (i, [vp, i, i, _dp]),
        "pulse_init": (i, [C.POINTER(CPulse), d, d, d, d, d, d, d, d, d]),
        "pulse_eval": (d, [C.POINTER(CPulse), i, d, d, d, d]),
        "run_normal_incidence": (i, [C.POINTER(CMode), i, C.POINTER(CSensorReq),
                                     C.POINTER(CNormalResult), _dp, _dp, _ip]),
        "run_single_particle": (i, [C.POINTER(CMode), i, C.POINTER(CSensorReq), _dp, _dp, _ip]),
        "run_custom": (i, [C.POINTER(CMode), i, C.POINTER(CSensorReq), _dp, _dp, _ip]),
        "add_oscillator": (i, [vp, i, i, i, i, i, i, i, d, d, _ip]),
        "mode_finalize": (i, [C.POINTER(CMode), i, i, C.POINTER(CSensorReq)]),
        # engine only
        "run_create": (i, [C.POINTER(vp), i, C.POINTER(CMode), i, C.POINTER(CSensorReq)]),
        "run_step": (i, [vp, i]),
        "run_cancel": (i, [vp]),
        "run_progress": (i, [vp, _ip, _ip]),
        "run_grid": (vp, [vp]),
        "run_error": (C.c_char_p, [vp]),
        "run_finish": (i, [vp, C.POINTER(CNormalResult), _dp, _dp]),
        "run_destroy": (i, [vp]),
        "run_write": (i, [vp, C.c_char_p, C.c_char_p, C.POINTER(C.c_char_p), _dp, C.POINTER(CNormalResult)]),
        "synchronize": (i, [vp]),
        "profile_enable": (i, [vp, i]),
        "profile_kinds": (i, [vp, i]),
        "profile_read": (i, [vp, C.c_char_p, _dp, C.POINTER(C.c_long)]),
        "device_bytes": (i, [vp, C.POINTER(C.c_size_t)]),
        "stats": (i, [vp, C.POINTER(C.c_long)]),
        "autotune": (i, [vp, i, _ip]),
        "set_storage_scale": (i, [vp, _ip]),
        "get_storage_scale": (i, [vp, _ip]),
        "sensor_out_len": (C.c_size_t, [C.POINTER(CSensorReq)]),
        "set_tiles": (i, [vp, i, i]),
        "set_step_tiles": (i, [vp, i]),
        "get_step_tiles": (i, [vp, _ip]),
        "batch_create": (i, [C.POINTER(vp), i]),
        "batch_destroy": (i, [vp]),
        "batch_error": (C.c_char_p, [vp]),
        "batch_add": (i, [vp, vp]),
        "batch_remove": (i, [vp, vp]),
        "batch_step": (i, [vp, i]),
        "batch_running": (i, [vp]),
        "batch_synchronize": (i, [vp]),
        "batch_set_tiles": (i, [vp, i, i]),
        "batch_autotune": (i, [vp, _ip]),
        "batch_set_skew": (i, [vp, i]),
        "batch_info": (i, [vp, _ip]),
        "batch_profile": (i, [vp, i, _dp, C.POINTER(C.c_long)]),
        "server_create": (i, [C.POINTER(vp), C.c_size_t]),
        "server_destroy": (i, [vp]),
        "server_submit": (i, [vp, i, C.POINTER(CMode), i, C.POINTER(CSensorReq), C.POINTER(C.c_long)]),
        "server_cancel": (i, [vp, C.c_long]),
        "server_wait": (i, [vp, C.c_long, d, _ip]),
        "server_result": (i, [vp, C.c_long, _dp, C.c_size_t, _dp]),
        "server_error": (C.c_char_p, [vp, C.c_long]),
        "server_release": (i, [vp, C.c_long]),
        "server_stats": (i, [vp, _dp]),
        "estimate_bytes": (i, [i, C.POINTER(CMode), i, C.POINTER(CSensorReq), C.POINTER(C.c_size_t)]),
        "stream_bandwidth": (i, [C.c_size_t, i, _dp]),
        "stream_mix_bandwidth": (i, [C.c_size_t, i, _dp]),
        "discretize": (i, [C.POINTER(CStructure), i, i, i, d, d, d, C.POINTER(C.c_uint32)]),
        # oracle only
        "step": (i, [vp, i]),
        "set_threads": (None, [i]),
        "get_threads": (i, []),
        "material_coeffs": (i, [C.POINTER(CMaterial), d, d, d, d, _dp, _ip]),
        "material_eval": (i, [C.POINTER(CMaterial), d, _dp, _dp]),
        "pml_profile": (i, [i, i, i, d, d, _dp, _dp, _dp, d, _dp]),
        "s_curve": (d, [d, d, d]),
        "nearest_integer": (i, [d]),
    }
    for name, (res, args) in table.items():
        f = getattr(lib, p + name, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args
    return lib


_ENGINE = None


def engine():
    """Loads the HIP engine library (raises if it has not been built)."""
    global _ENGINE
    if _ENGINE is None:
        if not os.path.exists(ENGINE_SO):
            raise RuntimeError(f"HIP engine library missing: {ENGINE_SO} (run __graft_entry__.build())")
        _ENGINE = Backend(C.CDLL(ENGINE_SO), "aefdtd_")
    return _ENGINE


class FDTDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Backend:
    """A loaded library plus its function prefix."""

    def __init__(self, lib, prefix):
        self.lib = _sig(lib, prefix)
        self.prefix = prefix

    def fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def has(self, name):
        return hasattr(self.lib, self.prefix + name)

    def check(self, rc, handle=None):
        if rc != OK:
            msg = self.fn("last_error")(handle)
            raise FDTDError(rc, msg.decode() if msg else "")


# --------------------------------------------------------------------------
# Python-side descriptions of materials, modes and sensors
# --------------------------------------------------------------------------
@dataclass
class Material:
    """eps_inf plus pole lists; the FDTD-compatible part of Aether's Material."""
    eps_inf: float = 1.0
    debye: list = field(default_factory=list)      # (ds, t0)
    drude: list = field(default_factory=list)      # (wd, g)
    lorentz: list = field(default_factory=list)    # (A, O, G)
    critpoint: list = field(default_factory=list)  # (A, O, P, G)

    def to_c(self, keep):
        def arr(v, w):
            a = np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1, w) if len(v) else np.zeros((0, w)))
            keep.append(a)
            return len(a), a.ctypes.data_as(_dp)
        m = CMaterial()
        m.eps_inf = self.eps_inf
        m.n_debye, m.debye = arr(self.debye, 2)
        m.n_drude, m.drude = arr(self.drude, 2)
        m.n_lorentz, m.lorentz = arr(self.lorentz, 3)
        m.n_critpoint, m.critpoint = arr(self.critpoint, 4)
        return m


@dataclass
class SensorReq:
    box: bool = True
    orientation: int = NORMAL_Z
    x1: int = 0
    x2: int = 0
    y1: int = 0
    y2: int = 0
    z1: int = 0
    z2: int = 0
    disable: tuple = (0, 0, 0, 0, 0, 0)
    Nl: int = 64
    lambda_min: float = 400e-9
    lambda_max: float = 800e-9
    kind: int = SENSOR_POYNTING

    def to_c(self):
        s = CSensorReq()
        s.box = int(self.box)
        s.orientation = self.orientation
        s.x1, s.x2, s.y1, s.y2, s.z1, s.z2 = self.x1, self.x2, self.y1, self.y2, self.z1, self.z2
        for f in range(6):
            s.disable[f] = int(self.disable[f])
        s.Nl, s.lambda_min, s.lambda_max = self.Nl, self.lambda_min, self.lambda_max
        s.kind = self.kind
        return s

    def spans(self):
        """In-plane spans of a field map (FieldMap::initialize, s_fieldmap.cpp:433-447)."""
        d = (max(0, self.x2 - self.x1), max(0, self.y2 - self.y1), max(0, self.z2 - self.z1))
        n = self.orientation // 2
        return {0: (d[1], d[2]), 1: (d[0], d[2])}.get(n, (d[0], d[1]))

    def out_len(self):
        """Doubles this request adds to a driver's sensor_out."""
        if self.kind == SENSOR_FIELDBLOCK:
            return 12 * max(0, self.x2 - self.x1) * max(0, self.y2 - self.y1) * max(0, self.z2 - self.z1)
        if self.kind == SENSOR_FIELDMAP:
            s1, s2 = self.spans()
            return 6 * s1 * s2
        if self.kind == SENSOR_FIELDMAP2:
            s1, s2 = self.spans()
            return 12 * self.Nl * s1 * s2
        return 7 * self.Nl if self.box else self.Nl


@dataclass
class SourceReq:
    """An oscillator source (Source_generator with SOURCE_GEN_OSCILLATOR), in structure cells."""
    orientation: int = NORMAL_Z
    x1: int = 0
    x2: int = 0
    y1: int = 0
    y2: int = 0
    z1: int = 0
    z2: int = 0
    lambda_min: float = 400e-9
    lambda_max: float = 800e-9
    type: int = 0

    def to_c(self):
        s = CSourceReq()
        s.type, s.orientation = self.type, self.orientation
        s.x1, s.x2, s.y1, s.y2, s.z1, s.z2 = self.x1, self.x2, self.y1, self.y2, self.z1, self.z2
        s.lambda_min, s.lambda_max = self.lambda_min, self.lambda_max
        return s


@dataclass
class ModeSpec:
    """The FDTD_Mode fields Aether's drivers read, plus the discretised structure."""
    matsgrid: np.ndarray                      # (nz, ny, nx) uint32, x fastest
    materials: list
    Nt: int = 5000
    Dx: float = 5e-9
    Dy: float = 5e-9
    Dz: float = 5e-9
    time_mod: float = 1.0
    tapering: int = 0
    pml: tuple = (0, 0, 0, 0, 0, 0)
    pad: tuple = (0, 0, 0, 0, 0, 0)
    kappa: tuple = (25.0,) * 6
    sigma: tuple = (1.0,) * 6
    alpha: tuple = (0.2,) * 6
    Nl: int = 481
    lambda_min: float = 370e-9
    lambda_max: float = 850e-9
    polarization: int = 0                     # 0 TE, 1 TM, 2 mix
    polar_angle: float = 0.0
    disable_fields: tuple = ()
    precision: int = FP64
    sources: list = field(default_factory=list)   # SourceReq, custom mode
    time_type: int = 0                        # 1: auto_tsteps, energy criterion; 2: spectral criterion
    cc_step: int = 100
    cc_coeff: float = 1e-3
    cc_lmin: float = 370e-9                   # spectral criterion (FDTD_Mode defaults, lua_fdtd.cpp:255-257)
    cc_lmax: float = 850e-9
    cc_quant: int = 500
    cc_layout: str = "nnb"
    cc_seed: int = 0

    def to_c(self, keep):
        mg = np.ascontiguousarray(self.matsgrid, dtype=np.uint32)
        keep.append(mg)
        mats = (CMaterial * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            mats[i] = m.to_c(keep)
        keep.append(mats)
        dis = np.ascontiguousarray(np.asarray(self.disable_fields, dtype=np.int32).reshape(-1))
        keep.append(dis)
        c = CMode()
        c.Nt = self.Nt
        c.Dx, c.Dy, c.Dz = self.Dx, self.Dy, self.Dz
        c.time_mod = self.time_mod
        c.tapering = self.tapering
        for f in range(6):
            c.pml[f], c.pad[f] = self.pml[f], self.pad[f]
            c.kappa[f], c.sigma[f], c.alpha[f] = self.kappa[f], self.sigma[f], self.alpha[f]
        c.Nl, c.lambda_min, c.lambda_max = self.Nl, self.lambda_min, self.lambda_max
        c.polarization, c.polar_angle = self.polarization, self.polar_angle
        c.n_disable = len(dis)
        c.disable_fields = dis.ctypes.data_as(_ip)
        c.nz, c.ny, c.nx = mg.shape
        c.matsgrid = mg.ctypes.data_as(C.POINTER(C.c_uint32))
        c.n_materials = len(self.materials)
        c.materials = mats
        c.precision = self.precision
        srcs = (CSourceReq * max(1, len(self.sources)))(*[q.to_c() for q in self.sources])
        keep.append(srcs)
        c.n_sources = len(self.sources)
        c.sources = srcs
        c.time_type, c.cc_step, c.cc_coeff = self.time_type, self.cc_step, self.cc_coeff
        c.cc_lmin, c.cc_lmax, c.cc_quant = self.cc_lmin, self.cc_lmax, self.cc_quant
        c.cc_layout = self.cc_layout.encode()[:3]
        c.cc_seed = int(self.cc_seed) & 0xFFFFFFFF
        return c


class Structure:
    """A structure as a structure script builds it: the operation list made by
    the script's add_* calls (structure/lua_structure.cpp:292-547), the default
    material, the nominal size lx, ly, lz and the flip / loop modifiers
    (structure/structure.cpp:326-352).  Method names and argument order follow
    the script functions; lengths in metres."""

    def __init__(self, lx=300e-9, ly=300e-9, lz=300e-9):   # Structure::Structure defaults (structure.cpp:52-66)
        self.ops = []
        self.default = 0
        self.lx, self.ly, self.lz = lx, ly, lz
        self.flips = (0, 0, 0)
        self.loops = (0, 0, 0)

    def _add(self, kind, params, index):
        p = list(map(float, params)) + [0.0] * (7 - len(params))
        self.ops.append((kind, int(index), p))
        return self

    def add_block(self, x1, x2, y1, y2, z1, z2, index):
        return self._add(OP_BLOCK, (x1, x2, y1, y2, z1, z2), index)

    def add_sphere(self, x, y, z, r, index):
        return self._add(OP_SPHERE, (x, y, z, r), index)

    def add_cylinder(self, x, y, z, dx, dy, dz, r, index):
        return self._add(OP_CYLINDER, (x, y, z, dx, dy, dz, r), index)

    def add_cone(self, x, y, z, dx, dy, dz, r, index):
        return self._add(OP_CONE, (x, y, z, dx, dy, dz, r), index)

    def add_ellipsoid(self, x, y, z, rx, ry, rz, index):
        return self._add(OP_ELLIPSOID, (x, y, z, rx, ry, rz), index)

    def add_layer(self, axis, z1, z2, index):
        """axis "X"/"x", "Y"/"y", anything else z, as add_layer (lua_structure.cpp:420-440)."""
        a = 0 if axis in ("X", "x") else (1 if axis in ("Y", "y") else 2)
        return self._add(OP_LAYER, (a, z1, z2), index)

    def add_coating(self, theta, phi, thickness, delta, origin, index):
        """add_coating (lua_structure.cpp:314-331): angles in degrees."""
        return self._add(OP_COATING, (theta, phi, thickness, delta, int(origin)), index)

    def default_material(self, index):
        self.default = int(index)
        return self

    def flip(self, x, y, z):
        self.flips = (int(x != 0), int(y != 0), int(z != 0))
        return self

    def loop(self, x, y, z):
        self.loops = (int(x != 0), int(y != 0), int(z != 0))
        return self

    def to_c(self, keep):
        arr = (CStructureOp * max(1, len(self.ops)))()
        for n, (kind, index, p) in enumerate(self.ops):
            arr[n].type, arr[n].mat_index = kind, index
            for q in range(7):
                arr[n].p[q] = p[q]
        keep.append(arr)
        c = CStructure()
        c.n_ops, c.ops = len(self.ops), arr
        c.default_material = self.default
        c.lx, c.ly, c.lz = self.lx, self.ly, self.lz
        for d in range(3):
            c.flip[d], c.loop[d] = self.flips[d], self.loops[d]
        return c

    def cells(self, Dx, Dy, Dz):
        """FD_Mode::compute_discretization (finite_differences/core/lua_fd.cpp:69-82):
        max(1, nearest_integer(l / D)) per axis, rounding halves away from zero."""
        def ni(v):
            return int(math.floor(v + 0.5)) if v >= 0 else -int(math.floor(-v + 0.5))
        return tuple(max(1, ni(l / D)) for l, D in ((self.lx, Dx), (self.ly, Dy), (self.lz, Dz)))


def discretize(be: Backend, structure: Structure, N, D):
    """Structure::discretize over an N = (Nx, Ny, Nz) grid of cell sizes D:
    returns the (Nz, Ny, Nx) uint32 material grid."""
    keep = []
    cs = structure.to_c(keep)
    out = np.zeros(tuple(N)[::-1], dtype=np.uint32)
    rc = be.fn("discretize")(C.byref(cs), int(N[0]), int(N[1]), int(N[2]), float(D[0]), float(D[1]), float(D[2]),
                             out.ctypes.data_as(C.POINTER(C.c_uint32)))
    be.check(rc, None)
    return out


def mode_finalize(be: Backend, spec: ModeSpec, kind, sensors=()):
    """FDTD_Mode::finalize's grid sizing: returns a copy of spec with pml/pad updated."""
    import copy
    keep = []
    cm = spec.to_c(keep)
    sreq = (CSensorReq * max(1, len(sensors)))(*[q.to_c() for q in sensors])
    be.check(be.fn("mode_finalize")(C.byref(cm), kind, len(sensors), sreq), None)
    out = copy.copy(spec)
    out.pml = tuple(cm.pml[f] for f in range(6))
    out.pad = tuple(cm.pad[f] for f in range(6))
    return out


def _flag_ptr(end_flag):
    """A ctypes c_int (the caller's end_computation flag) as an int pointer, or None."""
    return None if end_flag is None else C.cast(C.byref(end_flag), _ip)


def run_custom(be: Backend, spec: ModeSpec, sensors=(), want_fields=False, end_flag=None):
    """Runs the custom-mode driver (spec already finalized); sensor outputs and fields.
    end_flag: optional ctypes.c_int polled after every step (non-zero ends the run)."""
    keep = []
    cm = spec.to_c(keep)
    sreq = (CSensorReq * max(1, len(sensors)))(*[q.to_c() for q in sensors])
    nsout = sum(q.out_len() for q in sensors)
    sout = np.zeros(max(1, nsout))
    fields = None
    if want_fields:
        fields = np.zeros((6,) + custom_dims(spec)[::-1])
    rc = be.fn("run_custom")(C.byref(cm), len(sensors), sreq, _dptr(sout),
                             _dptr(fields) if fields is not None else None, _flag_ptr(end_flag))
    if rc not in (OK, ECANCELLED):
        be.check(rc, None)
    out = {"sensors": sout[:nsout], "cancelled": rc == ECANCELLED}
    if fields is not None:
        out["fields"] = fields
    return out


def custom_dims(spec):
    """Full-grid (Nx, Ny, Nz) of a finalized custom-mode spec (FDTD ctor extension)."""
    nz, ny, nx = spec.matsgrid.shape
    n = [nx, ny, nz]
    for a in range(3):
        n[a] += spec.pml[2 * a] + spec.pml[2 * a + 1] + spec.pad[2 * a] + spec.pad[2 * a + 1]
    return tuple(n)


def _dptr(a):
    return a.ctypes.data_as(_dp)


def run_normal_incidence(be: Backend, spec: ModeSpec, sensors=(), want_fields=False, end_flag=None):
    """Runs the normal-incidence driver; returns dict of spectra (and fields).
    end_flag as run_custom."""
    keep = []
    cm = spec.to_c(keep)
    Nl = spec.Nl
    out = {k: np.zeros(Nl) for k in ("lambda", "R", "T")}
    cplx = {k: np.zeros(2 * Nl) for k in ("refx", "refy", "transx", "transy")}
    res = CNormalResult()
    res.lambda_, res.R, res.T = _dptr(out["lambda"]), _dptr(out["R"]), _dptr(out["T"])
    res.refx, res.refy = _dptr(cplx["refx"]), _dptr(cplx["refy"])
    res.transx, res.transy = _dptr(cplx["transx"]), _dptr(cplx["transy"])
    sreq = (CSensorReq * max(1, len(sensors)))(*[s.to_c() for s in sensors])
    nsout = sum(s.out_len() for s in sensors)
    sout = np.zeros(max(1, nsout))
    fields = None
    if want_fields:
        fields = _normal_fields_buffer(spec)
    rc = be.fn("run_normal_incidence")(C.byref(cm), len(sensors), sreq, C.byref(res), _dptr(sout),
                                        _dptr(fields) if fields is not None else None, _flag_ptr(end_flag))
    if rc not in (OK, ECANCELLED):
        be.check(rc, None)
    for k, v in cplx.items():
        out[k] = v[0::2] + 1j * v[1::2]
    out["index_sup"], out["index_sub"], out["Nt_run"] = res.index_sup, res.index_sub, res.Nt_run
    out["sensors"] = sout[:nsout]
    out["cancelled"] = rc == ECANCELLED
    if fields is not None:
        out["fields"] = fields
    return out


def _normal_fields_buffer(spec):
    nz, ny, nx = spec.matsgrid.shape
    pzm, pzp = max(5, spec.pad[4]), max(5, spec.pad[5])
    if spec.pml[4] or spec.pml[5]:
        pzm, pzp = max(5, pzm), max(5, pzp)
    Nz = nz + spec.pml[4] + spec.pml[5] + pzm + pzp
    return np.zeros((6, Nz, ny, nx))


def single_particle_dims(spec):
    nz, ny, nx = spec.matsgrid.shape
    pad = [max(10, p) for p in spec.pad]
    if max(spec.pml[0], spec.pml[1]) <= 0:
        pad[0] = pad[1] = 0
    if max(spec.pml[2], spec.pml[3]) <= 0:
        pad[2] = pad[3] = 0
    for a in range(3):
        if spec.pml[2 * a] or spec.pml[2 * a + 1]:
            pad[2 * a], pad[2 * a + 1] = max(5, pad[2 * a]), max(5, pad[2 * a + 1])
    Nx = nx + spec.pml[0] + spec.pml[1] + pad[0] + pad[1]
    Ny = ny + spec.pml[2] + spec.pml[3] + pad[2] + pad[3]
    Nz = nz + spec.pml[4] + spec.pml[5] + pad[4] + pad[5]
    return Nx, Ny, Nz


def run_single_particle(be: Backend, spec: ModeSpec, sensors=(), want_fields=False, end_flag=None):
    """Runs the single-particle driver; end_flag as run_custom."""
    keep = []
    cm = spec.to_c(keep)
    sreq = (CSensorReq * max(1, len(sensors)))(*[s.to_c() for s in sensors])
    nsout = sum(s.out_len() for s in sensors)
    sout = np.zeros(max(1, nsout))
    fields = None
    if want_fields:
        Nx, Ny, Nz = single_particle_dims(spec)
        fields = np.zeros((6, Nz, Ny, Nx))
    rc = be.fn("run_single_particle")(C.byref(cm), len(sensors), sreq, _dptr(sout),
                                       _dptr(fields) if fields is not None else None, _flag_ptr(end_flag))
    if rc not in (OK, ECANCELLED):
        be.check(rc, None)
    out = {"sensors": sout[:nsout], "cancelled": rc == ECANCELLED}
    if fields is not None:
        out["fields"] = fields
    return out


class Run:
    """A mode driver split into create / step / finish (engine only)."""

    def __init__(self, be: Backend, kind, spec: ModeSpec, sensors=()):
        self.be, self.kind, self.spec, self.sensors = be, kind, spec, list(sensors)
        self._keep = []
        self._cm = spec.to_c(self._keep)
        self._sreq = (CSensorReq * max(1, len(self.sensors)))(*[s.to_c() for s in self.sensors])
        h = C.c_void_p()
        rc = be.fn("run_create")(C.byref(h), kind, C.byref(self._cm), len(self.sensors), self._sreq)
        be.check(rc, None)
        self.h = h

    def step(self, n):
        """Advances up to n steps; False once the run was cancelled."""
        rc = self.be.fn("run_step")(self.h, n)
        if rc == ECANCELLED:
            return False
        if rc != OK:
            raise FDTDError(rc, (self.be.fn("run_error")(self.h) or b"").decode())
        return True

    def cancel(self):
        self.be.check(self.be.fn("run_cancel")(self.h), None)

    def progress(self):
        t, nt = C.c_int(), C.c_int()
        self.be.fn("run_progress")(self.h, C.byref(t), C.byref(nt))
        return t.value, nt.value

    def grid(self):
        return self.be.fn("run_grid")(self.h)

    def synchronize(self):
        self.be.check(self.be.fn("synchronize")(self.grid()), self.grid())

    def stats(self):
        out = (C.c_long * 4)()
        self.be.fn("stats")(self.grid(), out)
        return list(out)

    def layout(self):
        lay = (C.c_int * 16)()
        self.be.fn("layout")(self.grid(), lay)
        return list(lay)

    def profile(self, on, kinds=31):
        self.be.fn("profile_kinds")(self.grid(), int(kinds))
        self.be.fn("profile_enable")(self.grid(), int(on))

    def autotune(self, reps=3):
        out = (C.c_int * 2)()
        self.be.check(self.be.fn("autotune")(self.grid(), reps, out), self.grid())
        return list(out)

    def set_tiles(self, tile_e, tile_h):
        self.be.check(self.be.fn("set_tiles")(self.grid(), tile_e, tile_h), self.grid())

    def profile_read(self, kernel):
        ms, n = C.c_double(), C.c_long()
        self.be.fn("profile_read")(self.grid(), kernel.encode(), C.byref(ms), C.byref(n))
        return ms.value, n.value

    def finish(self, want_fields=False):
        nsout = sum(s.out_len() for s in self.sensors)
        sout = np.zeros(max(1, nsout))
        out = {}
        res = None
        if self.kind == 0:
            Nl = self.spec.Nl
            for k in ("lambda", "R", "T"):
                out[k] = np.zeros(Nl)
            cplx = {k: np.zeros(2 * Nl) for k in ("refx", "refy", "transx", "transy")}
            res = CNormalResult()
            res.lambda_, res.R, res.T = _dptr(out["lambda"]), _dptr(out["R"]), _dptr(out["T"])
            res.refx, res.refy = _dptr(cplx["refx"]), _dptr(cplx["refy"])
            res.transx, res.transy = _dptr(cplx["transx"]), _dptr(cplx["transy"])
        fields = None
        if want_fields:
            lay = self.layout()
            fields = np.zeros((6, lay[2], lay[1], lay[0]))
        rc = self.be.fn("run_finish")(self.h, C.byref(res) if res is not None else None, _dptr(sout),
                                       _dptr(fields) if fields is not None else None)
        if rc != OK:
            raise FDTDError(rc, (self.be.fn("run_error")(self.h) or b"").decode())
        if res is not None:
            for k, v in cplx.items():
                out[k] = v[0::2] + 1j * v[1::2]
            out["index_sup"], out["index_sub"], out["Nt_run"] = res.index_sup, res.index_sub, res.Nt_run
        out["sensors"] = sout[:nsout]
        if fields is not None:
            out["fields"] = fields
        self._finished = (res, sout)
        return out

    def write(self, directory, prefix="", names=None):
        """The result files of the finished run in the reference's formats
        (aefdtd_run_write); names: one per sensor request."""
        res, sout = self._finished
        arr = None
        if names is not None:
            arr = (C.c_char_p * len(names))(*[n.encode() for n in names])
        rc = self.be.fn("run_write")(self.h, str(directory).encode(), prefix.encode(), arr, _dptr(sout),
                                     C.byref(res) if res is not None else None)
        if rc != OK:
            raise FDTDError(rc, (self.be.fn("run_error")(self.h) or b"").decode())

    def close(self):
        if self.h:
            self.be.fn("run_destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Custom-mode runs stepped together by the engine (one launch per half step
    for all members, 64-step chunks replayed from a HIP graph)."""

    def __init__(self, be: Backend, precision=FP32):
        self.be = be
        h = C.c_void_p()
        be.check(be.fn("batch_create")(C.byref(h), precision), None)
        self.h = h
        self.runs = []

    def _c(self, rc):
        if rc != OK:
            raise FDTDError(rc, (self.be.fn("batch_error")(self.h) or b"").decode())

    def add(self, run: "Run"):
        self._c(self.be.fn("batch_add")(self.h, run.h))
        self.runs.append(run)

    def remove(self, run: "Run"):
        self._c(self.be.fn("batch_remove")(self.h, run.h))
        self.runs.remove(run)

    def step(self, chunks=1):
        self._c(self.be.fn("batch_step")(self.h, chunks))

    def running(self):
        return self.be.fn("batch_running")(self.h)

    def synchronize(self):
        self._c(self.be.fn("batch_synchronize")(self.h))

    def set_tiles(self, tile_e, tile_h):
        self._c(self.be.fn("batch_set_tiles")(self.h, tile_e, tile_h))

    def autotune(self):
        out = (C.c_int * 3)()
        self._c(self.be.fn("batch_autotune")(self.h, out))
        return list(out)

    def set_skew(self, code):
        self._c(self.be.fn("batch_set_skew")(self.h, code))

    def info(self):
        out = (C.c_int * 4)()
        self._c(self.be.fn("batch_info")(self.h, out))
        return dict(zip(("members", "running", "skewed", "skew"), list(out)))

    def profile(self, on):
        ms, n = C.c_double(), C.c_long()
        self._c(self.be.fn("batch_profile")(self.h, on, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def close(self):
        if self.h:
            self.be.fn("batch_destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED, REQ_CANCELLED = range(5)


def estimate_bytes(be: Backend, kind, spec: ModeSpec, sensors=()):
    keep = []
    cm = spec.to_c(keep)
    sreq = (CSensorReq * max(1, len(sensors)))(*[q.to_c() for q in sensors])
    out = C.c_size_t()
    be.check(be.fn("estimate_bytes")(kind, C.byref(cm), len(sensors), sreq, C.byref(out)), None)
    return out.value


class Server:
    """The engine's request server: FIFO admission under a device-memory budget,
    resident runs stepped together in 64-step rounds (continuous batching)."""

    def __init__(self, be: Backend, budget_bytes):
        self.be = be
        h = C.c_void_p()
        be.check(be.fn("server_create")(C.byref(h), int(budget_bytes)), None)
        self.h = h
        self.out_len = {}

    def submit(self, kind, spec: ModeSpec, sensors=()):
        keep = []
        cm = spec.to_c(keep)
        sreq = (CSensorReq * max(1, len(sensors)))(*[q.to_c() for q in sensors])
        rid = C.c_long()
        rc = self.be.fn("server_submit")(self.h, kind, C.byref(cm), len(sensors), sreq, C.byref(rid))
        if rc != OK:
            raise FDTDError(rc, "submit refused")
        n = sum(q.out_len() for q in sensors) + (2 * spec.Nl if kind == KIND_NORMAL else 0)
        self.out_len[rid.value] = n
        return rid.value

    def cancel(self, rid):
        return self.be.fn("server_cancel")(self.h, rid)

    def wait(self, rid, timeout=-1.0):
        st = C.c_int()
        rc = self.be.fn("server_wait")(self.h, rid, timeout, C.byref(st))
        if rc != OK:
            raise FDTDError(rc, "unknown request")
        return st.value

    def result(self, rid):
        """(outputs, times) of a finished request; times = (queue wait, admission to
        spectrum, submit to spectrum, resident bytes).  Raises for a failed one."""
        n = self.out_len.get(rid, 0)
        out = np.zeros(max(1, n))
        t = np.zeros(4)
        rc = self.be.fn("server_result")(self.h, rid, _dptr(out), n, _dptr(t))
        if rc != OK:
            raise FDTDError(rc, (self.be.fn("server_error")(self.h, rid) or b"").decode())
        return out[:n], t

    def error(self, rid):
        return (self.be.fn("server_error")(self.h, rid) or b"").decode()

    def release(self, rid):
        return self.be.fn("server_release")(self.h, rid)

    def stats(self):
        out = np.zeros(8)
        self.be.check(self.be.fn("server_stats")(self.h, _dptr(out)), None)
        keys = ("queued", "resident", "done", "failed", "bytes_resident", "peak_bytes", "budget", "rounds")
        return dict(zip(keys, out.tolist()))

    def close(self):
        if self.h:
            self.be.fn("server_destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Grid:
    """Thin object wrapper over one grid handle of a backend (engine or oracle)."""

    def __init__(self, be: Backend, N, Nt, D, Dt, mode=MODE_CUSTOM, pml=(0,) * 6, pad=(0,) * 6, precision=FP64):
        self.be = be
        h = C.c_void_p()
        p6 = (C.c_int * 6)(*pml)
        q6 = (C.c_int * 6)(*pad)
        rc = be.fn("create")(C.byref(h), N[0], N[1], N[2], Nt, D[0], D[1], D[2], Dt, mode, p6, q6, precision)
        be.check(rc, None)
        self.h = h
        lay = (C.c_int * 16)()
        be.fn("layout")(h, lay)
        self.layout = list(lay)
        self.Nx, self.Ny, self.Nz = self.layout[:3]
        self._keep = []

    def _c(self, rc):
        self.be.check(rc, self.h)

    def close(self):
        if self.h:
            self.be.fn("destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_pml(self, face, kappa=25.0, sigma=1.0, alpha=0.2):
        self._c(self.be.fn("set_pml")(self.h, face, kappa, sigma, alpha))

    def set_tapering(self, n):
        self._c(self.be.fn("set_tapering")(self.h, n))

    def set_matsgrid(self, mg):
        mg = np.ascontiguousarray(mg, dtype=np.uint32)
        nz, ny, nx = mg.shape
        self._c(self.be.fn("set_matsgrid")(self.h, mg.ctypes.data_as(C.POINTER(C.c_uint32)), nx, ny, nz))

    def set_material(self, idx, m: Material):
        keep = []
        cm = m.to_c(keep)
        self._c(self.be.fn("set_material")(self.h, idx, C.byref(cm)))

    def disable_fields(self, fields):
        a = (C.c_int * max(1, len(fields)))(*fields)
        self._c(self.be.fn("disable_fields")(self.h, a, len(fields)))

    def bootstrap(self):
        self._c(self.be.fn("bootstrap")(self.h))

    def update_E(self):
        self._c(self.be.fn("update_E")(self.h))

    def update_H(self):
        self._c(self.be.fn("update_H")(self.h))

    def step(self, n=1):
        for _ in range(n):
            self.update_E()
            self.update_H()

    def read(self, f):
        out = np.zeros((self.Nz, self.Ny, self.Nx))
        self._c(self.be.fn("read_field")(self.h, f, _dptr(out)))
        return out

    def write(self, f, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.Nz, self.Ny, self.Nx)
        self._c(self.be.fn("write_field")(self.h, f, _dptr(a)))

    def fields(self):
        return np.stack([self.read(f) for f in range(6)])

    def set_plane_injection(self, k, dEx, dEy, dHx, dHy):
        arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (dEx, dEy, dHx, dHy)]
        self._keep += arrs
        self._c(self.be.fn("set_plane_injection")(self.h, k, len(arrs[0]), *[_dptr(a) for a in arrs]))

    def set_tfsf(self, aux, box, periodic_x=False, periodic_y=False):
        self._c(self.be.fn("set_tfsf")(self.h, aux.h, *box, int(periodic_x), int(periodic_y)))

    def add_spect_poynting(self, orientation, box, lambdas):
        lam = np.ascontiguousarray(lambdas, dtype=np.float64)
        i = C.c_int()
        self._c(self.be.fn("add_spect_poynting")(self.h, orientation, *box, len(lam), _dptr(lam), C.byref(i)))
        return i.value

    def add_oscillator(self, orientation, box, lambda_min, lambda_max):
        i = C.c_int()
        self._c(self.be.fn("add_oscillator")(self.h, orientation, *box, lambda_min, lambda_max, C.byref(i)))
        return i.value

    def feed_sensors(self):
        self._c(self.be.fn("feed_sensors")(self.h))

    def add_field_block(self, box, wavelength):
        i = C.c_int()
        self._c(self.be.fn("add_field_block")(self.h, *box, wavelength, C.byref(i)))
        return i.value

    def add_field_map(self, orientation, box, wavelength, cumulative=False, magnetic=False):
        i = C.c_int()
        self._c(self.be.fn("add_field_map")(self.h, orientation, *box, wavelength, int(cumulative),
                                            int(magnetic), C.byref(i)))
        return i.value

    def field_sensor(self, fid):
        """Accumulators of a field sensor as complex (ncomp, s3, s2, s1)."""
        dims = (C.c_int * 4)()
        self._c(self.be.fn("field_sensor_read")(self.h, fid, None, dims))
        nc, s1, s2, s3 = dims
        out = np.zeros(nc * s1 * s2 * s3 * 2)
        self._c(self.be.fn("field_sensor_read")(self.h, fid, _dptr(out), dims))
        return (out[0::2] + 1j * out[1::2]).reshape(nc, s3, s2, s1)

    def field_sensor_mats(self, fid):
        dims = (C.c_int * 4)()
        self._c(self.be.fn("field_sensor_read")(self.h, fid, None, dims))
        _, s1, s2, s3 = dims
        out = np.zeros(s1 * s2 * s3, dtype=np.uint32)
        self._c(self.be.fn("field_sensor_mats")(self.h, fid, out.ctypes.data_as(C.POINTER(C.c_uint32))))
        return out.reshape(s3, s2, s1)

    def sensor_poynting(self, sid, Nl):
        out = np.zeros(Nl)
        self._c(self.be.fn("sensor_poynting")(self.h, sid, _dptr(out)))
        return out

    def sensor_spectra(self, sid):
        dims = (C.c_int * 3)()
        self._c(self.be.fn("sensor_spectra")(self.h, sid, None, dims))
        s1, s2, nl = dims
        out = np.zeros(6 * nl * s1 * s2 * 2)
        self._c(self.be.fn("sensor_spectra")(self.h, sid, _dptr(out), dims))
        v = out[0::2] + 1j * out[1::2]
        return v.reshape(6, nl, s2, s1)
