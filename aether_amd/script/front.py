"""Aether's script front end over the MI355X engine.

Runs Aether simulation scripts (``MODE("fdtd_normal")`` ... ``:compute()``),
structure scripts and material library files unchanged, with the interpreter
of :mod:`aether_amd.script.lua`, and hands the FDTD modes to the engine's
drivers through the C ABI.  The objects and their script methods follow
Aether's Lua bindings:

- ``MODE`` (src/base.cpp:75-136, FDTD types only) and the FDTD mode methods of
  ``FDTD_Mode_create_metatable`` (finite_differences/FDTD/modes/
  lua_fdtd.cpp:363-411) with the FD_Mode setters and defaults of
  finite_differences/core/lua_fd.cpp:33-56,117-206,476-672;
- ``Material`` and the material-file functions (materials/lua_material.cpp:
  26-58,60-494);
- ``Structure(path)`` with ``:parameter`` and the structure-script functions
  (structure/lua_structure.cpp:210-282,292-690, structure/structure.cpp:
  140-221); the structure is discretised on the device (aefdtd_discretize);
- ``create_sensor`` / ``create_source`` and their methods
  (sensors/lua_sensors.cpp:26-147, sources/lua_sources.cpp:26-140).

``:compute()`` runs what FDTD_Mode::process runs for the three provided mode
types ("fdtd_normal", "fdtd_single_particle", "fdtd"): compute_discretization,
Structure::discretize, the sensors' and sources' to_discrete, the sizing of
FDTD_Mode::finalize, the driver, and the result files.  Methods of modes,
sensors and materials the engine does not provide raise a LuaError naming
them; nothing falls back to another solver.
"""
from __future__ import annotations

import copy
import math
import os

from .. import capi
from .lua import Interpreter, LuaError, to_integer, to_number, tostring

MODE_KINDS = {"fdtd_normal": capi.KIND_NORMAL, "fdtd_single_particle": capi.KIND_SINGLE, "fdtd": capi.KIND_CUSTOM}


def nearest_integer(x):
    """common/mathUT.cpp:1169-1173: rounds halves away from zero."""
    x = float(x)
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def _num(v, what):
    n = to_number(v)
    if n is None:
        raise LuaError(f"{what}: number expected, got {tostring(v)}")
    return float(n)


def _int(v, what):
    n = to_number(v)
    if n is None:
        raise LuaError(f"{what}: integer expected, got {tostring(v)}")
    i = to_integer(n)
    return i if i is not None else int(n)   # lua_tointeger truncates a float argument through the C API


# Aether's executable and installed-resources directories (where its mat_lib
# lives) for the script being run: Front(resources=...) or, by default, the
# AETHER_RESOURCES variable (os.pathsep-separated).
_resources = []


def _resource_dirs():
    env = os.environ.get("AETHER_RESOURCES", "")
    return _resources + [d for d in env.split(os.pathsep) if d]


def locate(path, caller_dir):
    """PathManager::locate_file (common/filehdl.cpp:292-339): an absolute path
    as given; a relative one from the working directory, else from the calling
    script's directory, else from the executable and resources directories."""
    if os.path.isabs(path) or os.path.exists(path):
        return path
    dirs = [caller_dir] + _resource_dirs()
    for base in dirs:
        p = os.path.join(base, path)
        if os.path.exists(p):
            return p
    raise LuaError(f"file not found: {path} (searched the working directory, {', '.join(dirs)})")


def _orientation(s):
    """is_x_pos / is_x_neg ... of common/string_tools.cpp:253-300."""
    s = tostring(s)
    for axis, pos, neg in (("x", capi.NORMAL_X, capi.NORMAL_XM), ("y", capi.NORMAL_Y, capi.NORMAL_YM),
                           ("z", capi.NORMAL_Z, capi.NORMAL_ZM)):
        if s in (axis, axis.upper(), "+" + axis, "+" + axis.upper(), axis + "+", axis.upper() + "+"):
            return pos
        if s in ("-" + axis, "-" + axis.upper(), axis + "-", axis.upper() + "-"):
            return neg
    return None


# ---------------------------------------------------------------- materials

class Material:
    """Material with the dielectric models the FDTD path uses (eps_inf and
    Debye, Drude, Lorentz, critical-point poles)."""

    def __init__(self):
        self.reset()

    def reset(self):
        self.eps_inf = 1.0
        self.debye, self.drude, self.lorentz, self.critpoint = [], [], [], []
        self.name, self.description, self.script_path = "", "", ""
        self.validity = (0.0, 0.0)
        self.unsupported = []

    def load(self, path):
        """lua_material::Loader::load (lua_material.cpp:455-487): reset, then run
        the file with the material functions bound to this material."""
        self.reset()
        self.script_path = path
        it = Interpreter(chunk=os.path.basename(path))
        for name, fn in MATERIAL_FUNCTIONS.items():
            it.G[name] = (lambda f: (lambda *a: f(self, *a)))(fn)
        it.G["add_effective_component"] = lambda *a: self.unsupported.append("add_effective_component")
        self.caller_dir = os.path.dirname(os.path.abspath(path))
        with open(path) as fh:
            it.execute(fh.read())

    def to_capi(self, what):
        if self.unsupported:
            raise LuaError(f"{what}: the FDTD engine takes dielectric models only; "
                           f"{', '.join(sorted(set(self.unsupported)))} is not provided")
        return capi.Material(eps_inf=self.eps_inf, debye=list(self.debye), drude=list(self.drude),
                             lorentz=list(self.lorentz), critpoint=list(self.critpoint))


def _mat_set(attr):
    def f(m, v):
        setattr(m, attr, tostring(v))
    return f


def _unsupported(name):
    def f(m, *a):
        m.unsupported.append(name)
    return f


MATERIAL_FUNCTIONS = {
    "add_crit_point": lambda m, A, O, P, G: m.critpoint.append(tuple(_num(x, "add_crit_point") for x in (A, O, P, G))),
    "add_debye": lambda m, ds, t0: m.debye.append((_num(ds, "add_debye"), _num(t0, "add_debye"))),
    "add_drude": lambda m, wd, g: m.drude.append((_num(wd, "add_drude"), _num(g, "add_drude"))),
    "add_lorentz": lambda m, A, O, G: m.lorentz.append(tuple(_num(x, "add_lorentz") for x in (A, O, G))),
    "description": _mat_set("description"),
    "epsilon_infinity": lambda m, e: setattr(m, "eps_inf", _num(e, "epsilon_infinity")),
    "epsilon_infty": lambda m, e: setattr(m, "eps_inf", _num(e, "epsilon_infty")),
    "index_infty": lambda m, n: setattr(m, "eps_inf", _num(n, "index_infty") * _num(n, "index_infty")),
    "name": _mat_set("name"),
    "refractive_index": lambda m, n: setattr(m, "eps_inf", _num(n, "refractive_index") * _num(n, "refractive_index")),
    "validity_range": lambda m, a, b: setattr(m, "validity", (_num(a, "validity_range"), _num(b, "validity_range"))),
    "set_dielectric": lambda m, *a: None,
    "add_cauchy": _unsupported("add_cauchy"),
    "add_sellmeier": _unsupported("add_sellmeier"),
    "add_data_epsilon": _unsupported("add_data_epsilon"),
    "add_data_index": _unsupported("add_data_index"),
    "effective_host": _unsupported("effective_host"),
    "effective_type": _unsupported("effective_type"),
}
MATERIAL_FUNCTIONS["index_infinity"] = MATERIAL_FUNCTIONS["index_infty"]


def _material_methods():
    meth = {name: (lambda f: (lambda self, *a: f(self, *a)))(fn) for name, fn in MATERIAL_FUNCTIONS.items()
            if name != "set_dielectric"}

    def load_script(self, path):
        self.load(locate(tostring(path), self.caller_dir))
    meth["load_script"] = load_script
    return meth


Material.lua_methods = _material_methods()


# ---------------------------------------------------------------- structures

class Structure:
    """Structure(path): a structure script plus the parameters set on it."""

    lua_methods = {}

    def __init__(self, path):
        self.path = path
        self.params = {}        # name -> value, in the order first seen
        self.built = None

    def finalize(self):
        """Structure::finalize (structure.cpp:140-221): runs the script with the
        structure functions bound, then reads lx, ly, lz."""
        s = capi.Structure()
        it = Interpreter(chunk=os.path.basename(self.path))
        params = self.params

        def declare_parameter(name, default):
            name = tostring(name)
            if name not in params:
                params[name] = _num(default, "declare_parameter")
            it.G[name] = params[name]
            return params[name]

        def op(fn, n_num, what):
            def f(*a):
                if len(a) < n_num + 1:
                    raise LuaError(f"{what}: {n_num + 1} arguments expected")
                fn(*[_num(x, what) for x in a[:n_num]], _int(a[n_num], what))
            return f

        def add_layer(axis, z1, z2, index):
            s.add_layer(tostring(axis), _num(z1, "add_layer"), _num(z2, "add_layer"), _int(index, "add_layer"))

        def add_coating(*a):
            # lua_structure.cpp:314-331: four numbers, then two integers
            if len(a) < 6:
                raise LuaError("add_coating: 6 arguments expected")
            s.add_coating(*[_num(x, "add_coating") for x in a[:4]], _int(a[4], "add_coating"),
                          _int(a[5], "add_coating"))

        G = it.G
        G.update(add_block=op(s.add_block, 6, "add_block"), add_sphere=op(s.add_sphere, 4, "add_sphere"),
                 add_cylinder=op(s.add_cylinder, 7, "add_cylinder"), add_cone=op(s.add_cone, 7, "add_cone"),
                 add_ellipsoid=op(s.add_ellipsoid, 6, "add_ellipsoid"), add_layer=add_layer,
                 add_coating=add_coating,
                 default_material=lambda i: s.default_material(_int(i, "default_material")),
                 flip=lambda x, y, z: s.flip(_int(x, "flip"), _int(y, "flip"), _int(z, "flip")),
                 loop=lambda x, y, z: s.loop(_int(x, "loop"), _int(y, "loop"), _int(z, "loop")),
                 declare_parameter=declare_parameter, nearest_integer=nearest_integer)
        for name in ("add_conformal_coating", "add_lua_def", "add_mesh", "add_sin_layer",
                     "add_vect_block", "add_vect_tri", "random_packing", "trapz"):
            G[name] = (lambda n: (lambda *a: (_ for _ in ()).throw(
                LuaError(f"{n}: this structure operation is not provided by the device discretiser"))))(name)
        with open(self.path) as fh:
            it.execute(fh.read())
        s.lx = _num(G.get("lx"), "lx")
        s.ly = _num(G.get("ly"), "ly")
        s.lz = _num(G.get("lz"), "lz")
        self.built = s
        return s


def _structure_parameter(self, name, value):
    self.params[tostring(name)] = _num(value, "parameter")


Structure.lua_methods = {"parameter": _structure_parameter, "finalize": lambda self: self.finalize()}


# ---------------------------------------------------------------- sensors and sources

_SENSOR_KINDS = {"box_spectral_poynting": capi.SENSOR_POYNTING, "planar_spectral_poynting": capi.SENSOR_POYNTING,
                 "fieldblock": capi.SENSOR_FIELDBLOCK, "fieldmap": capi.SENSOR_FIELDMAP,
                 "fieldmap2": capi.SENSOR_FIELDMAP2}


class SensorGen:
    """Sensor_generator (sensors/sensors.cpp:20-161) for the spectral Poynting
    and the field sensors (Sensor_type names, lua_sensors.cpp:126-147)."""

    def __init__(self, kind):
        self.kind = kind
        self.name = "sensor"
        self.orientation = capi.NORMAL_Z
        self.Nl, self.lambda_min, self.lambda_max = 481, 470e-9, 850e-9
        self.real = False
        self.g = [0, 0, 0, 0, 0, 0]
        self.r = [0.0] * 6
        self.disable = [0] * 6

    def to_req(self, D):
        """Sensor_generator::to_discrete (sensors.cpp:142-161)."""
        if self.kind not in _SENSOR_KINDS:
            raise LuaError(f"sensor type {self.kind} is not provided by the engine")
        g = list(self.g)
        if self.real:
            for a in range(3):
                lo = nearest_integer(self.r[2 * a] / D[a])
                hi = nearest_integer(self.r[2 * a + 1] / D[a])
                if hi <= lo:
                    hi = lo + 1
                g[2 * a], g[2 * a + 1] = lo, hi
        return capi.SensorReq(box=self.kind == "box_spectral_poynting", orientation=self.orientation,
                              x1=g[0], x2=g[1], y1=g[2], y2=g[3], z1=g[4], z2=g[5], disable=tuple(self.disable),
                              Nl=self.Nl, lambda_min=self.lambda_min, lambda_max=self.lambda_max,
                              kind=_SENSOR_KINDS[self.kind])


def _sensor_location(real):
    def f(self, *a):
        self.real = real
        if len(a) == 3:
            if real:
                self.r[0::2] = [_num(v, "location") for v in a]
            else:
                self.g[0::2] = [_int(v, "location_grid") for v in a]
        elif len(a) == 6:
            if real:
                self.r = [_num(v, "location") for v in a]
            else:
                self.g = [_int(v, "location_grid") for v in a]
        return self
    return f


def _sensor_disable(self, d):
    o = _orientation(d)
    idx = {capi.NORMAL_XM: 0, capi.NORMAL_X: 1, capi.NORMAL_YM: 2, capi.NORMAL_Y: 3,
           capi.NORMAL_ZM: 4, capi.NORMAL_Z: 5}.get(o)
    if idx is not None:
        self.disable[idx] = 1


def _sensor_orientation(self, d):
    o = _orientation(d)
    if o is not None:
        self.orientation = o


def _sensor_spectrum(self, lmin, lmax, Nl):
    self.lambda_min, self.lambda_max, self.Nl = _num(lmin, "spectrum"), _num(lmax, "spectrum"), _int(Nl, "spectrum")


def _sensor_wavelength(self, lam):
    """Sensor_generator::set_wavelength (sensors.cpp:135-140): the analysis
    wavelength of fieldblock / fieldmap sensors."""
    self.lambda_min = self.lambda_max = _num(lam, "wavelength")


SensorGen.lua_methods = {
    "location": _sensor_location(True), "location_grid": _sensor_location(False), "disable": _sensor_disable,
    "name": lambda self, n: setattr(self, "name", tostring(n)), "orientation": _sensor_orientation,
    "spectrum": _sensor_spectrum, "wavelength": _sensor_wavelength,
}


class SourceGen:
    """Source_generator (sources/sources.cpp:20-101), oscillator type."""

    def __init__(self, kind):
        self.kind = kind
        self.orientation = capi.NORMAL_Z
        self.lambda_min, self.lambda_max = 370e-9, 850e-9
        self.real = False
        self.g = [0, 1, 0, 1, 0, 1]
        self.r = [0.0] * 6

    def to_req(self, D):
        """Source_generator::to_discrete (sources.cpp:84-101)."""
        if self.kind != "oscillator":
            raise LuaError(f"source type {self.kind} is not provided by the engine")
        g = list(self.g)
        if self.real:
            for a in range(3):
                lo = nearest_integer(self.r[2 * a] / D[a])
                hi = nearest_integer(self.r[2 * a + 1] / D[a])
                if hi <= lo:
                    hi = lo + 1
                g[2 * a], g[2 * a + 1] = lo, hi
        return capi.SourceReq(orientation=self.orientation, x1=g[0], x2=g[1], y1=g[2], y2=g[3], z1=g[4], z2=g[5],
                              lambda_min=self.lambda_min, lambda_max=self.lambda_max)


def _source_location(real):
    def f(self, *a):
        """source_set_location{,_real} (lua_sources.cpp:51-108): a point is
        one cell wide on the grid, zero wide in metres (widened by to_discrete)."""
        self.real = real
        if len(a) == 3:
            if real:
                v = [_num(x, "location") for x in a]
                self.r = [v[0], v[0], v[1], v[1], v[2], v[2]]
            else:
                v = [_int(x, "location_grid") for x in a]
                self.g = [v[0], v[0] + 1, v[1], v[1] + 1, v[2], v[2] + 1]
        elif len(a) == 6:
            if real:
                self.r = [_num(x, "location") for x in a]
            else:
                self.g = [_int(x, "location_grid") for x in a]
        return self
    return f


def _source_spectrum(self, lmin, lmax):
    self.lambda_min, self.lambda_max = _num(lmin, "spectrum"), _num(lmax, "spectrum")


SourceGen.lua_methods = {"location": _source_location(True), "location_grid": _source_location(False),
                         "orientation": _sensor_orientation, "spectrum": _source_spectrum}


# ---------------------------------------------------------------- modes

class FDTDMode:
    """FDTD_Mode with the defaults of FD_Mode::FD_Mode (lua_fd.cpp:33-56) and
    FDTD_Mode::reset (lua_fdtd.cpp:234-266)."""

    def __init__(self, name, front):
        self.kind = MODE_KINDS[name]
        self.front = front
        self.D = [5e-9, 5e-9, 5e-9]
        self.Nt, self.tapering, self.time_mod = 5000, 0, 1.0
        self.time_type, self.cc_step, self.cc_coeff = 0, 500, 1e-3
        self.cc_lmin, self.cc_lmax, self.cc_quant, self.cc_layout = 370e-9, 850e-9, 500, "nnb"
        self.Nl, self.lambda_min, self.lambda_max = 481, 370e-9, 850e-9
        self.polarization, self.polar_angle = "TE", 0.0
        self.prefix, self.directory = "", ""
        self.pml = [0] * 6
        self.pad = [0] * 6
        self.kappa, self.sigma, self.alpha = [25.0] * 6, [1.0] * 6, [0.2] * 6
        self.materials = {}
        self.disable = []
        self.sensors, self.sources = [], []
        self.structure = None
        self.precision = front.precision
        self.result = None

    def dims(self):
        s = self.structure.built
        return s.cells(*self.D)

    def spec(self):
        """ModeSpec of the mode: compute_discretization and the discretised
        structure (on the device), the material list, sensors and sources
        taken to the grid, then FDTD_Mode::finalize's sizing."""
        if self.structure is None:
            raise LuaError("compute: no structure set")
        s = self.structure.built or self.structure.finalize()
        N = s.cells(*self.D)
        mg = capi.discretize(self.front.be, s, N, self.D)
        nmat = int(mg.max()) + 1
        mats = []
        for i in range(nmat):
            m = self.materials.get(i)
            mats.append(m.to_capi(f"material {i}") if m is not None else capi.Material(1.0))
        pol = {"TE": 0, "TM": 1, "mix": 2}[self.polarization]
        sensors = [q.to_req(self.D) for q in self.sensors]
        sources = [q.to_req(self.D) for q in self.sources]
        spec = capi.ModeSpec(matsgrid=mg, materials=mats, Nt=self.Nt, Dx=self.D[0], Dy=self.D[1], Dz=self.D[2],
                             time_mod=self.time_mod, tapering=self.tapering, pml=tuple(self.pml),
                             pad=tuple(self.pad), kappa=tuple(self.kappa), sigma=tuple(self.sigma),
                             alpha=tuple(self.alpha), Nl=self.Nl, lambda_min=self.lambda_min,
                             lambda_max=self.lambda_max, polarization=pol, polar_angle=self.polar_angle,
                             disable_fields=tuple(self.disable), precision=self.precision, sources=sources,
                             time_type=self.time_type, cc_step=self.cc_step, cc_coeff=self.cc_coeff,
                             cc_lmin=self.cc_lmin, cc_lmax=self.cc_lmax, cc_quant=self.cc_quant,
                             cc_layout=self.cc_layout, cc_seed=self.front.seed)
        spec = capi.mode_finalize(self.front.be, spec, self.kind, sensors)
        return spec, sensors

    def compute(self):
        if self.front.dry:
            # the inputs as they stand now (later script lines may change the structure)
            self.front.results.append((self, self.spec()))
            return
        spec, sensors = self.spec()
        run = capi.Run(self.front.be, self.kind, spec, sensors)
        try:
            t, nt = run.progress()
            while t < nt:
                run.step(64)
                t, nt = run.progress()
            out = run.finish()
            directory = self.directory or self.front.out_dir
            os.makedirs(directory, exist_ok=True)
            run.write(directory, self.prefix, [q.name for q in self.sensors])
        finally:
            run.close()
        out["steps"] = t
        self.result = out
        self.front.results.append((self, out))


def _pml(face):
    def f(self, N, kap, sig, alp):
        """FD_Mode::set_pml_* (lua_fd.cpp:154-206): also pads the face to 5 cells."""
        self.pml[face] = _int(N, "pml")
        self.kappa[face], self.sigma[face], self.alpha[face] = (_num(kap, "pml"), _num(sig, "pml"),
                                                                _num(alp, "pml"))
        self.pad[face] = max(self.pad[face], 5)
    return f


def _set_D(axes):
    def f(self, d):
        for a in axes:
            self.D[a] = _num(d, "D")
    return f


def _polarization(self, p):
    """FD_mode_set_polarization (lua_fd.cpp:619-644): a number sets "mix" at
    that angle; strings other than te/tm variants leave TE."""
    if to_number(p) is not None and not isinstance(p, str):
        self.polar_angle = _num(p, "polarization")
        self.polarization = "mix"
        return
    t = tostring(p)
    pol = "TE"
    if t in ("tm", "TM", "Tm", "tM"):
        pol = "TM"
    self.polarization = pol


def _prefix(self, p):
    p = tostring(p)
    if not p.endswith("_"):
        p += "_"
    self.prefix = p


def _material(self, index, mat):
    i = _int(index, "material")
    if i < 0:
        return
    if isinstance(mat, str):
        m = Material()
        m.caller_dir = self.front.caller_dir
        m.load(locate(mat, self.front.caller_dir))
        self.materials[i] = m
    elif isinstance(mat, Material):
        self.materials[i] = copy.deepcopy(mat)     # FD_Mode::set_material copies (lua_fd.cpp:121-127)
    else:
        raise LuaError("material: a Material or a material file expected")


def _disable_fields(self, *names):
    ids = {"ex": capi.EX, "ey": capi.EY, "ez": capi.EZ, "hx": capi.HX, "hy": capi.HY, "hz": capi.HZ}
    for n in names:
        f = ids.get(tostring(n).lower())
        if f is not None:
            self.disable.append(f)


def _spectrum(self, lmin, lmax, Nl=None):
    self.lambda_min, self.lambda_max = _num(lmin, "spectrum"), _num(lmax, "spectrum")
    if Nl is not None:
        self.Nl = _int(Nl, "spectrum")


def _auto_tsteps(self, Nt, step, *rest):
    """FDTD_mode_set_auto_tsteps (lua_fdtd.cpp:440-466): (Nt, step, coeff) selects
    the energy criterion; (Nt, step, lmin, lmax, coeff, quant [, layout]) the
    spectral one, whose cells are drawn from the generator seeded by rng_seed."""
    self.Nt, self.cc_step = _int(Nt, "auto_tsteps"), _int(step, "auto_tsteps")
    if len(rest) == 1:
        self.cc_coeff = _num(rest[0], "auto_tsteps")
        self.time_type = 1
    elif len(rest) in (4, 5):
        self.cc_lmin, self.cc_lmax = _num(rest[0], "auto_tsteps"), _num(rest[1], "auto_tsteps")
        self.cc_coeff, self.cc_quant = _num(rest[2], "auto_tsteps"), _int(rest[3], "auto_tsteps")
        self.cc_layout = tostring(rest[4]) if len(rest) == 5 else "nnb"
        self.time_type = 2
    else:
        raise LuaError("auto_tsteps: 3, 6 or 7 arguments expected")


def _padding(self, *p):
    self.pad = [_int(v, "padding") for v in p[:6]]


def _length(axis):
    def f(self):
        """FD_Mode::get_Lx (lua_fd.cpp:69-92): N*D after compute_discretization."""
        if self.structure is None:
            raise LuaError("no structure set")
        s = self.structure.built or self.structure.finalize()
        return s.cells(*self.D)[axis] * self.D[axis]
    return f


def _count(axis):
    def f(self):
        s = self.structure.built or self.structure.finalize()
        return s.cells(*self.D)[axis]
    return f


def _structure(self, st):
    """FD_Mode::set_structure (lua_fd.cpp:465-469): finalizes the structure now."""
    if not isinstance(st, Structure):
        raise LuaError("structure: a Structure expected")
    self.structure = st
    st.finalize()


FDTDMode.lua_methods = {
    "auto_tsteps": _auto_tsteps, "compute": lambda self: self.compute(),
    "display_step": lambda self, *a: None,
    "Dx": _set_D((0,)), "Dy": _set_D((1,)), "Dz": _set_D((2,)), "Dxyz": _set_D((0, 1, 2)),
    "Lx": _length(0), "Ly": _length(1), "Lz": _length(2), "Nx": _count(0), "Ny": _count(1), "Nz": _count(2),
    "material": _material, "N_tsteps": lambda self, n: setattr(self, "Nt", _int(n, "N_tsteps")),
    "output_directory": lambda self, d: setattr(self, "directory", tostring(d)),
    "padding": _padding,
    "pml_xm": _pml(0), "pml_xp": _pml(1), "pml_ym": _pml(2), "pml_yp": _pml(3), "pml_zm": _pml(4), "pml_zp": _pml(5),
    "polarization": _polarization, "prefix": _prefix, "structure": _structure,
    "tapering": lambda self, n: setattr(self, "tapering", _int(n, "tapering")),
    "time_mod": lambda self, v: setattr(self, "time_mod", _num(v, "time_mod")),
    "disable_fields": _disable_fields,
    "register_sensor": lambda self, s: self.sensors.append(s),
    "register_source": lambda self, s: self.sources.append(s),
    "spectrum": _spectrum,
}


class TestFailure(LuaError):
    pass


class Front:
    """One script run: the global functions of src/base.cpp:203-282 that the
    FDTD path needs, bound to the engine backend `be` (or the oracle)."""

    def __init__(self, be=None, out_dir=".", precision=capi.FP64, dry=False, resources=()):
        self.be = be if be is not None else capi.engine()
        self.dry = dry                      # compute() records the mode without running it
        self.resources = list(resources)    # Aether's executable / share/Aether directories
        self.out_dir = out_dir
        self.precision = precision
        self.results = []
        self.caller_dir = "."
        self.log = []
        self.seed = 0               # rng_seed(n); the reference seeds its generator with the time

    def run_file(self, path):
        self.caller_dir = os.path.dirname(os.path.abspath(path))
        with open(path) as fh:
            return self.run(fh.read(), chunk=os.path.basename(path))

    def run(self, source, chunk="script"):
        global _resources
        saved, _resources = _resources, self.resources
        try:
            return self._run(source, chunk)
        finally:
            _resources = saved

    def _run(self, source, chunk):
        it = Interpreter(chunk=chunk, out=self.log)
        front = self

        def mode(name):
            name = tostring(name)
            if name not in MODE_KINDS:
                raise LuaError(f"MODE(\"{name}\"): only fdtd_normal, fdtd_single_particle and fdtd are provided")
            return FDTDMode(name, front)

        def material(arg=None):
            """lua_material::allocate (lua_material.cpp:60-74): a number sets a
            constant index, a string loads a material file."""
            m = Material()
            m.caller_dir = front.caller_dir
            if arg is not None:
                if isinstance(arg, str):
                    m.load(locate(arg, front.caller_dir))
                else:
                    n = _num(arg, "Material")
                    m.eps_inf = n * n
            return m

        def structure(path):
            return Structure(locate(tostring(path), front.caller_dir))

        def const_material(n):
            m = Material()
            v = _num(n, "const_material")
            m.eps_inf = v * v
            return m

        def fail_test(*a):
            raise TestFailure("fail_test() called by the script")

        def rng_seed(n):
            front.seed = _int(n, "rng_seed") & 0xFFFFFFFF

        it.G.update(MODE=mode, Material=material, Structure=structure, create_sensor=lambda t: SensorGen(tostring(t)),
                    create_source=lambda t: SourceGen(tostring(t)), nearest_integer=nearest_integer,
                    fail_test=fail_test, rng_seed=rng_seed)
        # const_material returns a material script path in the reference; here the
        # material itself, which FDTD_Mode:material accepts the same way
        it.G["const_material"] = const_material
        it.execute(source)
        return it


def run_script(path, be=None, out_dir=".", precision=capi.FP64, resources=()):
    """Runs an Aether simulation script; returns the Front (results, log)."""
    f = Front(be, out_dir, precision, resources=resources)
    f.run_file(path)
    return f


__all__ = ["Front", "run_script", "Material", "Structure", "FDTDMode", "nearest_integer"]
