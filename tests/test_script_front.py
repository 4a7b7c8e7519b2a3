"""The script front end (aether_amd/script): Aether's simulation, structure and
material scripts run unchanged.

Lua itself (5.4.4) is not installed, so the interpreter is pinned by
known answers of the Lua 5.4 manual's semantics (operator precedence and
associativity, the integer/float split, ``//`` and ``%`` rounding towards
minus infinity, closures, varargs, numeric-for bounds).  The bindings are
pinned by the reference's own front-end test (tests/fdtd/parametric_structure.lua,
which calls fail_test() on a wrong value) and its sample structures and
material library, read where the reference tree is present; by hand-derived
grids of the test scripts under tests/scripts; and, on the GPU, by running
the test scripts on the engine and the same inputs through the oracle's
drivers (FP64 bars: R/T 1e-9, sensor spectra 1e-9 of their peak)."""
import math
import os

import numpy as np
import pytest

import oracle
from aether_amd import capi
from aether_amd.script import Front
from aether_amd.script.front import Material, Structure
from aether_amd.script.lua import Interpreter, LuaError, tostring

HERE = os.path.dirname(os.path.abspath(__file__))
SCRIPTS = os.path.join(HERE, "scripts")
REF = "/root/reference"
needs_ref = pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")


def lua(src):
    it = Interpreter()
    it.execute(src)
    return it


# ---------------------------------------------------------------- the language

def test_operators_follow_the_manual():
    it = lua("""
a = 2^3^2            -- right associative: 2^9
b = -2^2             -- ^ binds tighter than unary minus
c = 7 // 2; d = -7 // 2; e = 7 // 2.0
f = 7 % -3; g = -7 % 3; h = 5.5 % 2
i = 1 .. 2 .. 3      -- right associative, numbers converted
j = 10 / 2           -- / always gives a float
k = 3 | 5; l = 6 & 3; m = 1 << 4; n = 5 ~ 3
o = not nil and 1 or 2
p = 1 + 2 * 3 - 4 / 2
q = "10" + 5         -- string coerced by arithmetic
""")
    G = it.G
    assert G["a"] == 512.0 and G["b"] == -4.0
    assert (G["c"], G["d"], G["e"]) == (3, -4, 3.0) and isinstance(G["c"], int) and isinstance(G["e"], float)
    assert (G["f"], G["g"], G["h"]) == (-2, 2, 1.5)
    assert G["i"] == "123" and G["j"] == 5.0 and tostring(G["j"]) == "5.0"
    assert (G["k"], G["l"], G["m"], G["n"]) == (7, 2, 16, 6)
    assert G["o"] == 1 and G["p"] == 5.0 and G["q"] == 15


def test_control_flow_closures_and_tables():
    it = lua("""
local function counter() local n = 0 return function() n = n + 1 return n end end
local c1, c2 = counter(), counter()
c1(); c1(); x = c1() + c2()
s = 0; for i = 10, 1, -3 do s = s + i end          -- 10 7 4 1
fs = 0; for v = 0.5, 2 do fs = fs + v end         -- float loop 0.5 1.5
t = {3, 5, 7, n = "x", [10] = 1}; len = #t
acc = 0; for i, v in ipairs(t) do acc = acc + i * v end
keys = 0; for k, v in pairs(t) do keys = keys + 1 end
function va(...) return select("#", ...), ... end
cnt, a1, a2 = va("p", nil, "q")
w = 0; repeat local z = w; w = z + 1 until w >= 4
obj = {v = 3}; function obj:get(k) return self.v * k end; mv = obj:get(2)
str = string.format("%d:%5.1f:%s", 7, 2.25, "z") .. ("ab"):rep(2)
fl, ce = math.floor(-2.5), math.ceil(2.1)
""")
    G = it.G
    assert G["x"] == 4 and G["s"] == 22 and G["fs"] == 2.0
    assert G["len"] == 3 and G["acc"] == 3 + 10 + 21 and G["keys"] == 5
    assert (G["cnt"], G["a1"], G.get("a2")) == (3, "p", None)
    assert G["w"] == 4 and G["mv"] == 6 and G["str"] == "7:  2.2:zabab"
    assert (G["fl"], G["ce"]) == (-3, 3)


def test_number_to_string_and_errors():
    assert [tostring(v) for v in (1e15, 1e16, 0.1, 2.0, -0.0, 200e-9)] == \
        ["1e+15", "1e+16", "0.1", "2.0", "-0.0", "2e-07"]
    with pytest.raises(LuaError):
        lua("x = nil + 1")
    with pytest.raises(LuaError):
        lua("undefined_function(1)")
    with pytest.raises(LuaError):
        lua("x = 1 //")


# ---------------------------------------------------------------- bindings against the reference's files

@needs_ref
def test_reference_parametric_structure_test(oracle_be):
    """tests/fdtd/parametric_structure.lua: Lx/Ly/Lz of a structure with
    declared parameters, then after :parameter() overrides; fail_test() on a
    mismatch."""
    f = Front(be=oracle_be, dry=True)
    f.run_file(os.path.join(REF, "tests/fdtd/parametric_structure.lua"))
    assert not f.log


@needs_ref
def test_reference_sample_script_from_its_installation(oracle_be, monkeypatch):
    """samples/fdtd/scripts/nanorods_grid.lua as Aether runs it when installed:
    its material file (mat_lib/...) and structure script (samples/...) are
    found under the resources directory, as PathManager::locate_file
    (common/filehdl.cpp:292-339) searches the working directory, the script's
    directory, then the executable's and resources directories.  Dry run:
    each of the three compute() calls records its normal-incidence mode."""
    monkeypatch.delenv("AETHER_RESOURCES", raising=False)
    path = os.path.join(REF, "samples/fdtd/scripts/nanorods_grid.lua")
    with pytest.raises(LuaError, match="Au_1m_5m_Vial"):
        Front(be=oracle_be, dry=True).run_file(path)
    f = Front(be=oracle_be, dry=True, resources=[REF])
    f.run_file(path)
    assert len(f.results) == 3
    for mode, (spec, sensors) in f.results:
        assert mode.kind == 0 and mode.Nt == 20000 and mode.Nl == 2001 and mode.pml[4:] == [25, 25]
        gold = mode.materials[2]
        assert gold.eps_inf == pytest.approx(6.0031) and len(gold.drude) == 1
        assert spec.matsgrid.shape == (30, 60, 60) and int(spec.matsgrid.max()) == 2
    assert [m.polarization for m, _ in f.results] == ["TE", "TM", "TE"]
    # the same through the environment variable
    monkeypatch.setenv("AETHER_RESOURCES", REF)
    g = Front(be=oracle_be, dry=True)
    g.run_file(path)
    assert len(g.results) == 3


@needs_ref
def test_reference_sample_structures():
    s = Structure(os.path.join(REF, "samples/fdtd/structures/nanorods_grid.lua")).finalize()
    assert (s.lx, s.ly, s.lz) == (300e-9, 300e-9, 150e-9) and s.default == 0
    assert [op[0] for op in s.ops] == [capi.OP_LAYER, capi.OP_BLOCK, capi.OP_BLOCK]
    assert s.ops[1][2][:6] == [105e-9, 235e-9, 0.0, 35e-9, 50e-9, 70e-9] and s.ops[1][1] == 2
    st = Structure(os.path.join(REF, "samples/fdtd/structures/cyl_var_h_var_r_var_p.lua"))
    st.params["h"] = 80e-9
    s = st.finalize()
    assert s.lz == 200e-9 + 80e-9 and s.ops[0][0] == capi.OP_CYLINDER and s.ops[0][2][5] == 80e-9
    # the coated variant: add_coating(60, 20, 20e-9, 1e-9, 0, 3) after the rods
    s = Structure(os.path.join(REF, "samples/fdtd/structures/cyl_var_h_var_r_var_p_coated.lua")).finalize()
    assert s.ops[-1][0] == capi.OP_COATING and s.ops[-1][1] == 3
    assert s.ops[-1][2][:5] == [60.0, 20.0, 20e-9, 1e-9, 0.0]


@needs_ref
def test_reference_material_library():
    """Every dielectric-model file of mat_lib loads; Au_Rakic.lua's poles are
    its own arithmetic (wp = 9.03 eV, 1/hbar in eV s)."""
    lib = os.path.join(REF, "mat_lib")
    loaded = 0
    for name in sorted(os.listdir(lib)):
        if not name.endswith(".lua"):
            continue
        m = Material()
        try:
            m.load(os.path.join(lib, name))
        except LuaError:
            continue                           # a file calling a function the reference lacks too
        if not m.unsupported:
            m.to_capi(name)
            loaded += 1
    assert loaded >= 25
    m = Material()
    m.load(os.path.join(lib, "Au_Rakic.lua"))
    ih = 1.0 / 6.58211951440e-16
    assert m.eps_inf == 1 and len(m.drude) == 1 and len(m.lorentz) == 5
    assert m.drude[0] == (math.sqrt(0.760) * 9.03 * ih, 0.053 * ih)
    C1 = 9.03 / 0.415
    assert m.lorentz[0] == (0.024 * C1 * C1, 0.415 * ih, 0.241 * ih)


# ---------------------------------------------------------------- the test scripts, CPU side

def test_scripts_build_the_expected_inputs(oracle_be):
    f = Front(be=oracle_be, dry=True)
    f.run_file(os.path.join(SCRIPTS, "slab_normal.lua"))
    (m1, (s1, _)), (m2, (s2, _)), (m3, (s3, _)) = f.results
    assert [m.prefix for m in (m1, m2, m3)] == ["slab_TE_", "slab_TM_", "slab_thick_"]
    assert (s1.polarization, s2.polarization, s3.polarization) == (0, 1, 0)
    # 40 x 30 x 200 nm at 5 nm: 8 x 6 x 40 cells; slab from z=60 nm, Lorentz cap on top
    assert s1.matsgrid.shape == (40, 6, 8) and s1.pml == (0, 0, 0, 0, 8, 8)
    assert "".join(map(str, s1.matsgrid[:, 0, 0])) == "0" * 12 + "1" * 12 + "2" * 4 + "0" * 12
    assert "".join(map(str, s3.matsgrid[:, 0, 0])) == "0" * 12 + "1" * 16 + "2" * 4 + "0" * 8
    assert s1.materials[1].eps_inf == 1.5 * 1.5 and s1.materials[2].eps_inf == 2
    assert len(s1.materials[2].lorentz) == 1 and s1.tapering == 50 and s1.Nl == 12
    f = Front(be=oracle_be, dry=True)
    f.run_file(os.path.join(SCRIPTS, "sphere_single.lua"))
    (m, (spec, sens)), = f.results
    assert spec.matsgrid.shape == (20, 20, 20) and spec.pml == (5,) * 6
    assert (sens[0].box, sens[0].x1, sens[0].x2, sens[0].Nl) == (True, 0, 20, 10)
    assert len(spec.materials[1].critpoint) == 2 and spec.materials[1].critpoint[1][0] == 1.6
    f = Front(be=oracle_be, dry=True)
    f.run_file(os.path.join(SCRIPTS, "osc_custom.lua"))
    (m, (spec, sens)), = f.results
    src = spec.sources[0]
    assert (src.orientation, src.x1, src.x2, src.z1, src.z2) == (capi.NORMAL_X, 0, 20, 2, 3)
    mg = spec.matsgrid
    assert mg.shape == (16, 18, 20) and int(mg.sum()) == 4 * 7 * 7      # 20 x 35 x 35 nm block at 5 nm


def test_unprovided_mode_and_operations_fail_loudly(oracle_be, tmp_path):
    p = tmp_path / "s.lua"
    p.write_text('f = MODE("fdfd")\n')
    with pytest.raises(LuaError, match="only fdtd_normal"):
        Front(be=oracle_be, dry=True).run_file(str(p))
    q = tmp_path / "st.lua"
    q.write_text('lx=1e-7 ly=1e-7 lz=1e-7\nadd_sin_layer(0,1,2,3,4,5,6,7,1)\n')
    with pytest.raises(LuaError, match="not provided"):
        Structure(str(q)).finalize()


# ---------------------------------------------------------------- GPU: engine runs of the scripts

def _runs(script, tmp_path):
    out = tmp_path / "out"
    eng = Front(be=capi.engine(), out_dir=str(out))
    eng.run_file(os.path.join(SCRIPTS, script))
    dry = Front(be=oracle.load(), dry=True)
    dry.run_file(os.path.join(SCRIPTS, script))
    assert len(eng.results) == len(dry.results)
    return out, zip(eng.results, dry.results)


@pytest.mark.gpu
def test_script_normal_incidence_fp64(tmp_path):
    print("script front: slab_normal.lua", flush=True)
    out, pairs = _runs("slab_normal.lua", tmp_path)
    Rs = []
    for (m, g), (_, (spec, sens)) in pairs:
        r = capi.run_normal_incidence(oracle.load(), spec, sens)
        for k in ("R", "T"):
            assert np.max(np.abs(g[k] - r[k])) <= 1e-9, (m.prefix, k)
        assert (out / (m.prefix + "spectdata_norm")).is_file()
        assert (out / (m.prefix + "spectdata_norm2")).is_file()
        Rs.append(g["R"])
    assert np.max(np.abs(Rs[0] - Rs[2])) > 1e-3          # the thicker slab reflects differently


@pytest.mark.gpu
def test_script_spectral_auto_tsteps_fp64(tmp_path):
    print("script front: slab_auto_ft.lua", flush=True)
    out, pairs = _runs("slab_auto_ft.lua", tmp_path)
    for (m, g), (_, (spec, sens)) in pairs:
        assert (spec.time_type, spec.cc_quant, spec.cc_layout, spec.cc_seed) == (2, 60, "nnb", 5)
        r = capi.run_normal_incidence(oracle.load(), spec, sens)
        assert g["steps"] == oracle.last_steps() < spec.Nt and g["Nt_run"] == r["Nt_run"]
        for k in ("R", "T"):
            assert np.max(np.abs(g[k] - r[k])) <= 1e-9, k


@pytest.mark.gpu
def test_script_auto_tsteps_fp64(tmp_path):
    print("script front: slab_auto.lua", flush=True)
    out, pairs = _runs("slab_auto.lua", tmp_path)
    for (m, g), (_, (spec, sens)) in pairs:
        assert (spec.time_type, spec.cc_step, spec.polarization, spec.polar_angle) == (1, 10, 2, 30.0)
        r = capi.run_normal_incidence(oracle.load(), spec, sens)
        assert g["steps"] == oracle.last_steps() < spec.Nt and g["Nt_run"] == r["Nt_run"]
        for k in ("R", "T"):
            assert np.max(np.abs(g[k] - r[k])) <= 1e-9, k


@pytest.mark.gpu
def test_script_single_particle_fp64(tmp_path):
    print("script front: sphere_single.lua", flush=True)
    out, pairs = _runs("sphere_single.lua", tmp_path)
    for (m, g), (_, (spec, sens)) in pairs:
        r = capi.run_single_particle(oracle.load(), spec, sens)
        assert np.max(np.abs(g["sensors"] - r["sensors"])) <= 1e-9 * np.max(np.abs(r["sensors"]))
        assert (out / "box_bspdft").is_file() and (out / "box.m").is_file()


@pytest.mark.gpu
def test_script_custom_oscillator_fp64(tmp_path):
    print("script front: osc_custom.lua", flush=True)
    out, pairs = _runs("osc_custom.lua", tmp_path)
    for (m, g), (_, (spec, sens)) in pairs:
        r = capi.run_custom(oracle.load(), spec, sens)
        assert np.max(np.abs(g["sensors"] - r["sensors"])) <= 1e-9 * np.max(np.abs(r["sensors"]))
        assert (out / "plane_pspdft").is_file()
