"""GPU parity: the HIP engine (FP64) against the CPU oracle on the same seeded
inputs.  Bar: <= 1e-10 relative L2 per field, <= 1e-9 absolute on spectra
(normalised spectra R, T, and Poynting powers relative to their scale).
FP32 storage: spectra within 1e-4 of the FP64 oracle (relative to the
spectrum's peak)."""
import numpy as np
import pytest

from aether_amd import capi, workloads
from aether_amd.capi import Grid, Material

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-10
GOLD = workloads.GOLD
LORENTZ = Material(eps_inf=1.3, lorentz=[(1.5, 3e15, 1e14)])
DEBYE = Material(eps_inf=2.0, debye=[(1.2, 3e-15)])


def rel_l2(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def build_pair(be_e, be_o, n, pml, pad, mats, mg, seed, Nt=50, D=(5e-9, 5e-9, 5e-9), disable=(),
               precision=capi.FP64, kap=(25, 25, 25, 25, 25, 25), sig=(1,) * 6, alp=(0.2,) * 6):
    Dt = min(D) / (np.sqrt(3.0) * 299792458.0) * 0.99
    grids = []
    for be, prec in ((be_e, precision), (be_o, capi.FP64)):
        g = Grid(be, n, Nt, D, Dt, capi.MODE_CUSTOM, pml, pad, prec)
        for f in range(6):
            g.set_pml(f, kap[f], sig[f], alp[f])
        g.set_matsgrid(mg)
        for i, m in enumerate(mats):
            g.set_material(i, m)
        if disable:
            g.disable_fields(list(disable))
        g.bootstrap()
        grids.append(g)
    rng = np.random.default_rng(seed)
    e, o = grids
    for f in range(6):
        if f in disable:
            continue
        scale = 1.0 if f < 3 else 1.0 / 376.73
        a = rng.standard_normal((e.Nz, e.Ny, e.Nx)) * scale
        e.write(f, a)
        o.write(f, a)
    return e, o


def compare(e, o, tol=FIELD_TOL):
    for f in range(6):
        a, b = e.read(f), o.read(f)
        err = rel_l2(a, b)
        assert err <= tol, (f, err)


CASES = [
    # name, structure n, pml, pad, disable
    ("vacuum_periodic", (12, 10, 9), (0,) * 6, (0,) * 6, ()),
    ("cpml_all_faces", (10, 11, 12), (4, 5, 3, 6, 4, 4), (0,) * 6, ()),
    ("cpml_z_only", (9, 8, 14), (0, 0, 0, 0, 5, 3), (0, 0, 0, 0, 2, 1), ()),
    ("kappa_quirk_xp_only", (10, 9, 8), (0, 5, 0, 0, 0, 0), (0,) * 6, ()),
    ("kappa_quirk_yp_only", (10, 9, 8), (0, 0, 0, 5, 0, 0), (0,) * 6, ()),
    ("tm_2d", (20, 18, 1), (5, 5, 4, 4, 0, 0), (0,) * 6, (0, 1, 5)),
    ("te_2d", (20, 18, 1), (5, 5, 4, 4, 0, 0), (0,) * 6, (2, 3, 4)),
    ("column_1d", (1, 1, 40), (0, 0, 0, 0, 6, 6), (0,) * 6, ()),
    ("wide_x", (70, 5, 6), (3, 3, 2, 2, 2, 2), (0,) * 6, ()),
]


@pytest.mark.parametrize("name,n,pml,pad,disable", CASES, ids=[c[0] for c in CASES])
def test_grid_steps_fp64(engine_be, oracle_be, name, n, pml, pad, disable):
    """Random fields, three materials (constant, Drude+2CP, Lorentz), 30 steps."""
    print(f"case {name} n={n} pml={pml}", flush=True)
    rng = np.random.default_rng(11)
    nx, ny, nz = n
    mg = rng.integers(0, 4, size=(nz, ny, nx)).astype(np.uint32)
    mats = [Material(1.7), GOLD, LORENTZ, DEBYE]
    e, o = build_pair(engine_be, oracle_be, n, pml, pad, mats, mg, seed=5, disable=disable)
    for _ in range(30):
        e.update_E(); o.update_E()
        e.update_H(); o.update_H()
    compare(e, o)


def test_grid_half_step(engine_be, oracle_be):
    """update_E alone then update_H alone, checked after each half step."""
    rng = np.random.default_rng(3)
    mg = rng.integers(0, 2, size=(7, 6, 8)).astype(np.uint32)
    e, o = build_pair(engine_be, oracle_be, (8, 6, 7), (3, 2, 2, 3, 3, 2), (0,) * 6, [Material(2.0), GOLD], mg, 9)
    for _ in range(3):
        e.update_E(); o.update_E()
        compare(e, o)
        e.update_H(); o.update_H()
        compare(e, o)


def test_plane_injection_and_sensor(engine_be, oracle_be):
    rng = np.random.default_rng(4)
    mg = rng.integers(0, 3, size=(16, 9, 10)).astype(np.uint32)
    e, o = build_pair(engine_be, oracle_be, (10, 9, 16), (0, 0, 0, 0, 5, 5), (0,) * 6,
                      [Material(1.0), Material(2.25), GOLD], mg, 2, Nt=40)
    tab = [rng.standard_normal(40) for _ in range(4)]
    lam = np.linspace(400e-9, 800e-9, 7)
    for g in (e, o):
        g.set_tapering(10)
    ids = []
    for g in (e, o):
        g.set_plane_injection(20, *tab)
        ids.append([g.add_spect_poynting(capi.NORMAL_Z, (1, 9, 2, 8, 4, 5), lam),
                    g.add_spect_poynting(capi.NORMAL_XM, (3, 4, 0, 9, 2, 12), lam),
                    g.add_spect_poynting(capi.NORMAL_Y, (0, 10, 5, 6, 0, 16), lam)])
    for _ in range(40):
        for g in (e, o):
            g.update_E(); g.update_H(); g.feed_sensors()
    compare(e, o)
    for s in range(3):
        se, so = e.sensor_spectra(ids[0][s]), o.sensor_spectra(ids[1][s])
        assert rel_l2(se, so) <= FIELD_TOL
        pe, po = e.sensor_poynting(ids[0][s], 7), o.sensor_poynting(ids[1][s], 7)
        assert np.max(np.abs(pe - po)) <= 1e-9 * max(1e-30, np.max(np.abs(po))), (pe, po)


def test_normal_incidence_small_fp64(engine_be, oracle_be):
    spec = workloads.slab_normal_incidence(n=(8, 6, 40), slab=(12, 24), Nt=400, Nl=12, pml_z=8)
    spec.tapering = 50
    sens = [capi.SensorReq(box=False, orientation=capi.NORMAL_Z, x1=0, x2=8, y1=0, y2=6, z1=30, z2=31, Nl=12,
                           lambda_min=400e-9, lambda_max=800e-9)]
    g = capi.run_normal_incidence(engine_be, spec, sens, want_fields=True)
    r = capi.run_normal_incidence(oracle_be, spec, sens, want_fields=True)
    assert rel_l2(g["fields"], r["fields"]) <= FIELD_TOL
    for k in ("R", "T"):
        assert np.max(np.abs(g[k] - r[k])) <= 1e-9, k
    for k in ("refx", "refy", "transx", "transy"):
        assert np.max(np.abs(g[k] - r[k])) <= 1e-9, k
    assert np.max(np.abs(g["sensors"] - r["sensors"])) <= 1e-9 * np.max(np.abs(r["sensors"]))


def test_config1_slab_in_full_fp64(engine_be, oracle_be):
    """The 64^3 slab configuration run in full: 64x64x90 grid (8-cell CPML and
    5-cell pads on +-z), 2000 steps, 50 wavelengths, FP64, against the oracle."""
    print("config 1: 64^3 slab, 2000 steps", flush=True)
    spec = workloads.slab_normal_incidence()
    g = capi.run_normal_incidence(engine_be, spec, want_fields=True)
    r = capi.run_normal_incidence(oracle_be, spec, want_fields=True)
    assert g["fields"].shape == (6, 90, 64, 64) and g["Nt_run"] == r["Nt_run"] == 2000
    assert rel_l2(g["fields"], r["fields"]) <= FIELD_TOL
    for k in ("R", "T", "refx", "refy", "transx", "transy"):
        assert np.max(np.abs(g[k] - r[k])) <= 1e-9, k


def test_single_particle_small_fp64(engine_be, oracle_be):
    spec, sens, _ = workloads.spheres_single_particle(n=20, D=2e-9, count=3, rmin=3e-9, rmax=6e-9, seed=5,
                                                      Nt=150, Nl=10, pml=5, precision=capi.FP64)
    g = capi.run_single_particle(engine_be, spec, sens, want_fields=True)
    r = capi.run_single_particle(oracle_be, spec, sens, want_fields=True)
    assert rel_l2(g["fields"], r["fields"]) <= FIELD_TOL
    assert np.max(np.abs(g["sensors"] - r["sensors"])) <= 1e-9 * np.max(np.abs(r["sensors"]))


def test_single_particle_fp32_spectra(engine_be, oracle_be):
    spec, sens, _ = workloads.spheres_single_particle(n=20, D=2e-9, count=3, rmin=3e-9, rmax=6e-9, seed=5,
                                                      Nt=300, Nl=10, pml=5, precision=capi.FP32)
    g = capi.run_single_particle(engine_be, spec, sens)
    spec.precision = capi.FP64
    r = capi.run_single_particle(oracle_be, spec, sens)
    scale = np.max(np.abs(r["sensors"]))
    err = np.max(np.abs(g["sensors"] - r["sensors"])) / scale
    print("fp32 relative spectral error", err)
    assert err <= 1e-4


@pytest.mark.parametrize("cfg", range(32))
def test_every_tile_shape_matches_oracle(engine_be, oracle_be, cfg):
    """Each compiled (BY, KZ, XCD-remap) tile of the E/H passes updates every cell."""
    import ctypes as C
    rng = np.random.default_rng(cfg)
    n = (70, 13, 37)
    mg = rng.integers(0, 3, size=(n[2], n[1], n[0])).astype(np.uint32)
    e, o = build_pair(engine_be, oracle_be, n, (3, 4, 2, 3, 5, 2), (0,) * 6, [Material(1.5), GOLD, LORENTZ], mg, 8)
    assert engine_be.fn("set_tiles")(e.h, cfg, cfg) == 0
    for _ in range(4):
        e.update_E(); o.update_E()
        e.update_H(); o.update_H()
    compare(e, o)


def test_autotune_keeps_state(engine_be, oracle_be):
    """The autotuner runs the passes on the grid itself and must restore its state."""
    import ctypes as C
    rng = np.random.default_rng(12)
    n = (40, 21, 30)
    mg = rng.integers(0, 2, size=(n[2], n[1], n[0])).astype(np.uint32)
    e, o = build_pair(engine_be, oracle_be, n, (4, 4, 4, 4, 4, 4), (0,) * 6, [Material(2.0), GOLD], mg, 3)
    for _ in range(3):
        e.step(); o.step()
    out = (C.c_int * 2)()
    assert engine_be.fn("autotune")(e.h, 2, out) == 0
    assert 0 <= out[0] < 32 and 0 <= out[1] < 32
    for _ in range(3):
        e.step(); o.step()
    compare(e, o)


def test_fp16_storage_round_trip(engine_be):
    """FP16 storage with per-component 2^e scale: write/read keeps 2^-11 relative precision."""
    D = (5e-9,) * 3
    Dt = 5e-9 / (np.sqrt(3.0) * 299792458.0) * 0.99
    g = Grid(engine_be, (24, 20, 16), 10, D, Dt, capi.MODE_CUSTOM, (0,) * 6, (0,) * 6, capi.FP16)
    g.set_matsgrid(np.zeros((16, 20, 24), dtype=np.uint32))
    g.set_material(0, Material(1.0))
    g.bootstrap()
    rng = np.random.default_rng(0)
    for f in range(6):
        a = rng.uniform(0.5, 1.0, size=(16, 20, 24)) * (1.0 if f < 3 else 1 / 377.0)
        g.write(f, a)
        b = g.read(f)
        assert np.max(np.abs(b - a) / np.abs(a)) <= 2.0 ** -11


def test_fp16_spectra_error_reported(engine_be, oracle_be):
    """FP16 storage, FP32 arithmetic: spectral error against the FP64 oracle (reported;
    guarded at 5e-2 of the spectrum's peak)."""
    spec, sens, _ = workloads.spheres_single_particle(n=20, D=2e-9, count=3, rmin=3e-9, rmax=6e-9, seed=5,
                                                      Nt=300, Nl=10, pml=5, precision=capi.FP16)
    g = capi.run_single_particle(engine_be, spec, sens)
    spec.precision = capi.FP64
    r = capi.run_single_particle(oracle_be, spec, sens)
    err = np.max(np.abs(g["sensors"] - r["sensors"])) / np.max(np.abs(r["sensors"]))
    print("fp16 relative spectral error", err)
    assert err <= 5e-2


def test_fp16_normal_incidence_spectra(engine_be, oracle_be):
    spec = workloads.slab_normal_incidence(n=(8, 6, 40), slab=(12, 24), Nt=600, Nl=12, pml_z=8)
    r = capi.run_normal_incidence(oracle_be, spec)
    spec.precision = capi.FP16
    g = capi.run_normal_incidence(engine_be, spec)
    spec.precision = capi.FP32
    f = capi.run_normal_incidence(engine_be, spec)
    e16 = max(np.max(np.abs(g["R"] - r["R"])), np.max(np.abs(g["T"] - r["T"])))
    e32 = max(np.max(np.abs(f["R"] - r["R"])), np.max(np.abs(f["T"] - r["T"])))
    print("normal incidence |R|,|T| error: fp32", e32, "fp16", e16)
    assert e32 <= 1e-4
    assert e16 <= 5e-2
