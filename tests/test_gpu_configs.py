"""The benchmark configurations at their own sizes against the oracle, and the
staged DFT flush over more wavelengths than one pass of its matrix blocks.

- The 256^3 scatterer configuration (300^3 grid, 40 Drude + 2 CP spheres,
  12-cell CPML on six faces, TFSF, 64-wavelength box sensor) in FP64 for 520
  steps (two 256-step DFT flushes and a partial one): fields <= 1e-10 relative
  L2, spectra <= 1e-9 of their peak against the oracle.
- The same configuration in FP32 over 3000 steps: box spectra within 1e-4 of
  the FP64 engine's (which the test above holds to the oracle).
- Sensors with 40, 64 and 128 wavelengths against the oracle in FP64: 5, 8
  and 16 blocks of 16 (re, im) rows of the matrix flush (eight blocks per
  pass, so 128 wavelengths take two passes), across two 256-step flushes and
  a partial one.
- The precision configuration's layer stack: normal incidence is invariant in
  x and y, so the full 384-cell stack (12 layers + Lorentz layer, 16-cell
  CPML, 8000 steps, 128 wavelengths) runs on a 4 x 4 cross-section in FP64
  against the oracle; the 512 x 512 runs in FP32 (1e-4) and FP16 storage
  (reported) are compared with it.
"""
import numpy as np
import pytest

from aether_amd import capi, workloads
from aether_amd.capi import Grid, Material

pytestmark = pytest.mark.gpu

GOLD = workloads.GOLD


def rel_l2(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def test_config2_full_size_fp64_against_oracle(engine_be, oracle_be):
    print("config 2 at size: 300^3 grid, 40 spheres, FP64, 520 steps", flush=True)
    spec, sens, sph = workloads.spheres_single_particle(Nt=520, precision=capi.FP64)
    assert len(sph) == 40 and sens[0].Nl == 64
    g = capi.run_single_particle(engine_be, spec, sens, want_fields=True)
    print("engine done; oracle (16 threads)", flush=True)
    oracle_be.fn("set_threads")(16)
    r = capi.run_single_particle(oracle_be, spec, sens, want_fields=True)
    oracle_be.fn("set_threads")(0)
    assert g["fields"].shape == (6, 300, 300, 300)
    for f in range(6):
        err = rel_l2(g["fields"][f], r["fields"][f])
        print(f"field {f}: rel-L2 {err:.3e}", flush=True)
        assert err <= 1e-10, (f, err)
    peak = np.max(np.abs(r["sensors"]))
    serr = np.max(np.abs(g["sensors"] - r["sensors"])) / peak
    print(f"box spectra (64 wavelengths, 7 columns): max diff / peak {serr:.3e}", flush=True)
    assert peak > 0 and serr <= 1e-9


def test_config2_full_size_fp32_spectra_within_1e4(engine_be):
    print("config 2 at size: FP32 against FP64, 3000 steps", flush=True)
    spec, sens, _ = workloads.spheres_single_particle(Nt=3000, precision=capi.FP64)
    r = capi.run_single_particle(engine_be, spec, sens)["sensors"]
    spec.precision = capi.FP32
    g = capi.run_single_particle(engine_be, spec, sens)["sensors"]
    Nl = sens[0].Nl
    peak = np.max(np.abs(r[:Nl]))
    err = np.max(np.abs(g - r)) / peak
    print(f"FP32 box spectra: max diff / peak of the total {err:.3e}", flush=True)
    assert peak > 0 and err <= 1e-4


@pytest.mark.parametrize("Nl", [40, 64, 128])
def test_sensors_many_wavelengths(engine_be, oracle_be, Nl):
    """Three planar sensors with Nl wavelengths on random fields with a dispersive
    material, 600 steps (two 256-step flushes and a partial one) in FP64."""
    rng = np.random.default_rng(Nl)
    n = (14, 12, 18)
    D = (5e-9,) * 3
    Dt = min(D) / (np.sqrt(3.0) * 299792458.0) * 0.99
    mg = rng.integers(0, 2, size=(n[2], n[1], n[0])).astype(np.uint32)
    grids = []
    for be in (engine_be, oracle_be):
        g = Grid(be, n, 150, D, Dt, capi.MODE_CUSTOM, (3, 3, 0, 0, 4, 4), (0,) * 6, capi.FP64)
        for f in range(6):
            g.set_pml(f)
        g.set_matsgrid(mg)
        g.set_material(0, Material(1.5))
        g.set_material(1, GOLD)
        g.set_tapering(40)
        g.bootstrap()
        grids.append(g)
    e = grids[0]
    for f in range(6):
        a = rng.standard_normal((e.Nz, e.Ny, e.Nx)) * (1.0 if f < 3 else 1 / 376.73)
        for g in grids:
            g.write(f, a)
    lam = np.linspace(400e-9, 800e-9, Nl)
    ids = [[g.add_spect_poynting(capi.NORMAL_Z, (2, 12, 1, 11, 12, 13), lam),
            g.add_spect_poynting(capi.NORMAL_XM, (4, 5, 0, 12, 2, 16), lam),
            g.add_spect_poynting(capi.NORMAL_Y, (0, 14, 7, 8, 3, 15), lam)] for g in grids]
    for _ in range(600):
        for g in grids:
            g.update_E(); g.update_H(); g.feed_sensors()
    e, o = grids
    for s in range(3):
        se, so = e.sensor_spectra(ids[0][s]), o.sensor_spectra(ids[1][s])
        err = rel_l2(se, so)
        assert err <= 1e-10, (s, err)
        # every wavelength chunk carries signal
        for l0 in range(0, Nl, 32):
            assert np.max(np.abs(so[:, l0:l0 + 32])) > 0
        pe, po = e.sensor_poynting(ids[0][s], Nl), o.sensor_poynting(ids[1][s], Nl)
        assert np.max(np.abs(pe - po)) <= 1e-9 * np.max(np.abs(po)), s


@pytest.mark.parametrize("Nl,Nt", [(8, 20), (17, 37), (50, 300), (128, 513)])
def test_fp32_matrix_core_flush_against_oracle(engine_be, oracle_be, Nl, Nt):
    """The FP32 flush (k_dft_mx: split-bf16 matrix products of 16 steps, 32
    wavelength columns per block, blocks of 4, 512-step staging) on runs whose
    last flush is partial and whose wavelength count fills 1, 2, 4 and 8
    blocks, some partly: FP32 spectra within 1e-4 of their peak of the FP64
    oracle.  Runs of fewer steps than the flush's prefetch depth end their
    single flush at the table's last row (the flush reads clamped rows past
    it)."""
    spec, sens, _ = workloads.spheres_single_particle(n=16, D=2e-9, count=2, rmin=3e-9, rmax=5e-9, seed=9,
                                                      Nt=Nt, Nl=Nl, pml=5, precision=capi.FP64)
    r = capi.run_single_particle(oracle_be, spec, sens)["sensors"]
    spec.precision = capi.FP32
    g = capi.run_single_particle(engine_be, spec, sens)["sensors"]
    peak = np.max(np.abs(r))
    err = np.max(np.abs(g - r)) / peak
    print(f"Nl {Nl}, {Nt} steps: FP32 flush vs FP64 oracle, max diff / peak {err:.3e}", flush=True)
    assert peak > 0 and err <= 1e-4


def test_box_sensor_64_wavelengths_single_particle(engine_be, oracle_be):
    spec, sens, _ = workloads.spheres_single_particle(n=20, D=2e-9, count=3, rmin=3e-9, rmax=6e-9, seed=5,
                                                      Nt=200, Nl=64, pml=5, precision=capi.FP64)
    g = capi.run_single_particle(engine_be, spec, sens)
    r = capi.run_single_particle(oracle_be, spec, sens)
    assert np.max(np.abs(g["sensors"] - r["sensors"])) <= 1e-9 * np.max(np.abs(r["sensors"]))


def test_config3_stack_cross_section_and_full_width(engine_be, oracle_be):
    print("config 3: 4x4 cross-section, FP64, 8000 steps, 128 wavelengths", flush=True)
    spec, info = workloads.multilayer_normal_incidence(n=(4, 4, 384), precision=capi.FP64)
    assert spec.Nl == 128 and spec.Nt == 8000 and len(spec.materials) == 14
    g = capi.run_normal_incidence(engine_be, spec, want_fields=True)
    r = capi.run_normal_incidence(oracle_be, spec, want_fields=True)
    assert rel_l2(g["fields"], r["fields"]) <= 1e-10
    for k in ("R", "T", "refx", "refy", "transx", "transy"):
        assert np.max(np.abs(g[k] - r[k])) <= 1e-9, k
    for name, prec, tol in (("FP32", capi.FP32, 1e-4), ("FP16 storage", capi.FP16, 5e-2)):
        print(f"config 3: 512x512 cross-section, {name}", flush=True)
        spec_w, _ = workloads.multilayer_normal_incidence(precision=prec)
        w = capi.run_normal_incidence(engine_be, spec_w)
        err = max(np.max(np.abs(w["R"] - r["R"])), np.max(np.abs(w["T"] - r["T"])))
        print(f"{name}: max |dR|, |dT| against the FP64 stack {err:.3e}", flush=True)
        assert err <= tol, (name, err)
