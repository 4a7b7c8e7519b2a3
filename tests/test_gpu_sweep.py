"""GPU parity of the custom mode (oscillator sources, 2D TE/TM) and of batched
execution.  Bars as the other GPU tests: FP64 fields <= 1e-10 relative L2 and
Poynting spectra <= 1e-9 relative to their peak against the CPU oracle; a run
inside a batch gives the same result as the same run alone (FP64: against the
oracle at the FP64 bar; FP32: against the FP32 run alone, <= 1e-6 relative,
and within 1e-4 of the FP64 oracle's spectrum)."""
import numpy as np
import pytest

from aether_amd import capi, workloads
from aether_amd.capi import Grid, Material

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-10


def rel_l2(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def spec_err(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


def small_run(seed, side, steps, precision=capi.FP64, pol=None):
    rng = workloads.MT19937_64(seed)
    spec, sens, info = workloads.sweep_run(rng, seed, side=side, steps=steps, Nl=16, precision=precision)
    if pol is not None and info["pol"] != pol:
        # flip the polarisation, keeping the structure
        tm = pol == "TM"
        spec.disable_fields = (capi.EX, capi.EY, capi.HZ) if tm else (capi.EZ, capi.HX, capi.HY)
        spec.sources[0].orientation = capi.NORMAL_Z if tm else capi.NORMAL_X
    return spec, sens, info


def finalized(spec, sens):
    return capi.mode_finalize(capi.engine(), spec, capi.KIND_CUSTOM, sens)


def engine_alone(spec, sens):
    run = capi.Run(capi.engine(), capi.KIND_CUSTOM, spec, sens)
    run.step(spec.Nt)
    out = run.finish(want_fields=True)
    run.close()
    return out


def test_oscillators_on_grid_fp64(engine_be, oracle_be):
    """Grid-level: random fields, CPML, three oscillators of different
    orientations (one added after 7 steps), 40 steps."""
    rng = np.random.default_rng(21)
    n = (18, 15, 12)
    D = (5e-9,) * 3
    Dt = 5e-9 / (np.sqrt(3.0) * 299792458.0) * 0.99
    mg = rng.integers(0, 2, size=n[::-1]).astype(np.uint32)
    grids = []
    for be in (engine_be, oracle_be):
        g = Grid(be, n, 60, D, Dt, capi.MODE_CUSTOM, (3, 3, 3, 3, 3, 3), (0,) * 6, capi.FP64)
        for f in range(6):
            g.set_pml(f)
        g.set_matsgrid(mg)
        g.set_material(0, Material(1.0))
        g.set_material(1, Material(3.5))
        g.bootstrap()
        g.add_oscillator(capi.NORMAL_Z, (2, 9, 4, 5, 3, 4), 400e-9, 800e-9)
        g.add_oscillator(capi.NORMAL_XM, (0, 18, 0, 15, 6, 7), 300e-9, 600e-9)
        grids.append(g)
    full = (grids[0].Nz, grids[0].Ny, grids[0].Nx)
    fields = [rng.standard_normal(full) * (1.0 if f < 3 else 1 / 377.0) for f in range(6)]
    for g in grids:
        for f in range(6):
            g.write(f, fields[f])
    e, o = grids
    for t in range(40):
        if t == 7:
            for g in grids:
                g.add_oscillator(capi.NORMAL_Y, (5, 6, 5, 6, 0, 12), 500e-9, 900e-9)
        for g in grids:
            g.update_E()
            g.update_H()
    for f in range(6):
        assert rel_l2(e.read(f), o.read(f)) <= FIELD_TOL, f


@pytest.mark.parametrize("pol", ["TM", "TE"])
def test_custom_2d_run_fp64(engine_be, oracle_be, pol):
    print(f"custom 2D {pol}", flush=True)
    spec, sens, _ = small_run(7, 40, 261, pol=pol)
    spec = finalized(spec, sens)
    g = engine_alone(spec, sens)
    r = capi.run_custom(oracle_be, spec, sens, want_fields=True)
    assert rel_l2(g["fields"], r["fields"]) <= FIELD_TOL
    assert spec_err(g["sensors"], r["sensors"]) <= 1e-9
    assert np.max(np.abs(r["sensors"])) > 0


def test_custom_3d_run_fp64(engine_be, oracle_be):
    """3D custom mode: dielectric block, oscillator plane source, box sensor."""
    n = (20, 18, 16)
    mg = np.zeros(n[::-1], dtype=np.uint32)
    mg[6:10, 5:12, 7:14] = 1
    src = capi.SourceReq(orientation=capi.NORMAL_X, x1=0, x2=20, y1=0, y2=18, z1=2, z2=3,
                         lambda_min=400e-9, lambda_max=800e-9)
    sens = [capi.SensorReq(box=True, x1=2, x2=18, y1=2, y2=16, z1=4, z2=14, Nl=9),
            capi.SensorReq(box=False, orientation=capi.NORMAL_Z, x1=0, x2=20, y1=0, y2=18, z1=15, z2=16, Nl=9)]
    spec = capi.ModeSpec(matsgrid=mg, materials=[Material(1.0), Material(4.0)], Nt=230, Dx=5e-9, Dy=5e-9,
                         Dz=5e-9, pml=(4, 4, 4, 4, 6, 6), tapering=40, sources=[src], precision=capi.FP64)
    spec = finalized(spec, sens)
    g = engine_alone(spec, sens)
    r = capi.run_custom(oracle_be, spec, sens, want_fields=True)
    assert rel_l2(g["fields"], r["fields"]) <= FIELD_TOL
    assert spec_err(g["sensors"], r["sensors"]) <= 1e-9


def test_custom_3d_dispersive_fp64(engine_be, oracle_be):
    """3D custom mode with a Drude + critical-point block, an oscillator and a box
    sensor: the RC pre-pass runs inside the sensor launch from the second step on."""
    print("custom 3D dispersive FP64", flush=True)
    n = (20, 18, 16)
    mg = np.zeros(n[::-1], dtype=np.uint32)
    mg[6:10, 5:12, 7:14] = 1
    src = capi.SourceReq(orientation=capi.NORMAL_Y, x1=0, x2=20, y1=2, y2=3, z1=0, z2=16,
                         lambda_min=400e-9, lambda_max=800e-9)
    sens = [capi.SensorReq(box=True, x1=3, x2=17, y1=4, y2=15, z1=4, z2=15, Nl=9)]
    spec = capi.ModeSpec(matsgrid=mg, materials=[Material(1.0), workloads.GOLD], Nt=260, Dx=5e-9, Dy=5e-9,
                         Dz=5e-9, pml=(6,) * 6, pad=(5,) * 6, tapering=40, sources=[src], precision=capi.FP64)
    g = engine_alone(spec, sens)
    r = capi.run_custom(oracle_be, spec, sens, want_fields=True)
    assert rel_l2(g["fields"], r["fields"]) <= FIELD_TOL
    assert spec_err(g["sensors"], r["sensors"]) <= 1e-9


def test_custom_3d_fp16_storage(engine_be, oracle_be):
    """FP16 field storage with an oscillator source (amplitude 50): finite fields
    and the planar spectrum within FP16's precision of the FP64 oracle."""
    print("custom 3D FP16 storage", flush=True)
    n = (20, 18, 16)
    mg = np.zeros(n[::-1], dtype=np.uint32)
    mg[6:10, 5:12, 7:14] = 1
    src = capi.SourceReq(orientation=capi.NORMAL_X, x1=0, x2=20, y1=0, y2=18, z1=2, z2=3,
                         lambda_min=400e-9, lambda_max=800e-9)
    sens = [capi.SensorReq(box=False, orientation=capi.NORMAL_Z, x1=0, x2=20, y1=0, y2=18, z1=15, z2=16, Nl=9)]
    spec = capi.ModeSpec(matsgrid=mg, materials=[Material(1.0), Material(4.0)], Nt=1200, Dx=5e-9, Dy=5e-9,
                         Dz=5e-9, pml=(6,) * 6, pad=(5,) * 6, tapering=40, sources=[src], precision=capi.FP16)
    g = engine_alone(spec, sens)
    r = capi.run_custom(oracle_be, capi.ModeSpec(**{**spec.__dict__, "precision": capi.FP64}), sens)
    assert np.all(np.isfinite(g["fields"]))
    err = spec_err(g["sensors"], r["sensors"])
    print("fp16 custom spectral error", err)
    assert err <= 1e-2


def _batch_runs(precision, specs):
    be = capi.engine()
    b = capi.Batch(be, precision)
    runs = []
    for spec, sens in specs:
        r = capi.Run(be, capi.KIND_CUSTOM, spec, sens)
        b.add(r)
        runs.append(r)
    return b, runs


def test_batch_fp64_matches_oracle(engine_be, oracle_be):
    """Three uneven 2D runs (TM/TE, step counts not multiples of 64) in one batch."""
    cases = [small_run(11, 36, 200, pol="TM"), small_run(12, 52, 131, pol="TE"), small_run(13, 44, 261, pol="TM")]
    specs = [(finalized(s, q), q) for s, q, _ in cases]
    b, runs = _batch_runs(capi.FP64, specs)
    while b.running():
        b.step(2)
    b.synchronize()
    for (spec, sens), run in zip(specs, runs):
        b.remove(run)
        assert run.progress() == (spec.Nt, spec.Nt)
        g = run.finish(want_fields=True)
        r = capi.run_custom(oracle_be, spec, sens, want_fields=True)
        assert rel_l2(g["fields"], r["fields"]) <= FIELD_TOL
        assert spec_err(g["sensors"], r["sensors"]) <= 1e-9
        run.close()
    b.close()


def test_batch_fp32_alone_equals_mixed(engine_be, oracle_be):
    """FP32: each run's result in the batch equals the same run alone; a run
    joins after two chunks of the others; spectra within 1e-4 of FP64."""
    cases = [small_run(21, 40, 300, capi.FP32), small_run(22, 60, 190, capi.FP32), small_run(23, 33, 257, capi.FP32)]
    specs = [(finalized(s, q), q) for s, q, _ in cases]
    be = capi.engine()
    b = capi.Batch(be, capi.FP32)
    runs = [capi.Run(be, capi.KIND_CUSTOM, s, q) for s, q in specs]
    b.add(runs[0]); b.add(runs[1])
    b.step(2)
    runs[2].step(64)          # the late run does its first chunk alone, then joins
    b.add(runs[2])
    while b.running():
        b.step(1)
    b.synchronize()
    for (spec, sens), run in zip(specs, runs):
        b.remove(run)
        mixed = run.finish(want_fields=True)
        run.close()
        alone = engine_alone(spec, sens)
        assert rel_l2(mixed["fields"], alone["fields"]) <= 1e-6
        # a grid stepped alone forms its DFT sums on the matrix cores from
        # split-bf16 operands (about 2^-17 of each term: ~1e-6 of the peak
        # here); the batch flush multiplies in FP64 on the vector ALU
        assert spec_err(mixed["sensors"], alone["sensors"]) <= 1e-5
        s64 = capi.ModeSpec(**{**spec.__dict__, "precision": capi.FP64})
        r = capi.run_custom(oracle_be, s64, sens)
        err = spec_err(mixed["sensors"], r["sensors"])
        print("batched fp32 spectral error vs fp64 oracle", err)
        assert err <= 1e-4
    b.close()


def test_batch_fp16_storage(engine_be, oracle_be):
    """FP16 storage in a batch (oscillator sources, per-step passes): each run as
    alone, finite, and within FP16's precision of the FP64 oracle."""
    print("batch FP16 storage", flush=True)
    cases = [small_run(31, 40, 300, capi.FP16, pol="TM"), small_run(32, 48, 200, capi.FP16, pol="TE")]
    specs = [(finalized(s, q), q) for s, q, _ in cases]
    b, runs = _batch_runs(capi.FP16, specs)
    while b.running():
        b.step(1)
    b.synchronize()
    for (spec, sens), run in zip(specs, runs):
        b.remove(run)
        mixed = run.finish(want_fields=True)
        run.close()
        assert np.all(np.isfinite(mixed["fields"]))
        alone = engine_alone(spec, sens)
        assert spec_err(mixed["sensors"], alone["sensors"]) <= 1e-6
        r = capi.run_custom(oracle_be, capi.ModeSpec(**{**spec.__dict__, "precision": capi.FP64}), sens)
        err = spec_err(mixed["sensors"], r["sensors"])
        print("batched fp16 spectral error vs fp64 oracle", err)
        assert err <= 1e-2
    b.close()


def test_batch_autotune_and_3d_member(engine_be, oracle_be):
    """The online tuner changes tile shapes between chunks without changing
    results; a 3D custom run can share the batch with 2D runs."""
    n = (20, 18, 16)
    mg = np.zeros(n[::-1], dtype=np.uint32)
    mg[6:10, 5:12, 7:14] = 1
    src = capi.SourceReq(orientation=capi.NORMAL_Y, x1=0, x2=20, y1=0, y2=18, z1=2, z2=3)
    sens3 = [capi.SensorReq(box=False, orientation=capi.NORMAL_Z, x1=0, x2=20, y1=0, y2=18, z1=15, z2=16, Nl=9)]
    spec3 = capi.ModeSpec(matsgrid=mg, materials=[Material(1.0), Material(4.0)], Nt=1500, Dx=5e-9, Dy=5e-9,
                          Dz=5e-9, pml=(4, 4, 4, 4, 6, 6), sources=[src], precision=capi.FP64)
    s2, q2, _ = small_run(31, 30, 1400)
    specs = [(finalized(spec3, sens3), sens3), (finalized(s2, q2), q2)]
    b, runs = _batch_runs(capi.FP64, specs)
    tiles = b.autotune()
    assert 0 <= tiles[0] < 32
    while b.running():
        b.step(4)
    for (spec, sens), run in zip(specs, runs):
        b.remove(run)
        g = run.finish(want_fields=True)
        r = capi.run_custom(oracle_be, spec, sens, want_fields=True)
        assert rel_l2(g["fields"], r["fields"]) <= FIELD_TOL
        assert spec_err(g["sensors"], r["sensors"]) <= 1e-9
        run.close()
    b.close()


def test_batch_refuses_what_it_cannot_step(engine_be):
    be = capi.engine()
    b = capi.Batch(be, capi.FP32)
    spec, sens, _ = small_run(41, 30, 100, capi.FP64)
    r64 = capi.Run(be, capi.KIND_CUSTOM, finalized(spec, sens), sens)
    with pytest.raises(capi.FDTDError, match="precision"):
        b.add(r64)
    sp = workloads.slab_normal_incidence(n=(4, 4, 40), slab=(10, 20), Nt=10, Nl=4)
    sp.precision = capi.FP32
    rs = capi.Run(be, capi.KIND_NORMAL, sp)
    with pytest.raises(capi.FDTDError, match="normal-incidence"):
        b.add(rs)
    spec, sens, _ = small_run(42, 30, 100, capi.FP32)
    r = capi.Run(be, capi.KIND_CUSTOM, finalized(spec, sens), sens)
    r.step(10)
    with pytest.raises(capi.FDTDError, match="multiple of 64"):
        b.add(r)
    for x in (r64, rs, r):
        x.close()
    b.close()


def _sensor_variety(seed, side, steps, pol):
    """A 2D run with a line sensor along x, one along y and a z-normal plane
    sensor, and a second oscillator on the other live E component for TE."""
    spec, sens, info = small_run(seed, side, steps, pol=pol)
    sens = sens + [capi.SensorReq(box=False, orientation=capi.NORMAL_XM, x1=side // 3, x2=side // 3 + 1, y1=2,
                                  y2=side - 3, z1=0, z2=1, Nl=5),
                   capi.SensorReq(box=False, orientation=capi.NORMAL_Z, x1=4, x2=side // 2, y1=side // 2,
                                  y2=side // 2 + 5, z1=0, z2=1, Nl=3)]
    if pol == "TE":
        spec.sources.append(capi.SourceReq(orientation=capi.NORMAL_Y, x1=side // 2, x2=side // 2 + 1, y1=3,
                                           y2=side - 3, z1=0, z2=1, lambda_min=450e-9, lambda_max=900e-9))
    return finalized(spec, sens), sens


# time-skew codes: tile height index (0, 1) | depth index (0: 2, 1: 4, 2: 8 steps) << 2;
# an 8-step skew needs the taller FP64 tile (16 rows leave no interior)
SKEW_CODES = [0, 1, 4, 5, 8, 9]
SKEW_CODES_FP64 = [0, 1, 4, 5, 9]


def test_batch_time_skew_rejects_unknown_codes(engine_be):
    for prec, bad in ((capi.FP32, (2, 3, 6, 7, 10, 11, 12, 16, -2)), (capi.FP64, (8, 2, 13, 17))):
        b = capi.Batch(capi.engine(), prec)
        try:
            for code in bad:
                with pytest.raises(capi.FDTDError):
                    b.set_skew(code)
        finally:
            b.close()


@pytest.mark.parametrize("code", [-1] + SKEW_CODES_FP64)
def test_batch_time_skew_codes_fp64(engine_be, oracle_be, code):
    """Every time-skew code (tile height, depth 2/4/8) and the per-step passes
    give the oracle's fields and spectra on uneven 2D TM/TE runs, including
    tiles larger than the grid and step counts that end inside a block."""
    cases = [_sensor_variety(51, 36, 203, "TM"), _sensor_variety(52, 150, 131, "TE"),
             _sensor_variety(53, 20, 77, "TE")]
    be = capi.engine()
    b = capi.Batch(be, capi.FP64)
    b.set_skew(code)
    runs = []
    for spec, sens in cases:
        r = capi.Run(be, capi.KIND_CUSTOM, spec, sens)
        b.add(r)
        runs.append(r)
    info = b.info()
    assert info["skewed"] == (0 if code < 0 else 3)
    while b.running():
        b.step(1)
    for (spec, sens), run in zip(cases, runs):
        b.remove(run)
        g = run.finish(want_fields=True)
        r = capi.run_custom(oracle_be, spec, sens, want_fields=True)
        assert rel_l2(g["fields"], r["fields"]) <= FIELD_TOL
        assert spec_err(g["sensors"], r["sensors"]) <= 1e-9
        run.close()
    b.close()


def _batch_results(prec, code, cases):
    be = capi.engine()
    b = capi.Batch(be, prec)
    b.set_skew(code)
    runs = []
    for spec, sens in cases:
        r = capi.Run(be, capi.KIND_CUSTOM, spec, sens)
        b.add(r)
        runs.append(r)
    while b.running():
        b.step(1)
    out = []
    for r in runs:
        b.remove(r)
        out.append(r.finish(want_fields=True))
        r.close()
    b.close()
    return out


@pytest.mark.parametrize("code", SKEW_CODES)
def test_batch_time_skew_codes_fp32(engine_be, code):
    """FP32 storage: every time-skew code against the per-step passes on the
    uneven TM/TE runs (FP32 tiles are twice as tall as FP64 ones)."""
    cases = [_sensor_variety(51, 36, 203, "TM"), _sensor_variety(52, 150, 131, "TE"),
             _sensor_variety(53, 20, 77, "TE")]
    for spec, _ in cases:
        spec.precision = capi.FP32
    ref = _batch_results(capi.FP32, -1, cases)
    got = _batch_results(capi.FP32, code, cases)
    for a, c in zip(ref, got):
        assert rel_l2(c["fields"], a["fields"]) <= 1e-5
        assert spec_err(c["sensors"], a["sensors"]) <= 1e-5


@pytest.mark.parametrize("prec,tol", [(capi.FP64, 1e-12), (capi.FP32, 1e-5)])
def test_batch_time_skew_large_grid(engine_be, prec, tol):
    """On grids with far more tiles than the GPU holds at once, the time-skewed
    passes give the per-step passes' result (tiles never read a neighbour's
    already-advanced halo)."""
    outs = []
    for code in (-1, None):
        be = capi.engine()
        b = capi.Batch(be, prec)
        if code is not None:
            b.set_skew(code)
        runs = []
        for seed, side, steps, pol in ((61, 1400, 300, "TM"), (62, 1100, 250, "TE")):
            spec, sens, _ = small_run(seed, side, steps, prec, pol=pol)
            spec = finalized(spec, sens)
            r = capi.Run(be, capi.KIND_CUSTOM, spec, sens)
            b.add(r)
            runs.append(r)
        assert b.info()["skewed"] == (0 if code == -1 else 2)
        while b.running():
            b.step(1)
        res = []
        for r in runs:
            b.remove(r)
            res.append(r.finish(want_fields=True))
            r.close()
        b.close()
        outs.append(res)
    for a, c in zip(*outs):
        assert rel_l2(c["fields"], a["fields"]) <= tol
        assert spec_err(c["sensors"], a["sensors"]) <= tol
