"""Host-side logic (CPU only): the C ABI exports, the XCD tile fold, workload generators."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from aether_amd import capi, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "aether_fdtd.h")).read()
    return sorted(set(re.findall(r"\b(aefdtd_[a-z_0-9]+)\s*\(", text)))


def test_engine_exports_every_declared_symbol():
    lib = C.CDLL(capi.ENGINE_SO)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_oracle_mirrors_the_grid_abi():
    import oracle
    lib = C.CDLL(oracle.ORACLE_SO)
    engine_only = {"aefdtd_profile_enable", "aefdtd_profile_read", "aefdtd_profile_kinds", "aefdtd_device_bytes", "aefdtd_stats",
                   "aefdtd_autotune", "aefdtd_set_tiles", "aefdtd_set_step_tiles", "aefdtd_get_step_tiles", "aefdtd_synchronize",
                   "aefdtd_sensor_poynting_noalloc",
                   "aefdtd_set_storage_scale", "aefdtd_get_storage_scale", "aefdtd_stream_bandwidth",
                   "aefdtd_stream_mix_bandwidth",
                   "aefdtd_sensor_out_len"}
    missing = [s for s in declared_symbols()
               if s not in engine_only and not s.startswith(("aefdtd_run_", "aefdtd_batch_", "aefdtd_server_"))
               and s != "aefdtd_estimate_bytes" or s in
               ("aefdtd_run_normal_incidence", "aefdtd_run_single_particle", "aefdtd_run_custom")]
    missing = [s for s in missing if not hasattr(lib, s.replace("aefdtd_", "aeoracle_"))]
    assert not missing, missing


def fold(w, B):
    """The XCD fold of tile_of() in kernels.hpp."""
    q, r = divmod(B, 8)
    x, loc = w % 8, w // 8
    return x * (q + 1) + loc if x < r else r * (q + 1) + (x - r) * q + loc


@pytest.mark.parametrize("B", [1, 7, 8, 9, 63, 64, 65, 1000, 14250, 14251])
def test_xcd_fold_is_a_permutation(B):
    m = [fold(w, B) for w in range(B)]
    assert sorted(m) == list(range(B))
    # each XCD receives one contiguous range of tiles
    for x in range(8):
        got = sorted(fold(w, B) for w in range(x, B, 8))
        assert got == list(range(got[0], got[0] + len(got))) if got else True


def test_sphere_voxelisation_corner_sampling():
    D = (2e-9,) * 3
    g = workloads.voxelize_spheres((10, 10, 10), D, [(10e-9, 10e-9, 10e-9, 4.5e-9, 1)])
    # ball of radius 2.25 cells around corner (5,5,5), sampled at cell corners (i*D, j*D, k*D)
    assert g[3, 5, 5] == 1 and g[5, 5, 7] == 1 and g[7, 6, 6] == 0 and g[6, 6, 6] == 1
    assert int(g.sum()) == sum(1 for i in range(10) for j in range(10) for k in range(10)
                               if (i - 5) ** 2 + (j - 5) ** 2 + (k - 5) ** 2 <= 5)


def test_random_spheres_do_not_overlap():
    sph = workloads.random_spheres(2, 40, (512e-9,) * 3, 8e-9, 20e-9)
    assert len(sph) == 40
    for a in range(40):
        for b in range(a):
            d = np.sqrt(sum((sph[a][c] - sph[b][c]) ** 2 for c in range(3)))
            assert d > sph[a][3] + sph[b][3]
        for c in range(3):
            assert sph[a][3] <= sph[a][c] <= 512e-9 - sph[a][3]


def test_engine_refuses_without_gpu_or_bad_args():
    """create fails with a status code (not a crash) on bad dimensions."""
    lib = capi.engine()
    h = C.c_void_p()
    p6 = (C.c_int * 6)(*([0] * 6))
    rc = lib.fn("create")(C.byref(h), 0, 4, 4, 10, 1e-9, 1e-9, 1e-9, 1e-18, capi.MODE_CUSTOM, p6, p6, capi.FP64)
    assert rc == capi.EINVAL
    assert b"dimension" in lib.fn("last_error")(None)
    # the bandwidth references check their arguments before any device call
    out = np.zeros(3)
    for name in ("stream_bandwidth", "stream_mix_bandwidth"):
        assert lib.fn(name)(1 << 10, 3, out.ctypes.data_as(capi._dp)) == capi.EINVAL
        assert lib.fn(name)(1 << 20, 0, out.ctypes.data_as(capi._dp)) == capi.EINVAL
        assert lib.fn(name)(1 << 20, 3, None) == capi.EINVAL


def test_discretize_and_run_checks_come_before_the_device():
    """Malformed structures and completion parameters are refused with EINVAL
    and a message before any device call (so also on a host without a GPU)."""
    be = capi.engine()
    for st, n, D in ((capi.Structure().add_cylinder(0, 0, 0, 0, 0, 0, 1e-9, 1), (4, 4, 4), (1e-9,) * 3),
                     (capi.Structure(0, 1, 1).loop(1, 0, 0), (4, 4, 4), (1e-9,) * 3),
                     (capi.Structure().add_sphere(0, 0, 0, -1, 1), (4, 4, 4), (1e-9,) * 3),
                     (capi.Structure(), (0, 4, 4), (1e-9,) * 3),
                     (capi.Structure(), (4, 4, 4), (0.0, 1e-9, 1e-9))):
        with pytest.raises(capi.FDTDError) as e:
            capi.discretize(be, st, n, D)
        assert e.value.code == capi.EINVAL and "discretize" in str(e.value)
    spec, sens = _tiny_custom()
    for bad in (dict(time_type=3), dict(time_type=2, cc_quant=0), dict(time_type=1, cc_step=0)):
        with pytest.raises(capi.FDTDError) as e:
            capi.Run(be, capi.KIND_CUSTOM, capi.ModeSpec(**{**spec.__dict__, **bad}), sens)
        assert e.value.code == capi.EINVAL


def _finalized(be, spec, kind, sensors):
    return capi.mode_finalize(be, spec, kind, sensors)


@pytest.mark.parametrize("backend", ["engine", "oracle"])
def test_mode_finalize_pads(backend):
    """FDTD_Mode::finalize (modes/lua_fdtd.cpp:74-203): a PML face gets a pad of
    at least 5, 5 beyond a sensor edge and 50 beyond an oscillator edge; a face
    without PML gets none; a one-cell axis is left alone; normal incidence
    drops the x/y PML and defaults the z PML to 25."""
    import oracle
    be = capi.engine() if backend == "engine" else oracle.load()
    n = 100
    g = np.zeros((1, n, n), dtype=np.uint32)
    src = capi.SourceReq(orientation=capi.NORMAL_Z, x1=0, x2=n, y1=0, y2=1, z1=0, z2=1)
    sens = capi.SensorReq(box=False, orientation=capi.NORMAL_Y, x1=0, x2=n, y1=n - 1, y2=n, z1=0, z2=1)
    spec = capi.ModeSpec(matsgrid=g, materials=[capi.Material(1.0)], pml=(16, 16, 16, 16, 0, 0),
                         pad=(0, 0, 0, 0, 7, 7), sources=[src])
    out = _finalized(be, spec, capi.KIND_CUSTOM, [sens])
    # x: oscillator spans [0, n): 50 - 0 and 50 + n - n; y: source at 0 -> 50,
    # sensor top edge n -> 5 + n - n = 5; z: one cell, pads untouched
    assert out.pad == (50, 50, 50, 5, 7, 7)
    assert out.pml == (16, 16, 16, 16, 0, 0)
    # a sensor 3 cells inside the lower x face needs 5 - 3 = 2 < 5: the minimum holds;
    # one at -4 (outside the structure) needs 9
    g3 = np.zeros((40, 30, 20), dtype=np.uint32)
    s2 = capi.SensorReq(box=True, x1=-4, x2=20, y1=3, y2=33, z1=0, z2=40)
    spec3 = capi.ModeSpec(matsgrid=g3, materials=[capi.Material(1.0)], pml=(8, 0, 8, 8, 8, 8), pad=(1, 9, 0, 0, 0, 6))
    out3 = _finalized(be, spec3, capi.KIND_SINGLE, [s2])
    assert out3.pad == (9, 0, 5, 8, 5, 6)
    # normal incidence: x/y PML removed (their pads follow), z PML defaulted
    spec4 = capi.ModeSpec(matsgrid=g3, materials=[capi.Material(1.0)], pml=(8, 8, 8, 8, 0, 12), pad=(3,) * 6)
    out4 = _finalized(be, spec4, capi.KIND_NORMAL, [])
    assert out4.pml == (0, 0, 0, 0, 25, 12)
    assert out4.pad == (0, 0, 0, 0, 5, 5)


def test_cylinder_packing():
    rng = workloads.MT19937_64(11)
    occ, disks = workloads.random_cylinders(rng, 200, 0.35)
    assert occ.mean() >= 0.35 * 0.97
    # disks never touch: centre distance > r1 + r2 (one clear cell between them)
    for a in range(len(disks)):
        xa, ya, ra = disks[a]
        assert 3.0 <= ra <= 15.0 and ra <= xa <= 200 - ra and ra <= ya <= 200 - ra
        for b in range(a):
            xb, yb, rb = disks[b]
            assert (xa - xb) ** 2 + (ya - yb) ** 2 > (ra + rb) ** 2


def test_sweep_steps_and_sizes():
    assert workloads.sweep_steps(256) == 4000 and workloads.sweep_steps(2048) == 12000
    assert workloads.sweep_steps(1152) == 8000
    rng = workloads.MT19937_64(4)
    spec, sens, info = workloads.sweep_run(rng, 0, side=300)
    assert spec.matsgrid.shape == (1, 300, 300) and info["steps"] == workloads.sweep_steps(300)
    assert set(spec.disable_fields) in ({capi.EX, capi.EY, capi.HZ}, {capi.EZ, capi.HX, capi.HY})


def _no_gpu():
    lib = capi.engine()
    h = C.c_void_p()
    p6 = (C.c_int * 6)(*([0] * 6))
    rc = lib.fn("create")(C.byref(h), 4, 4, 4, 10, 1e-9, 1e-9, 1e-9, 1e-18, capi.MODE_CUSTOM, p6, p6, capi.FP64)
    if rc == capi.OK:
        lib.fn("destroy")(h)
    return rc == capi.EDEVICE


def _tiny_custom(n=8, Nt=20, precision=capi.FP32):
    g = np.zeros((1, n, n), dtype=np.uint32)
    src = capi.SourceReq(orientation=capi.NORMAL_Z, x1=0, x2=n, y1=0, y2=1, z1=0, z2=1)
    sens = [capi.SensorReq(box=False, orientation=capi.NORMAL_Y, x1=0, x2=n, y1=n - 1, y2=n, z1=0, z2=1, Nl=4)]
    spec = capi.ModeSpec(matsgrid=g, materials=[capi.Material(1.0)], Nt=Nt, pml=(4, 4, 4, 4, 0, 0),
                         disable_fields=(capi.EX, capi.EY, capi.HZ), sources=[src], precision=precision)
    return capi.mode_finalize(capi.engine(), spec, capi.KIND_CUSTOM, sens), sens


def test_estimate_bytes_closed_form():
    """Fields, material ids and CPML profiles of a grid without PML, poles or
    sensors, plus the 2% + 8 MiB margin of the estimate.  The cell arrays of a
    3D grid have their 7-cell rows padded to 64 cells."""
    be = capi.engine()
    g = np.zeros((6, 5, 7), dtype=np.uint32)
    for prec, ss, cs in ((capi.FP64, 8, 8), (capi.FP32, 4, 4), (capi.FP16, 2, 4)):
        spec = capi.ModeSpec(matsgrid=g, materials=[capi.Material(2.0)], Nt=10, precision=prec)
        NA = 64 * 5 * 6
        b = 6 * NA * ss + NA + 6 * (7 + 5 + 6) * cs
        assert capi.estimate_bytes(be, capi.KIND_CUSTOM, spec) == b + b // 50 + (8 << 20)
    # a dispersive material adds the slot map and the per-cell pole state
    spec = capi.ModeSpec(matsgrid=np.ones((6, 5, 7), dtype=np.uint32), materials=[capi.Material(1.0), workloads.GOLD],
                         Nt=10, precision=capi.FP32)
    N, NA = 7 * 5 * 6, 64 * 5 * 6
    b = 6 * NA * 4 + NA + 6 * 18 * 4 + 4 * NA + N * (8 + 3 * 4 + 3 * 3 * 16)
    assert capi.estimate_bytes(be, capi.KIND_CUSTOM, spec) == b + b // 50 + (8 << 20)


def test_estimate_bytes_sensor_terms():
    """A plane spectral-Poynting sensor adds its FP64 accumulators (6 x Nl x P
    complex), the FP64 coefficient table (Nt x E|H x Nl complex), the matrix
    flush's copy of it in the arithmetic type (Nt x E|H x LP, LP = 2 Nl
    rounded up to 32) and the stage of 6 samples per plane cell (256 steps
    deep for FP64 arithmetic, 512 otherwise)."""
    be = capi.engine()
    g = np.zeros((6, 5, 7), dtype=np.uint32)
    Nt, Nl = 10, 10
    sens = [capi.SensorReq(box=False, orientation=capi.NORMAL_Z, x1=1, x2=5, y1=1, y2=4, z1=3, z2=3, Nl=Nl)]
    P, LP = 4 * 3, 32
    for prec, ss, cs in ((capi.FP64, 8, 8), (capi.FP32, 4, 4), (capi.FP16, 2, 4)):
        spec = capi.ModeSpec(matsgrid=g, materials=[capi.Material(2.0)], Nt=Nt, precision=prec)
        NA = 64 * 5 * 6
        b = 6 * NA * ss + NA + 6 * (7 + 5 + 6) * cs
        depth = 256 if prec == capi.FP64 else 512
        b += 6 * Nl * P * 16 + Nt * 2 * Nl * 16 + Nt * 2 * LP * cs + depth * 6 * P * cs
        assert capi.estimate_bytes(be, capi.KIND_CUSTOM, spec, sens) == b + b // 50 + (8 << 20), prec


def test_server_refuses_request_over_budget():
    """A request larger than the whole budget fails at admission with ENOMEM
    (no device work), and the server keeps serving."""
    spec, sens = _tiny_custom()
    srv = capi.Server(capi.engine(), 1 << 20)
    rid = srv.submit(capi.KIND_CUSTOM, spec, sens)
    assert srv.wait(rid, 30.0) == capi.REQ_FAILED
    with pytest.raises(capi.FDTDError) as e:
        srv.result(rid)
    assert e.value.code == capi.ENOMEM and "budget" in srv.error(rid)
    st = srv.stats()
    assert st["failed"] == 1 and st["resident"] == 0 and st["bytes_resident"] == 0
    assert srv.release(rid) == capi.OK
    srv.close()


def test_server_reports_device_failure_without_gpu():
    """Without a usable GPU, admission fails with EDEVICE and the error text."""
    if not _no_gpu():
        pytest.skip("a GPU is visible: the device-failure path is not reachable")
    spec, sens = _tiny_custom()
    srv = capi.Server(capi.engine(), 1 << 34)
    ids = [srv.submit(capi.KIND_CUSTOM, spec, sens) for _ in range(3)]
    for rid in ids:
        assert srv.wait(rid, 30.0) == capi.REQ_FAILED
        with pytest.raises(capi.FDTDError) as e:
            srv.result(rid)
        assert e.value.code == capi.EDEVICE
    assert srv.stats()["failed"] == 3
    srv.close()


def test_server_rejects_bad_submissions():
    spec, sens = _tiny_custom()
    srv = capi.Server(capi.engine(), 1 << 30)
    with pytest.raises(capi.FDTDError):
        srv.submit(7, spec, sens)
    bad = capi.ModeSpec(**{**spec.__dict__, "precision": 9})
    with pytest.raises(capi.FDTDError):
        srv.submit(capi.KIND_CUSTOM, bad, sens)
    with pytest.raises(capi.FDTDError):
        srv.wait(12345, 0.0)
    srv.close()


def test_oracle_cancel_flag_ends_after_the_step():
    """The oracle's drivers poll the caller's end_computation flag after every
    step (normal_incidence.cpp:431-435, single_particle.cpp:403): a flag set
    before the run starts ends it after one step, with the sensors of a
    one-step run (no tapering, so nothing else depends on Nt)."""
    import oracle
    be = oracle.load()
    spec, sens, _ = workloads.spheres_single_particle(n=12, D=2e-9, count=1, rmin=3e-9, rmax=4e-9, seed=5, Nt=40,
                                                      Nl=6, pml=4, precision=capi.FP64)
    flag = C.c_int(1)
    got = capi.run_single_particle(be, spec, sens, end_flag=flag)
    assert got["cancelled"] and oracle.last_steps() == 1
    spec1, _, _ = workloads.spheres_single_particle(n=12, D=2e-9, count=1, rmin=3e-9, rmax=4e-9, seed=5, Nt=1,
                                                    Nl=6, pml=4, precision=capi.FP64)
    one = capi.run_single_particle(be, spec1, sens)
    assert np.array_equal(got["sensors"], one["sensors"])
    flag.value = 0
    full = capi.run_single_particle(be, spec, sens, end_flag=flag)
    assert not full["cancelled"] and oracle.last_steps() == 40


def test_sensor_out_len_matches_the_python_view():
    """aefdtd_sensor_out_len (host code, no device call) agrees with
    SensorReq.out_len for every request kind."""
    from aether_amd import capi
    be = capi.engine()
    reqs = [capi.SensorReq(box=True, x1=0, x2=10, y1=0, y2=8, z1=0, z2=6, Nl=13),
            capi.SensorReq(box=False, orientation=capi.NORMAL_Y, x1=0, x2=10, y1=3, y2=4, z1=0, z2=6, Nl=7),
            capi.SensorReq(kind=capi.SENSOR_FIELDBLOCK, x1=1, x2=4, y1=2, y2=7, z1=0, z2=3, Nl=1),
            capi.SensorReq(kind=capi.SENSOR_FIELDMAP, orientation=capi.NORMAL_X, x1=2, x2=3, y1=0, y2=9, z1=1, z2=5,
                           Nl=1),
            capi.SensorReq(kind=capi.SENSOR_FIELDMAP2, orientation=capi.NORMAL_Z, x1=0, x2=6, y1=0, y2=5, z1=2, z2=3,
                           Nl=4)]
    for q in reqs:
        c = q.to_c()
        assert be.fn("sensor_out_len")(C.byref(c)) == q.out_len(), q.kind
