"""Automatic run length (Aether's CompletionSensor in energy mode).

Reference behaviour (sensors/s_completion.cpp:39-41,122-128,160-180 and the
drivers modes/default_fdtd.cpp:193-195, modes/single_particle.cpp:399-401,
modes/normal_incidence.cpp:417-423): every step's sensor feed sums
Ex^2+Ey^2+Ez^2 over the whole grid and keeps the running maximum; when
t % cc_step == 0 the driver asks for completion, which holds once the
sensor's step count (t + 1) is at least twice the longest grid axis and the
last sum is below cc_coeff times the maximum.  The custom and single-particle
drivers then stop with t + 1 steps run; normal incidence sets Nt = t, so its
R/T transforms use the first t samples.

CPU tests pin the oracle: the stop step against the rule applied to the
oracle's own per-step energies, those energies against the fields of fixed
runs of the same length, and the results of a stopped run against a
fixed-length run.  GPU tests hold the engine to the oracle: the same stop step
and FP64 results at the usual bars."""
import os
import numpy as np
import pytest

import oracle
from aether_amd import capi, workloads
from aether_amd.capi import Material

CC_STEP = 10


def custom_spec(precision=capi.FP64, Nt=4000, cc_coeff=1e-3):
    """A 14x12x10 dielectric block lit by an x-normal oscillator plane, CPML on
    every side (36x34x32 cells); the energy peaks near step 790."""
    n = (14, 12, 10)
    mg = np.zeros(n[::-1], dtype=np.uint32)
    mg[3:7, 4:8, 5:9] = 1
    src = capi.SourceReq(orientation=capi.NORMAL_X, x1=0, x2=14, y1=0, y2=12, z1=2, z2=3,
                         lambda_min=400e-9, lambda_max=800e-9)
    sens = [capi.SensorReq(box=True, x1=2, x2=12, y1=2, y2=10, z1=3, z2=8, Nl=7)]
    spec = capi.ModeSpec(matsgrid=mg, materials=[Material(1.0), Material(4.0)], Nt=Nt, Dx=5e-9, Dy=5e-9,
                         Dz=5e-9, pml=(6,) * 6, pad=(5,) * 6, tapering=0, sources=[src], precision=precision,
                         time_type=1, cc_step=CC_STEP, cc_coeff=cc_coeff)
    return spec, sens


def normal_spec(precision=capi.FP64):
    spec = workloads.slab_normal_incidence(n=(8, 6, 40), slab=(12, 24), Nt=4000, Nl=12, pml_z=8)
    spec.tapering = 0
    spec.precision = precision
    spec.time_type, spec.cc_step, spec.cc_coeff = 1, CC_STEP, 1e-3
    return spec


def single_spec(precision=capi.FP64):
    spec, sens, _ = workloads.spheres_single_particle(n=12, D=12e-9, count=2, rmin=18e-9, rmax=36e-9, seed=5,
                                                      Nt=3000, Nl=10, pml=5, precision=precision)
    spec.time_type, spec.cc_step, spec.cc_coeff = 1, CC_STEP, 1e-2
    return spec, sens


def expected_steps(trace, dims, cc_step, coeff, Nt):
    """The stop rule applied to a per-step energy trace: the number of steps
    the custom/single-particle drivers run."""
    mx = 0.0
    for t, e in enumerate(trace):
        mx = max(mx, e)
        if t % cc_step == 0 and t + 1 >= 2 * max(dims) and e / mx < coeff:
            return t + 1
    return Nt


def energy(fields):
    return float(np.sum(fields[0] ** 2 + fields[1] ** 2 + fields[2] ** 2))


def fixed(spec, **kw):
    return capi.ModeSpec(**{**spec.__dict__, "time_type": 0, **kw})


# ---------------------------------------------------------------- CPU: the oracle

def test_oracle_custom_stop_rule(oracle_be):
    spec, sens = custom_spec()
    r = capi.run_custom(oracle_be, spec, sens)
    steps, tr = oracle.last_steps(), oracle.energy_trace()
    dims = capi.custom_dims(spec)
    assert len(tr) == steps
    assert 2 * max(dims) < steps < spec.Nt          # stopped early, after the minimum length
    assert steps == expected_steps(tr, dims, CC_STEP, spec.cc_coeff, spec.Nt)
    assert (steps - 1) % CC_STEP == 0
    # the energies are the fields' sums of squares: fixed runs of two lengths
    for n in (steps, steps - 7):
        f = capi.run_custom(oracle_be, fixed(spec, Nt=n), sens, want_fields=True)
        assert abs(energy(f["fields"]) - tr[n - 1]) <= 1e-12 * tr[n - 1]
        if n == steps:
            # stopping is the same as a fixed run of that many steps (no tapering)
            assert np.array_equal(f["sensors"], r["sensors"])


def test_oracle_normal_incidence_sets_nt_to_t(oracle_be):
    spec = normal_spec()
    r = capi.run_normal_incidence(oracle_be, spec)
    steps, tr = oracle.last_steps(), oracle.energy_trace()
    Nz = spec.matsgrid.shape[0] + 2 * 8 + 2 * max(5, spec.pad[4])
    assert steps == expected_steps(tr, (8, 6, Nz), CC_STEP, spec.cc_coeff, spec.Nt) < spec.Nt
    assert r["Nt_run"] == steps - 1
    f = capi.run_normal_incidence(oracle_be, fixed(spec, Nt=steps - 1))
    for k in ("R", "T", "refx", "refy", "transx", "transy"):
        assert np.array_equal(f[k], r[k]), k


def test_oracle_fixed_runs_keep_their_length(oracle_be):
    spec, sens = custom_spec(Nt=300)
    capi.run_custom(oracle_be, fixed(spec), sens)
    assert oracle.last_steps() == 300


def test_oracle_refuses_unknown_criteria(oracle_be):
    spec, sens = custom_spec(Nt=20)
    for bad in (dict(time_type=3), dict(time_type=2, cc_quant=0), dict(time_type=2, cc_lmin=0.0),
                dict(time_type=1, cc_step=0)):
        with pytest.raises(capi.FDTDError):
            capi.run_custom(oracle_be, capi.ModeSpec(**{**spec.__dict__, **bad}), sens)


# ---------------------------------------------------------------- CPU: the spectral criterion

def ft_spec(coeff=1e-2, Nt=6000, quant=60, seed=5, layout="nnb"):
    spec = normal_spec()
    spec.Nt = Nt
    spec.time_type, spec.cc_step, spec.cc_coeff = 2, CC_STEP, coeff
    spec.cc_quant, spec.cc_lmin, spec.cc_lmax, spec.cc_seed, spec.cc_layout = quant, 400e-9, 800e-9, seed, layout
    return spec


def ft_read():
    import ctypes as C
    lib = C.CDLL(oracle.ORACLE_SO)
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    lib.aeoracle_ft_info(C.byref(a), C.byref(b), C.byref(c))
    navg, n, nch = a.value, b.value, c.value
    pts = np.zeros(3 * n, dtype=np.int32)
    lam = np.zeros(n)
    ring = np.zeros(6 * n * navg)
    checks = np.zeros(2 * nch, dtype=np.int32)
    ip = C.POINTER(C.c_int)
    dp = C.POINTER(C.c_double)
    lib.aeoracle_ft_read(pts.ctypes.data_as(ip), lam.ctypes.data_as(dp), ring.ctypes.data_as(dp),
                         checks.ctypes.data_as(ip))
    return pts.reshape(n, 3), lam, ring.reshape(6, n, navg), checks.reshape(nch, 2)


def settled(ring, coeff):
    """The FT check restated from s_completion.cpp:43-120, averages carried over between cells."""
    n = ring.shape[1]
    avg = np.zeros(6)
    nok = 0
    for p in range(n):
        for c in range(6):
            acc = avg[c]
            for v in ring[c, p]:
                acc += v
            avg[c] = abs(acc / ring.shape[2])
        amax = avg.max()
        tg = int(np.argmax(avg == amax))
        row = ring[tg, p]
        nok += abs(row.max() - row.min()) <= amax * coeff
    return nok


def test_oracle_ft_check_rule(oracle_be):
    spec = ft_spec()
    r = capi.run_normal_incidence(oracle_be, spec)
    steps = oracle.last_steps()
    pts, lam, ring, checks = ft_read()
    nz = spec.matsgrid.shape[0] + 2 * 8 + 2 * max(5, spec.pad[4])
    assert steps < spec.Nt and r["Nt_run"] == steps - 1 and (steps - 1) % CC_STEP == 0
    # checks from the first step count >= 2 max(N), every cc_step; the last one completes
    assert checks[0, 0] >= 2 * nz > checks[0, 0] - CC_STEP and checks[-1, 0] == steps
    assert np.all(np.diff(checks[:, 0]) == CC_STEP)
    assert checks[-1, 1] >= 0.9 * len(lam) > checks[-2, 1]
    assert settled(ring, spec.cc_coeff) == checks[-1, 1]
    # nnb: the cells sit on the structure's lowest and highest z planes
    zs_s = 8 + max(5, spec.pad[4])
    assert set(pts[:, 2]) == {zs_s, zs_s + spec.matsgrid.shape[0] - 1}
    assert np.all((lam >= 400e-9) & (lam <= 800e-9))


def test_oracle_ft_feed_is_a_running_dft(oracle_be):
    """The ring holds Re/Im of sum_s local_E(s) exp(i w s Dt): checked against the
    fields of fixed runs of every length on a small grid."""
    spec = ft_spec(coeff=1e-12, Nt=25, quant=6, layout="v")
    spec.matsgrid = spec.matsgrid[:, :2, :3].copy()
    capi.run_normal_incidence(oracle_be, spec)
    pts, lam, ring, _ = ft_read()
    navg = ring.shape[2]
    Dt = 5e-9 / (np.sqrt(3.0) * 299792458.0) * 0.99 * spec.time_mod
    w = 2.0 * np.pi * 299792458.0 / lam
    acc = np.zeros((6, len(lam)))
    for s in range(spec.Nt):
        f = capi.run_normal_incidence(oracle_be, capi.ModeSpec(**{**spec.__dict__, "time_type": 0, "Nt": s + 1}),
                                      want_fields=True)["fields"]
        nz, ny, nx = f.shape[1:]
        for p, (i, j, k) in enumerate(pts):
            ip, jp, kp = (i + 1) % nx, (j + 1) % ny, (k + 1) % nz
            loc = [(f[0, k, j, i] + f[0, k, jp, i] + f[0, kp, j, i] + f[0, kp, jp, i]) / 4,
                   (f[1, k, j, i] + f[1, k, j, ip] + f[1, kp, j, i] + f[1, kp, j, ip]) / 4,
                   (f[2, k, j, i] + f[2, k, j, ip] + f[2, k, jp, i] + f[2, k, jp, ip]) / 4]
            ph = np.exp(1j * w[p] * s * Dt)
            for c in range(3):
                acc[c, p] += (loc[c] * ph).real
                acc[3 + c, p] += (loc[c] * ph).imag
    last = (spec.Nt - 1) % navg
    assert np.max(np.abs(ring[:, :, last] - acc)) <= 1e-12 * np.max(np.abs(acc))
    assert np.max(np.abs(acc)) > 0


@pytest.mark.skipif(not os.path.exists(oracle.REF_SO), reason="oracle/_ref not built")
def test_oracle_ft_placement_follows_the_reference_generator(oracle_be):
    """Cells and wavelengths against Aether's own compiled randp/seedp
    (common/mathUT.cpp) driven through CompletionSensor::initialize's order."""
    import ctypes as C
    ref = C.CDLL(oracle.REF_SO)
    ref.aeref_randp.argtypes = [C.c_double, C.c_double]
    ref.aeref_randp.restype = C.c_double
    ref.aeref_seedp.argtypes = [C.c_int]
    for layout, seed in (("nnb", 5), ("bdn", 11), ("v", 3)):
        spec = ft_spec(Nt=2, quant=37, seed=seed, layout=layout)
        capi.run_normal_incidence(oracle_be, spec)
        pts, lam, _, _ = ft_read()
        nz, ny, nx = spec.matsgrid.shape
        zpad = 8 + max(5, spec.pad[4])
        Nx, Ny, Nz = nx, ny, nz + 2 * zpad
        xs, ys, zs = (0, nx), (0, ny), (zpad, zpad + nz)
        ref.aeref_seedp(seed)
        rp = lambda a, b: int(ref.aeref_randp(a, b) + 0.5)
        want = []
        if layout == "v":
            for _ in range(37):
                want.append((rp(xs[0], xs[1] - 1), rp(ys[0], ys[1] - 1), rp(zs[0], zs[1] - 1)))
        else:
            area = {0: Ny * Nz, 1: Nx * Nz, 2: Nx * Ny}
            n = []
            for a in range(3):
                n += [area[a] if layout[a] in "db" else 0, area[a] if layout[a] in "ub" else 0]
            tot = float(sum(n))
            left = 37
            for f in range(6):
                k = min(int(37 * n[f] / tot + 0.5), left)
                left -= k
                n[f] = k
            lo, hi = (xs[0], ys[0], zs[0]), (xs[1] - 1, ys[1] - 1, zs[1] - 1)
            for f in range(6):
                a = f // 2
                for _ in range(n[f]):
                    c = []
                    for b in range(3):
                        c.append((hi[b] if f % 2 else lo[b]) if b == a else rp(lo[b], hi[b]))
                    want.append(tuple(c))
            want += [(0, 0, 0)] * (37 - len(want))
        lam_ref = [ref.aeref_randp(400e-9, 800e-9) for _ in range(37)]
        assert [tuple(p) for p in pts] == want, layout
        assert list(lam) == lam_ref, layout


# ---------------------------------------------------------------- GPU: the engine

def run_engine(kind, spec, sens=(), chunk=64):
    run = capi.Run(capi.engine(), kind, spec, sens)
    t, nt = run.progress()
    while t < nt:
        run.step(chunk)
        t, nt = run.progress()
    out = run.finish(want_fields=kind != capi.KIND_NORMAL)
    run.close()
    return t, out


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.gpu
def test_engine_custom_stops_with_oracle_fp64(oracle_be):
    print("energy completion: custom FP64", flush=True)
    spec, sens = custom_spec()
    steps, g = run_engine(capi.KIND_CUSTOM, spec, sens)
    r = capi.run_custom(oracle_be, spec, sens, want_fields=True)
    assert steps == oracle.last_steps() < spec.Nt
    assert rel_l2(g["fields"], r["fields"]) <= 1e-10
    assert np.max(np.abs(g["sensors"] - r["sensors"])) <= 1e-9 * np.max(np.abs(r["sensors"]))
    # the one-shot driver stops at the same step
    d = capi.run_custom(capi.engine(), spec, sens)
    assert np.array_equal(d["sensors"], g["sensors"])


@pytest.mark.gpu
def test_engine_single_particle_stops_with_oracle_fp64(oracle_be):
    print("energy completion: single particle FP64", flush=True)
    spec, sens = single_spec()
    steps, g = run_engine(capi.KIND_SINGLE, spec, sens)
    r = capi.run_single_particle(oracle_be, spec, sens)
    assert steps == oracle.last_steps() < spec.Nt
    assert np.max(np.abs(g["sensors"] - r["sensors"])) <= 1e-9 * np.max(np.abs(r["sensors"]))


@pytest.mark.gpu
def test_engine_normal_incidence_stops_with_oracle_fp64(oracle_be):
    print("energy completion: normal incidence FP64", flush=True)
    spec = normal_spec()
    steps, g = run_engine(capi.KIND_NORMAL, spec)
    r = capi.run_normal_incidence(oracle_be, spec)
    assert steps == oracle.last_steps() and g["Nt_run"] == r["Nt_run"] == steps - 1
    for k in ("R", "T"):
        assert np.max(np.abs(g[k] - r[k])) <= 1e-9, k


@pytest.mark.gpu
def test_engine_2d_custom_stops_with_oracle_fp64(oracle_be):
    """A 2D TM run (the separate energy pass, not the fused 3D one)."""
    print("energy completion: 2D custom FP64", flush=True)
    rng = workloads.MT19937_64(7)
    spec, sens, _ = workloads.sweep_run(rng, 7, side=60, steps=6000, Nl=8, precision=capi.FP64)
    spec = capi.mode_finalize(oracle_be, spec, capi.KIND_CUSTOM, sens)
    spec.time_type, spec.cc_step, spec.cc_coeff = 1, CC_STEP, 1e-3
    steps, g = run_engine(capi.KIND_CUSTOM, spec, sens)
    r = capi.run_custom(oracle_be, spec, sens, want_fields=True)
    assert steps == oracle.last_steps() < spec.Nt
    assert rel_l2(g["fields"], r["fields"]) <= 1e-10
    assert np.max(np.abs(g["sensors"] - r["sensors"])) <= 1e-9 * np.max(np.abs(r["sensors"]))


@pytest.mark.gpu
def test_engine_fp16_storage_energy_stop(oracle_be):
    """FP16 field storage takes the separate energy pass (the fused sum needs
    the stored values in the arithmetic type): the stop lands within a few
    checks of the FP64 oracle's."""
    print("energy completion: custom FP16 storage", flush=True)
    spec, sens = custom_spec(precision=capi.FP16)
    steps, _ = run_engine(capi.KIND_CUSTOM, spec, sens)
    capi.run_custom(oracle_be, fixed(spec, precision=capi.FP64, time_type=1), sens)
    ref = oracle.last_steps()
    print("fp16 stop", steps, "oracle", ref)
    assert abs(steps - ref) <= 5 * CC_STEP and steps < spec.Nt


@pytest.mark.gpu
def test_engine_2d_ft_criterion_fp64(oracle_be):
    print("spectral completion: 2D custom FP64", flush=True)
    rng = workloads.MT19937_64(7)
    spec, sens, _ = workloads.sweep_run(rng, 7, side=60, steps=6000, Nl=8, precision=capi.FP64)
    spec = capi.mode_finalize(oracle_be, spec, capi.KIND_CUSTOM, sens)
    spec = capi.ModeSpec(**{**spec.__dict__, "time_type": 2, "cc_step": CC_STEP, "cc_coeff": 1e-1, "cc_quant": 40,
                            "cc_lmin": 500e-9, "cc_lmax": 1000e-9, "cc_seed": 9, "cc_layout": "bbn"})
    steps, g = run_engine(capi.KIND_CUSTOM, spec, sens)
    r = capi.run_custom(oracle_be, spec, sens)
    assert steps == oracle.last_steps() < spec.Nt
    assert np.max(np.abs(g["sensors"] - r["sensors"])) <= 1e-9 * np.max(np.abs(r["sensors"]))


@pytest.mark.gpu
def test_engine_fp32_stop_step(oracle_be):
    """FP32 fields change the energy sums in their last bits: the stop lands on
    the oracle's check or the next one."""
    print("energy completion: custom FP32", flush=True)
    spec, sens = custom_spec(precision=capi.FP32)
    steps, g = run_engine(capi.KIND_CUSTOM, spec, sens)
    capi.run_custom(oracle_be, fixed(spec, precision=capi.FP64, time_type=1), sens)
    ref = oracle.last_steps()
    print("fp32 stop", steps, "oracle", ref)
    assert abs(steps - ref) <= CC_STEP


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["normal", "single", "custom"])
def test_engine_ft_criterion_stops_with_oracle_fp64(oracle_be, kind):
    """The spectral criterion in the three drivers, FP64: the same stop step as
    the oracle and results at the FP64 bars."""
    print(f"spectral completion: {kind} FP64", flush=True)
    ft = dict(time_type=2, cc_step=CC_STEP, cc_quant=60, cc_lmin=400e-9, cc_lmax=800e-9, cc_seed=5)
    if kind == "normal":
        spec = ft_spec()
        steps, g = run_engine(capi.KIND_NORMAL, spec)
        r = capi.run_normal_incidence(oracle_be, spec)
        assert steps == oracle.last_steps() < spec.Nt and g["Nt_run"] == r["Nt_run"]
        for k in ("R", "T"):
            assert np.max(np.abs(g[k] - r[k])) <= 1e-9, k
        return
    if kind == "single":
        spec, sens = single_spec()
        spec = capi.ModeSpec(**{**spec.__dict__, **ft, "cc_coeff": 1e-2})
        steps, g = run_engine(capi.KIND_SINGLE, spec, sens)
        r = capi.run_single_particle(oracle_be, spec, sens)
    else:
        spec, sens = custom_spec()
        spec = capi.ModeSpec(**{**spec.__dict__, **ft, "cc_coeff": 1e-2, "cc_layout": "bbb"})
        steps, g = run_engine(capi.KIND_CUSTOM, spec, sens)
        r = capi.run_custom(oracle_be, spec, sens)
    assert steps == oracle.last_steps() < spec.Nt
    assert np.max(np.abs(g["sensors"] - r["sensors"])) <= 1e-9 * np.max(np.abs(r["sensors"]))


@pytest.mark.gpu
def test_engine_ft_criterion_fp32_and_served(oracle_be):
    """FP32 fields: the spectral check lands on the oracle's check or within a
    few; the server steps such a run on its own and returns the direct run's
    result."""
    print("spectral completion: normal FP32, served", flush=True)
    spec = ft_spec()
    capi.run_normal_incidence(oracle_be, spec)
    ref = oracle.last_steps()
    s32 = capi.ModeSpec(**{**spec.__dict__, "precision": capi.FP32})
    steps, direct = run_engine(capi.KIND_NORMAL, s32)
    print("fp32 spectral stop", steps, "oracle", ref)
    assert abs(steps - ref) <= 5 * CC_STEP and steps < spec.Nt
    srv = capi.Server(capi.engine(), 64 << 30)
    rid = srv.submit(capi.KIND_NORMAL, s32, [])
    assert srv.wait(rid, 120.0) == capi.REQ_DONE, srv.error(rid)
    out = srv.result(rid)[0]
    srv.close()
    assert np.array_equal(out[-2 * spec.Nl:], np.concatenate([direct["R"], direct["T"]]))


@pytest.mark.gpu
def test_batch_and_server_with_completion():
    """Batches step fixed-length runs only; the server runs a completion
    request on its own and gives the direct run's result."""
    print("energy completion: batch refusal, server", flush=True)
    spec, sens = custom_spec(precision=capi.FP32, cc_coeff=1e-2)
    b = capi.Batch(capi.engine(), capi.FP32)
    run = capi.Run(capi.engine(), capi.KIND_CUSTOM, spec, sens)
    with pytest.raises(capi.FDTDError):
        b.add(run)
    run.close()
    b.close()
    steps, direct = run_engine(capi.KIND_CUSTOM, spec, sens)
    srv = capi.Server(capi.engine(), 64 << 30)
    rid = srv.submit(capi.KIND_CUSTOM, spec, sens)
    assert srv.wait(rid, 120.0) == capi.REQ_DONE, srv.error(rid)
    served = srv.result(rid)[0]
    srv.close()
    assert steps < spec.Nt
    assert np.array_equal(served, direct["sensors"])
