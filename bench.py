"""Benchmark of the FDTD hot path on one MI355X.

Workload (the headline configuration): a 256^3-cell structure, 2 nm cells, 40
non-overlapping gold-like spheres (Drude + 2 critical-point poles, radii
8-20 nm, seed 2) in an eps=1.77 background, 12-cell CPML on all six faces and
10-cell pads (the single-particle driver's extension) -> a 300^3 full grid,
TFSF plane wave fed by the 1D auxiliary grid, box spectral-Poynting sensor on
the structure boundary at 64 wavelengths.  FP32 field storage and arithmetic,
FP64 recursive-convolution poles and DFT accumulators.  Synthetic structure,
no files.

A "step" is one full time step of the single-particle driver loop: auxiliary
E, main E (stencil + TFSF + RC + CPML), auxiliary H, main H, sensor sample and
(every 64 steps) the staged DFT.  value = full-grid cells x steps / seconds.

Also reported: the roofline of the dominant kernel (E pass) from HIP events
over the timed region, and the CPU baseline: the oracle (this project's C++
FP64 restatement of Aether's CPU solver, threaded as Aether) timed on this
host on a bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp32|fp64]
       [--no-cpu] [--cpu-steps S] [--form plain|precision|batching|serving|setup|slab]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, 8 TB/s
KS = 64                 # steps per batch chunk (and the batches' DFT flush)


def stage_depth(prec):
    """DFT staging depth of a grid stepped on its own (steps per accumulator
    flush): 512 for FP32 arithmetic, 256 for FP64."""
    from aether_amd import capi
    return 256 if prec == capi.FP64 else 512


def host_cpu_share():
    """Threads for the CPU baseline: the CPUs this process may run on (its
    affinity), capped by a cgroup CPU quota when one is set; with the host's
    CPU model, for the record."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(per)
    except (OSError, ValueError):
        pass
    threads = max(1, min(n, int(quota + 0.5))) if quota else n
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    note = f"{model}; affinity {n} CPUs, cgroup quota {quota if quota else 'none'}, os.cpu_count {os.cpu_count()}"
    return threads, note


def dft_bytes_per_step(spec, sensors, lay, stats, KD):
    """Algorithmic bytes per step of the sensor path: every plane cell's 6
    staged samples written once and read once, and the FP64 complex
    accumulators (16 B) of 6 components x Nl read and written once per KD-step
    block."""
    s = stats[3]
    tot = 0
    for q in sensors:
        nx, ny, nz = q.x2 - q.x1, q.y2 - q.y1, q.z2 - q.z1
        planes = [ny * nz, ny * nz, nx * nz, nx * nz, nx * ny, nx * ny] if q.box else [max(1, nx) * max(1, ny) * max(1, nz)]
        for P in planes:
            tot += P * 6 * 2 * s + P * 6 * q.Nl * 32 / KD
    return tot


def dft_flush_bytes(spec, sensors, KD, stage_bytes):
    """Algorithmic bytes of one DFT flush of KD staged steps: each plane cell's
    6 x KD staged samples read once and its 6 x Nl FP64 complex accumulators
    (16 B) read and written once."""
    tot = 0
    for q in sensors:
        nx, ny, nz = q.x2 - q.x1, q.y2 - q.y1, q.z2 - q.z1
        planes = [ny * nz, ny * nz, nx * nz, nx * nz, nx * ny, nx * ny] if q.box else [max(1, nx) * max(1, ny) * max(1, nz)]
        for P in planes:
            tot += P * 6 * (KD * stage_bytes + q.Nl * 32)
    return tot


def e_pass_bytes(lay, pml, stats):
    """Algorithmic HBM bytes of one E-pass launch (k_update_E): read E,H (6
    values) and the uint8 material id, write E (3 values); CPML slab cells
    read+write 2 psi per PML axis (E ranges: 1..pm-1 and N-pp+1..N-1);
    dispersive cells read a 4-byte slot index and the 3-value RC increment."""
    ncell, ndisp, npmax, s = stats
    Nx, Ny, Nz = lay[:3]
    b = ncell * (9 * s + 1)
    cross = (Ny * Nz, Nx * Nz, Nx * Ny)
    for a in range(3):
        pm, pp = pml[2 * a], pml[2 * a + 1]
        if pm or pp:
            planes = max(0, pm - 1) + max(0, pp - 1)
            b += planes * cross[a] * 2 * 2 * s
    b += ndisp * (4 + 3 * s)
    return b


def one_pass_bytes(e_bytes, h_bytes, stats):
    """Algorithmic HBM bytes of one one-pass step launch (k_step): the two
    passes' bytes less the second reads of E and H (each field is read once
    and written once, the material id read once; CPML psi and the dispersive
    increments as in the passes)."""
    ncell, _, _, s = stats
    return e_bytes + h_bytes - ncell * 6 * s


def rc_pass_bytes(stats):
    """RC pre-pass (k_rc_ante): per dispersive cell read cell index, material id
    and E (3), read+write 3 complex FP64 accumulators per pole, write 3 values."""
    ncell, ndisp, npmax, s = stats
    return ndisp * (4 + 1 + 3 * s + npmax * 96 + 3 * s)


def h_pass_bytes(lay, pml, stats):
    """H pass: read H,E, write H (no material id); CPML H ranges 0..pm-1 and N-pp..N-2."""
    ncell, _, _, s = stats
    Nx, Ny, Nz = lay[:3]
    b = ncell * 9 * s
    cross = (Ny * Nz, Nx * Nz, Nx * Ny)
    for a in range(3):
        pm, pp = pml[2 * a], pml[2 * a + 1]
        if pm or pp:
            planes = pm + max(0, pp - 1)
            b += planes * cross[a] * 2 * 2 * s
    return b


def pmc_traffic(prec, n, kernel="k_update_E"):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE runs, FETCH scaled by the copy-kernel calibration factor) on this
    workload; read from the committed counter summary, None when there is none."""
    from aether_amd import capi
    name = {capi.FP32: "fp32", capi.FP64: "fp64"}.get(prec)
    path = os.path.join(ROOT, "profiles", f"{kernel}_single_particle_256_{name}_pmc.txt")
    if name and n == 256 and os.path.exists(path):
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(float(line.split(":")[1]))
    return None


def cpu_baseline(spec, sensors, steps_lo, steps_hi, threads):
    """Times the oracle's single-particle driver at two run lengths on the same
    workload in FP64; the difference is the stepping cost (setup cancels)."""
    import oracle
    from aether_amd import capi
    be = oracle.load()
    be.fn("set_threads")(threads)
    import copy
    times = []
    for n in (steps_lo, steps_hi):
        s = copy.copy(spec)
        s.Nt = n
        s.precision = capi.FP64
        t0 = time.perf_counter()
        capi.run_single_particle(be, s, sensors)
        times.append(time.perf_counter() - t0)
    dt = (times[1] - times[0]) / (steps_hi - steps_lo)
    return dt, times


def precision_form(args):
    """The precision form: one multilayer normal-incidence stack (12 random layers +
    a Lorentz layer, 512x512x384 structure, 16-cell CPML, 128 wavelengths) run with
    FP64, FP32 and FP16 field storage; |r| and |t| spectra of FP32/FP16 are compared
    with the FP64 run (which matches the CPU oracle to <=1e-9, see the GPU tests)."""
    import numpy as np
    from aether_amd import capi, workloads
    out = {}
    for name, prec in (("f64", capi.FP64), ("f32", capi.FP32), ("f16", capi.FP16)):
        spec, info = workloads.multilayer_normal_incidence(Nt=args.nt, precision=prec)
        run = capi.Run(capi.engine(), 0, spec)
        ncell = run.stats()[0]
        run.synchronize()
        t0 = time.perf_counter()
        run.step(args.nt)
        run.synchronize()
        wall = time.perf_counter() - t0
        res = run.finish()
        run.close()
        out[name] = {"R": res["R"], "T": res["T"], "wall": wall, "mcells": ncell * args.nt / wall / 1e6}
        print(f"precision form {name}: {wall:.2f}s, {out[name]['mcells']:.0f} Mcell/s", file=sys.stderr, flush=True)
    rec = {"form": "precision", "workload": f"multilayer 512x512x384 (13 layers incl. Lorentz), "
                                            f"{ncell} cells, {args.nt} steps, 128 wavelengths",
           "thickness_scale": info["thickness_scale"]}
    for name in ("f64", "f32", "f16"):
        e = {"Mcell_updates_per_s": round(out[name]["mcells"], 1), "wall_s": round(out[name]["wall"], 2)}
        if name != "f64":
            e["max_abs_err_R"] = float(np.max(np.abs(out[name]["R"] - out["f64"]["R"])))
            e["max_abs_err_T"] = float(np.max(np.abs(out[name]["T"] - out["f64"]["T"])))
        rec[name] = e
    print(json.dumps(rec))


def batching_form(args):
    """The batching form: 64 independent 2D TE/TM runs (seed 4; sides
    log-uniform in 256..2048, dielectric cylinders, 16-cell CPML, oscillator
    line source, 4000-12000 steps, 64-wavelength transmission line sensor).
    Timed: (a) all runs in one FP32 batch (one launch per half step for every
    run, 64-step chunks replayed from a HIP graph, tiles chosen by the online
    tuner inside the timed region); (b) the same runs one at a time in FP64
    and (c) one at a time in FP32, each on the same GPU.  Device buffers are
    allocated and inputs uploaded before each timed region; the timed region
    covers stepping and the spectra.  runs/s = runs / wall seconds."""
    import numpy as np
    from aether_amd import capi, workloads
    be = capi.engine()
    t0 = time.perf_counter()
    runs = workloads.sweep_runs(args.runs, seed=4, precision=capi.FP32)
    specs = [(capi.mode_finalize(be, s, capi.KIND_CUSTOM, q), q, info) for s, q, info in runs]
    gen_s = time.perf_counter() - t0
    print(f"batching form: {len(specs)} runs generated in {gen_s:.1f}s", file=sys.stderr, flush=True)
    cell_steps = 0
    for spec, _, _ in specs:
        Nx, Ny, Nz = capi.custom_dims(spec)
        cell_steps += Nx * Ny * Nz * spec.Nt

    def with_prec(spec, prec):
        return capi.ModeSpec(**{**spec.__dict__, "precision": prec})

    out = {}
    # (a) one FP32 batch
    objs = [capi.Run(be, capi.KIND_CUSTOM, with_prec(sp, capi.FP32), q) for sp, q, _ in specs]
    b = capi.Batch(be, capi.FP32)
    for r in objs:
        b.add(r)
    b.synchronize()
    t0 = time.perf_counter()
    tiles = b.autotune()
    res_b = [None] * len(objs)
    live = list(range(len(objs)))
    while live:
        b.step(1)
        for i in [i for i in live if objs[i].progress()[0] >= objs[i].progress()[1]]:
            b.remove(objs[i])               # finished runs leave at once: no idle workgroups
            res_b[i] = objs[i].finish()["sensors"]
            live.remove(i)
    wall = time.perf_counter() - t0
    for r in objs:
        r.close()
    b.close()
    out["batched_fp32"] = {"wall_s": round(wall, 3), "runs_per_s": round(len(specs) / wall, 3),
                           "Mcell_updates_per_s": round(cell_steps / wall / 1e6, 1), "tiles": tiles}
    print(f"batched fp32: {wall:.2f}s", file=sys.stderr, flush=True)

    def one_at_a_time(prec):
        tot = 0.0
        res = []
        for sp, q, _ in specs:
            r = capi.Run(be, capi.KIND_CUSTOM, with_prec(sp, prec), q)
            r.synchronize()
            t1 = time.perf_counter()
            r.step(sp.Nt)
            res.append(r.finish()["sensors"])
            tot += time.perf_counter() - t1
            r.close()
        return tot, res

    for name, prec in (("alone_fp64", capi.FP64), ("alone_fp32", capi.FP32)):
        wall_a, res = one_at_a_time(prec)
        out[name] = {"wall_s": round(wall_a, 3), "runs_per_s": round(len(specs) / wall_a, 3),
                     "Mcell_updates_per_s": round(cell_steps / wall_a / 1e6, 1)}
        if prec == capi.FP64:
            res64 = res
        else:
            res32 = res
        print(f"{name}: {wall_a:.2f}s", file=sys.stderr, flush=True)
    err_b = [float(np.max(np.abs(a - r)) / np.max(np.abs(r))) for a, r in zip(res_b, res64)]
    same = [float(np.max(np.abs(a - r)) / np.max(np.abs(r))) for a, r in zip(res_b, res32)]
    rec = {"form": "batching",
           "workload": f"{len(specs)} 2D TE/TM runs (seed 4), sides 256-2048 log-uniform, cylinders, "
                       f"16-cell CPML, oscillator line source, 4000-12000 steps, 64-lambda line sensor",
           "cell_steps": cell_steps, "setup_s": round(gen_s, 1), **out,
           "speedup_vs_alone_fp64": round(out["alone_fp64"]["wall_s"] / out["batched_fp32"]["wall_s"], 3),
           "speedup_vs_alone_fp32": round(out["alone_fp32"]["wall_s"] / out["batched_fp32"]["wall_s"], 3),
           "spectral_err_vs_fp64": {"max": max(err_b), "median": float(np.median(err_b))},
           "batched_vs_alone_fp32_max_rel_diff": max(same)}
    print(json.dumps(rec))


def serving_form(args):
    """The serving form: a Poisson stream (4 requests/s, seed 5) of mixed 3D
    single-particle and 2D custom requests (FP32 fields, FP64 sensors) sent to
    the engine's request server with a 200 GB device-memory budget; reports
    time-to-spectrum p50/p99 (submit to spectra out, measured by the server)
    and runs/s (requests / (last spectrum - first arrival)).  Then the same
    requests one at a time in FP64 (the baseline: create, step, spectra per
    request, back to back), and two requests served alone to show that a
    result does not depend on the mix."""
    import numpy as np
    from aether_amd import capi, workloads
    be = capi.engine()
    t0 = time.perf_counter()
    reqs = workloads.serving_requests(args.requests, seed=5)
    reqs = [(t, k, capi.mode_finalize(be, sp, k, q), q, info) for t, k, sp, q, info in reqs]
    gen_s = time.perf_counter() - t0
    print(f"serving form: {len(reqs)} requests generated in {gen_s:.1f}s", file=sys.stderr, flush=True)
    cell_steps = 0
    for _, k, sp, q, info in reqs:
        Nx, Ny, Nz = (capi.single_particle_dims(sp) if k == capi.KIND_SINGLE else capi.custom_dims(sp))
        cell_steps += Nx * Ny * Nz * sp.Nt
    budget = int(args.budget_gb * 1e9)
    srv = capi.Server(be, budget)
    ids = []
    start = time.perf_counter()
    for t, k, sp, q, info in reqs:
        now = time.perf_counter() - start
        if t > now:
            time.sleep(t - now)
        ids.append(srv.submit(k, sp, q))
    outs, lat, waits = [], [], []
    for rid in ids:
        st = srv.wait(rid)
        if st != capi.REQ_DONE:
            raise RuntimeError(f"request {rid} ended in state {st}: {srv.error(rid)}")
        o, tm = srv.result(rid)
        outs.append(o)
        lat.append(tm[2])
        waits.append(tm[0])
    end = time.perf_counter()
    stats = srv.stats()
    srv.close()
    makespan = end - start
    print(f"served: {makespan:.2f}s", file=sys.stderr, flush=True)
    # one at a time, FP64
    base = None
    if not args.no_baseline:
        tb = 0.0
        for _, k, sp, q, info in reqs:
            s64 = capi.ModeSpec(**{**sp.__dict__, "precision": capi.FP64})
            r = capi.Run(be, k, s64, q)
            r.synchronize()
            t1 = time.perf_counter()
            r.step(s64.Nt)
            r.finish()
            tb += time.perf_counter() - t1
            r.close()
        base = {"wall_s": round(tb, 3), "runs_per_s": round(len(reqs) / tb, 4)}
        print(f"one at a time fp64: {tb:.2f}s", file=sys.stderr, flush=True)
    # alone == mixed: the smallest 3D and the smallest 2D request, each on a fresh server
    picks = []
    for kind in (capi.KIND_SINGLE, capi.KIND_CUSTOM):
        cand = [(i, r) for i, r in enumerate(reqs) if r[1] == kind]
        if cand:
            picks.append(min(cand, key=lambda c: c[1][4]["side"])[0])
    diff = 0.0
    for i in picks:
        _, k, sp, q, info = reqs[i]
        s1 = capi.Server(be, budget)
        rid = s1.submit(k, sp, q)
        s1.wait(rid)
        o, _ = s1.result(rid)
        s1.close()
        diff = max(diff, float(np.max(np.abs(o - outs[i])) / max(1e-300, np.max(np.abs(outs[i])))))
    lat = np.array(lat)
    rec = {"form": "serving",
           "workload": f"{len(reqs)} requests, Poisson 4/s (seed 5): 3D single-particle side 96-384 "
                       f"(Drude+CP or Lorentz spheres), 2D custom side 512-4096 (cylinders), 2000-10000 steps, FP32",
           "requests": len(reqs), "n_3d": sum(1 for r in reqs if r[1] == capi.KIND_SINGLE),
           "cell_steps": cell_steps, "budget_GB": args.budget_gb,
           "peak_resident_GB": round(stats["peak_bytes"] / 1e9, 2), "rounds": int(stats["rounds"]),
           "time_to_spectrum_s": {"p50": round(float(np.percentile(lat, 50)), 3),
                                  "p99": round(float(np.percentile(lat, 99)), 3),
                                  "max": round(float(lat.max()), 3)},
           "queue_wait_s_p50": round(float(np.percentile(waits, 50)), 3),
           "runs_per_s": round(len(reqs) / makespan, 4),
           "Mcell_updates_per_s": round(cell_steps / makespan / 1e6, 1),
           "one_at_a_time_fp64": base,
           "speedup_vs_one_at_a_time_fp64": round(base["wall_s"] / makespan, 3) if base else None,
           "alone_vs_mixed_max_rel_diff": diff,
           "setup_s": round(gen_s, 1)}
    print(json.dumps(rec))


def slab_form(args):
    """The 64^3 slab configuration in full (normal incidence, 64x64x90 grid,
    2000 steps, 50 wavelengths, FP64): the engine's driver end to end (run
    creation, stepping, R/T transforms) against the oracle's FP64 driver on the
    host (all host threads of the oracle's pool), and their largest R/T
    difference."""
    import numpy as np
    from aether_amd import capi, workloads
    be = capi.engine()
    spec = workloads.slab_normal_incidence()
    capi.run_normal_incidence(be, spec)                 # warm
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        g = capi.run_normal_incidence(be, spec)
        ts.append(time.perf_counter() - t0)
    cells = 64 * 64 * 90 * spec.Nt
    rec = {"form": "slab", "workload": "64^3 slab (eps 2.25, z cells 22-42), 8-cell CPML +-z, 2000 steps, "
                                       "50 wavelengths, FP64, normal incidence",
           "gpu_wall_s": round(min(ts), 4), "gpu_Mcell_updates_per_s": round(cells / min(ts) / 1e6, 1)}
    if not args.no_cpu:
        import oracle
        ob = oracle.load()
        t0 = time.perf_counter()
        r = capi.run_normal_incidence(ob, spec)
        dt = time.perf_counter() - t0
        rec.update(cpu_wall_s=round(dt, 3), cpu_Mcell_updates_per_s=round(cells / dt / 1e6, 1),
                   cpu_threads=int(ob.fn("get_threads")()),
                   max_abs_diff_RT=float(max(np.max(np.abs(g[k] - r[k])) for k in ("R", "T"))))
    print(json.dumps(rec))


SCRIPT_STRUCTURE = """lx = 256e-9
ly = 256e-9
lz = 256e-9
default_material(0)
for i = 1, 3 do
    add_sphere(64e-9 * i, 128e-9, 100e-9 + 20e-9 * i, 30e-9, 1)
end
"""
SCRIPT_MODE = """st = Structure("structure.lua")
gold = Material()
gold:epsilon_infty(1.1)
gold:add_drude(1.37e16, 1.07e14)
box = create_sensor("box_spectral_poynting")
box:location(0, 256e-9, 0, 256e-9, 0, 256e-9)
box:spectrum(400e-9, 800e-9, 64)
box:name("box")
f = MODE("fdtd_single_particle")
f:structure(st)
f:Dxyz(2e-9)
f:N_tsteps(200)
f:polarization("TM")
for _, face in ipairs({"xm", "xp", "ym", "yp", "zm", "zp"}) do f["pml_" .. face](f, 12, 25, 1, 0.2) end
f:material(0, Material(1.33))
f:material(1, gold)
f:register_sensor(box)
f:compute()
"""


def script_front_cost(be):
    """A single-particle script (128^3 cells of 2 nm, three Drude spheres, box
    sensor, 200 steps, FP32) through the script front end: the whole run, and
    its parts before the time loop (interpretation, device discretisation,
    sizing) and the run itself."""
    import tempfile
    from aether_amd import capi
    from aether_amd.script import Front
    with tempfile.TemporaryDirectory() as d:
        for name, text in (("structure.lua", SCRIPT_STRUCTURE), ("mode.lua", SCRIPT_MODE)):
            with open(os.path.join(d, name), "w") as fh:
                fh.write(text)
        Front(be, out_dir=d, precision=capi.FP32).run_file(os.path.join(d, "mode.lua"))      # warm
        t0 = time.perf_counter()
        dry = Front(be, out_dir=d, precision=capi.FP32, dry=True)
        dry.run_file(os.path.join(d, "mode.lua"))
        t1 = time.perf_counter()
        Front(be, out_dir=d, precision=capi.FP32).run_file(os.path.join(d, "mode.lua"))
        t2 = time.perf_counter()
    return {"inputs_ms": round(1e3 * (t1 - t0), 1), "whole_script_ms": round(1e3 * (t2 - t1), 1)}


def setup_form(args):
    """The setup form: what a run costs before and around its time loop.
    (1) Structure::discretize of the headline structure (256^3 cells, 40
    spheres) through aefdtd_discretize (allocation, kernel, copy of the uint32
    grid to the host; median of 5 after one warm call) against the serial loop
    of the oracle (the reference's algorithm, one thread, one call); (2) run
    creation (aefdtd_run_create of the single-particle driver: grid, material
    upload, CPML, aux column, sensors) at 64^3, 128^3 and 256^3 structures, warm
    process, median of 3; (3) the energy completion check (auto_tsteps, checked
    every 500 steps as the reference's default) on the headline 300^3 run: ms per
    step with and without it over 200 steps after 100 warm-up steps."""
    import statistics
    import numpy as np
    from aether_amd import capi, workloads
    be = capi.engine()
    n, D = 256, 2e-9
    sph = workloads.random_spheres(2, 40, (n * D,) * 3, 8e-9, 20e-9)
    st = capi.Structure(n * D, n * D, n * D)
    for (x, y, z, r, m) in sph:
        st.add_sphere(x, y, z, r, m)
    g = capi.discretize(be, st, (n, n, n), (D, D, D))
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        g = capi.discretize(be, st, (n, n, n), (D, D, D))
        ts.append(time.perf_counter() - t0)
    disc = {"engine_ms": round(1e3 * statistics.median(ts), 2)}
    if not args.no_cpu:
        import oracle
        t0 = time.perf_counter()
        o = capi.discretize(oracle.load(), st, (n, n, n), (D, D, D))
        disc["oracle_serial_ms"] = round(1e3 * (time.perf_counter() - t0), 1)
        disc["identical"] = bool(np.array_equal(o, g))
    print(f"setup form: discretize {disc}", file=sys.stderr, flush=True)
    create = {}
    for side in (64, 128, 256):
        spec, sensors, _ = workloads.spheres_single_particle(n=side, count=max(2, 40 * side ** 3 // 256 ** 3),
                                                             Nt=1000, precision=capi.FP32)
        cs = []
        for _ in range(3):
            t0 = time.perf_counter()
            r = capi.Run(be, capi.KIND_SINGLE, spec, sensors)
            r.synchronize()
            cs.append(time.perf_counter() - t0)
            r.close()
        create[str(side)] = round(1e3 * statistics.median(cs), 1)
    print(f"setup form: run creation ms {create}", file=sys.stderr, flush=True)
    spec, sensors, _ = workloads.spheres_single_particle(n=256, Nt=400, precision=capi.FP32)
    per = {}
    for tt in (0, 1):
        sp = capi.ModeSpec(**{**spec.__dict__, "time_type": tt, "cc_step": 500, "cc_coeff": 0.0})
        r = capi.Run(be, capi.KIND_SINGLE, sp, sensors)
        r.step(100)
        r.synchronize()
        t0 = time.perf_counter()
        r.step(200)
        r.synchronize()
        per[tt] = (time.perf_counter() - t0) / 200 * 1e3
        r.close()
    # result files (64 wavelengths, box sensor) and the sizing of FDTD_Mode::finalize
    import tempfile
    spec_w, sens_w, _ = workloads.spheres_single_particle(n=64, count=4, Nt=20, precision=capi.FP32)
    t0 = time.perf_counter()
    for _ in range(100):
        capi.mode_finalize(be, spec_w, capi.KIND_SINGLE, sens_w)
    finalize_us = (time.perf_counter() - t0) / 100 * 1e6
    r = capi.Run(be, capi.KIND_SINGLE, spec_w, sens_w)
    r.step(spec_w.Nt)
    r.finish()
    with tempfile.TemporaryDirectory() as d:
        t0 = time.perf_counter()
        r.write(d, "", ["box"])
        write_ms = (time.perf_counter() - t0) * 1e3
    r.close()
    script = script_front_cost(be)
    print(f"setup form: script front end {script}", file=sys.stderr, flush=True)
    rec = {"form": "setup", "discretize_256_40_spheres": disc, "run_create_ms": create, "script_front_end": script,
           "mode_finalize_us": round(finalize_us, 1), "write_box_64_lambda_ms": round(write_ms, 2),
           "energy_check": {"ms_per_step_fixed": round(per[0], 4), "ms_per_step_energy": round(per[1], 4),
                            "overhead": round(per[1] / per[0] - 1.0, 4)}}
    print(json.dumps(rec))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=640)
    ap.add_argument("--warmup", type=int, default=192)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--n", type=int, default=256, help="structure cells per side")
    ap.add_argument("--prof-steps", type=int, default=100,
                    help="steps of the profiled segment after the timed region (per-kernel averages)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 leg of the headline workload")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile shapes")
    ap.add_argument("--cpu-steps", type=int, default=200,
                    help="CPU baseline sample: (52 + S)-step run minus 52-step run of the oracle (the shorter run "
                         "holds the setup and 50 warm-up steps)")
    ap.add_argument("--step", type=int, default=None,
                    help="one-pass step tile code (-1: the E and H passes); with --tiles or alone it skips the tuner")
    ap.add_argument("--tiles", type=int, nargs=2, default=None, metavar=("E", "H"),
                    help="fixed tile configurations (0..31) instead of the autotuner's")
    ap.add_argument("--form", default="plain", choices=["plain", "precision", "batching", "serving", "setup", "slab"],
                    help="plain: the headline single-particle run; precision: the 512x512x384 multilayer "
                         "stack in FP64, FP32 and FP16 storage with spectral errors against FP64; "
                         "batching: 64 2D runs batched in FP32 against one-at-a-time FP64 and FP32; "
                         "serving: a Poisson request stream through the request server; "
                         "setup: discretiser, run creation and the energy completion check; "
                         "slab: the 64^3 slab configuration in full, FP64, engine against the oracle")
    ap.add_argument("--runs", type=int, default=64, help="runs of the batching form")
    ap.add_argument("--requests", type=int, default=24, help="requests of the serving form")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="device-memory budget of the serving form")
    ap.add_argument("--no-baseline", action="store_true", help="serving form without the FP64 one-at-a-time leg")
    ap.add_argument("--nt", type=int, default=8000, help="time steps of the precision form")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="only run a plain device copy of this many GiB (rocprof counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(the engine runs one simulation per GPU with no collectives)", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration > 0:
        import torch
        n = int(args.copy_calibration * (1 << 30)) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": n * 4, "launches": 5}))
        return

    import numpy as np
    from aether_amd import capi, workloads
    if args.form == "precision":
        return precision_form(args)
    if args.form == "batching":
        return batching_form(args)
    if args.form == "serving":
        return serving_form(args)
    if args.form == "setup":
        return setup_form(args)
    if args.form == "slab":
        return slab_form(args)
    prec = capi.FP32 if args.precision == "fp32" else capi.FP64
    # whole DFT staging blocks: the warm-up ends and the timed region starts
    # with an empty stage, and the timed region holds K/KD flushes
    KD = stage_depth(prec)
    K = -(-max(1, args.steps) // KD) * KD
    W = -(-max(0, args.warmup) // KD) * KD
    P = -(-max(1, args.prof_steps) // KD) * KD
    main_rec = headline_run(args, prec, W, K, P)
    info = main_rec.pop("_info")
    rec = {
        "metric": "Mcell-updates/s (256^3 Drude+CP spheres, CPML, TFSF, 64-wavelength box DFT)",
        "value": main_rec["value"],
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": main_rec["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == capi.FP32 else "f64",
        "data": "synthetic (seeded random sphere packing, no files)",
        "config": {"workload": info["workload"], "cells": info["cells"], "dispersive_cells": info["ndisp"],
                   "poles": info["poles"], "parallelism": "single GPU",
                   "timed_window": f"{K} steps = {K // KD} whole {KD}-step DFT blocks (requested {args.steps}), "
                                   f"after {W} warm-up steps (requested {args.warmup})"},
    }
    rec.update({k: v for k, v in main_rec.items() if k not in ("value", "ms_per_step")})
    if prec == capi.FP32 and not args.no_fp64:
        # the same workload with FP64 storage and arithmetic (the reference's precision)
        KD64 = stage_depth(capi.FP64)
        f64 = headline_run(args, capi.FP64, KD64, 2 * KD64, KD64)
        f64.pop("_info")
        rec["fp64"] = {k: f64[k] for k in ("value", "ms_per_step", "roofline", "kernels_ms_avg", "dft_roofline", "tiles",
                                             "step_tiles")}
        rec["fp64"]["steps"] = 2 * KD64
    if not args.no_cpu:
        threads, cpu_note = host_cpu_share()
        lo, hi = 52, 52 + args.cpu_steps
        spec, sensors = info["spec"], info["sensors"]
        dt, times = cpu_baseline(spec, sensors, lo, hi, threads)
        rec["cpu_baseline"] = {"value": round(info["cells"] / dt / 1e6, 2), "unit": "Mcell-updates/s",
                               "cores": threads, "kind": "port",
                               "sample": f"oracle FP64 single-particle driver on the same {info['dims']} workload, "
                                         f"{hi}-step run minus {lo}-step run ({times[1]:.1f}s vs {times[0]:.1f}s): "
                                         f"{hi - lo} steps after setup and 50 warm-up steps",
                               "host": cpu_note}
    print(json.dumps(rec))


def headline_run(args, prec, W, K, P):
    """One headline run: W warm-up steps, the tile autotuner, K timed steps with
    HIP events around the E and H pass launches on the engine's stream, then P
    steps with events around every launch for the other kernels' averages."""
    import numpy as np
    from aether_amd import capi, workloads
    t_setup = time.perf_counter()
    spec, sensors, spheres = workloads.spheres_single_particle(n=args.n, Nt=W + K + P, precision=prec)
    run = capi.Run(capi.engine(), 1, spec, sensors)
    setup_s = time.perf_counter() - t_setup
    lay = run.layout()
    stats = run.stats()
    ncell = stats[0]
    bw = np.zeros(3)
    capi.engine().fn("stream_bandwidth")(1 << 31, 3, bw.ctypes.data_as(capi._dp))
    bwm = np.zeros(2)
    capi.engine().fn("stream_mix_bandwidth")(128 << 20, 3, bwm.ctypes.data_as(capi._dp))
    run.step(W)
    be = capi.engine()
    if args.tiles is not None or args.step is not None:
        if args.tiles is not None:
            run.set_tiles(*args.tiles)
        be.check(be.fn("set_step_tiles")(run.grid(), -1 if args.step is None else args.step), run.grid())
        tiles = list(args.tiles) if args.tiles is not None else None
    else:
        tiles = run.autotune(3) if not args.no_tune else None
    code = capi.C.c_int(-1)
    be.fn("get_step_tiles")(run.grid(), capi.C.byref(code))
    step_code = code.value
    run.synchronize()
    run.profile(True, kinds=3)
    t0 = time.perf_counter()
    run.step(K)
    run.synchronize()
    wall = time.perf_counter() - t0
    e_ms, e_n = run.profile_read("E")
    h_ms, h_n = run.profile_read("H")
    run.profile(True)
    run.step(P)
    run.synchronize()
    s_ms, s_n = run.profile_read("sensor")
    d_ms, d_n = run.profile_read("dft")
    r_ms, r_n = run.profile_read("rc")
    run.profile(False)
    run.close()
    e_bytes = e_pass_bytes(lay, spec.pml, stats)
    h_bytes = h_pass_bytes(lay, spec.pml, stats)
    e_avg = e_ms / max(1, e_n) / 1e3
    h_avg = h_ms / max(1, h_n) / 1e3
    one_pass = step_code >= 0 and h_n == 0
    if one_pass:
        # the one-pass step (its launches are profiled as "E"): each field read
        # and written once, the material id read once
        kname, kbytes, kavg = "k_step", one_pass_bytes(e_bytes, h_bytes, stats), e_avg
        pass_bytes = kbytes
    elif h_avg > e_avg:
        # the roofline names the longer of the two passes
        kname, kbytes, kavg = "k_update_H", h_bytes, h_avg
        pass_bytes = e_bytes + h_bytes
    else:
        kname, kbytes, kavg = "k_update_E", e_bytes, e_avg
        pass_bytes = e_bytes + h_bytes
    achieved = kbytes / kavg / 1e9
    traffic = pmc_traffic(prec, args.n, kname)
    KD = stage_depth(prec)
    step_bytes = pass_bytes + rc_pass_bytes(stats) + dft_bytes_per_step(spec, sensors, lay, stats, KD)
    flush_bytes = dft_flush_bytes(spec, sensors, KD, 8 if prec == capi.FP64 else 4)
    flush_s = d_ms / max(1, d_n) / 1e3
    rec = {
        "value": round(ncell * K / wall / 1e6, 1),
        "ms_per_step": round(wall / K * 1e3, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": kname, "bytes_per_launch": kbytes, "avg_launch_ms": round(kavg * 1e3, 5),
                     **({"moved_GBps": round(traffic / kavg / 1e9, 1),
                         "moved_vs_stream_mix": round(traffic / kavg / 1e9 / float(bwm[0]), 4)}
                        if traffic and bwm[0] > 0 else {})},
        "kernels_ms_avg": {**({"step": round(e_avg * 1e3, 5)} if one_pass else
                              {"E": round(e_avg * 1e3, 5), "H": round(h_avg * 1e3, 5)}),
                           "sensor": round(s_ms / max(1, s_n), 5),
                           "dft_per_flush": round(d_ms / max(1, d_n), 5),
                           "dft_us_per_step": round(d_ms * 1e3 / max(1, P), 3),
                           "dft_steps_per_flush": KD,
                           "rc": round(r_ms / max(1, r_n), 5) if r_n else "in the sensor launch"},
        "dft_roofline": {"bound": "hbm", "achieved": round(flush_bytes / flush_s / 1e9, 1) if d_n else None,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(flush_bytes / flush_s / 1e9 / HBM_PEAK_GBS, 4) if d_n else None,
                         "kernel": "k_dft_mx" if prec != capi.FP64 else "k_dft_mx64", "bytes_per_flush": flush_bytes},
        "step_hbm_fraction": round(step_bytes / (wall / K) / 1e9 / HBM_PEAK_GBS, 4),
        **({"step_pass_GBps": round(kbytes / kavg / 1e9, 1)} if one_pass else
           {"e_pass_GBps": round(e_bytes / e_avg / 1e9, 1), "h_pass_GBps": round(h_bytes / h_avg / 1e9, 1)}),
        "stream_copy_GBps": {"plain": round(float(bw[0]), 1), "nontemporal": round(float(bw[1]), 1),
                             "note": "best 2 GiB device copy (read+write), the achievable-bandwidth reference"},
        "stream_mix_GBps": {"value": round(float(bwm[0]), 1),
                            "note": "best streaming kernel with the E/H passes' mix: six 128 MiB arrays read, "
                                    "three written, 4 B per lane, non-temporal"},
        "setup_s": round(setup_s, 2),
        "tiles": tiles,
        "step_tiles": step_code,
        "_info": {"workload": f"single-particle {args.n}^3 structure -> {lay[0]}x{lay[1]}x{lay[2]} grid, "
                              f"{len(spheres)} Drude+2CP spheres, 12-cell CPML x6, 64-lambda box DFT",
                  "cells": ncell, "ndisp": stats[1], "poles": stats[2], "spec": spec, "sensors": sensors,
                  "dims": f"{lay[0]}x{lay[1]}x{lay[2]}"},
    }
    return rec


if __name__ == "__main__":
    main()
