"""Batches of 3D runs: single-particle runs (TFSF plane wave fed by their
source column, which joins the batch with them) and dispersive custom runs
(RC pre-pass in its own batched launch) next to 2D runs.  Every member in
FP64 against the CPU oracle run of the same inputs (fields <= 1e-10 relative
L2, spectra <= 1e-9 of their peak), one of them after 64 steps on its own
(the pre-pass of its next step already ran in its last sensor feed); an FP32
single-particle member equals the same run stepped alone, field for field."""
import numpy as np
import pytest

from aether_amd import capi, workloads
from aether_amd.capi import Material

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-10


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(1e-300, np.linalg.norm(b))


def spec_err(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


def _single(seed, n, Nt, precision):
    spec, sens, _ = workloads.spheres_single_particle(n=n, D=2e-9, count=3, rmin=3e-9, rmax=5e-9, seed=seed, Nt=Nt,
                                                      Nl=8, pml=5, precision=precision)
    return spec, sens


def _custom_dispersive(Nt, precision):
    n = (20, 18, 16)
    mg = np.zeros(n[::-1], dtype=np.uint32)
    mg[6:10, 5:12, 7:14] = 1
    src = capi.SourceReq(orientation=capi.NORMAL_Y, x1=0, x2=20, y1=2, y2=3, z1=0, z2=16,
                         lambda_min=400e-9, lambda_max=800e-9)
    sens = [capi.SensorReq(box=True, x1=3, x2=17, y1=4, y2=15, z1=4, z2=15, Nl=9)]
    spec = capi.ModeSpec(matsgrid=mg, materials=[Material(1.0), workloads.GOLD], Nt=Nt, Dx=5e-9, Dy=5e-9,
                         Dz=5e-9, pml=(6,) * 6, pad=(5,) * 6, tapering=40, sources=[src], precision=precision)
    return capi.mode_finalize(capi.engine(), spec, capi.KIND_CUSTOM, sens), sens


def _custom_2d(seed, side, Nt, precision):
    rng = workloads.MT19937_64(seed)
    spec, sens, _ = workloads.sweep_run(rng, seed, side=side, steps=Nt, Nl=16, precision=precision)
    return capi.mode_finalize(capi.engine(), spec, capi.KIND_CUSTOM, sens), sens


def _drain(b, runs):
    out = [None] * len(runs)
    live = list(range(len(runs)))
    while live:
        b.step(1)
        for i in [i for i in live if runs[i].progress()[0] >= runs[i].progress()[1]]:
            b.remove(runs[i])
            out[i] = runs[i].finish(want_fields=True)
            live.remove(i)
    return out


def test_batch_3d_members_fp64_against_oracle(engine_be, oracle_be):
    be = engine_be
    cases = [("single", capi.KIND_SINGLE) + _single(5, 20, 200, capi.FP64),
             ("single_joined_late", capi.KIND_SINGLE) + _single(6, 16, 150, capi.FP64),
             ("custom_dispersive", capi.KIND_CUSTOM) + _custom_dispersive(130, capi.FP64),
             ("custom_2d", capi.KIND_CUSTOM) + _custom_2d(7, 40, 100, capi.FP64)]
    runs = [capi.Run(be, kind, spec, sens) for _, kind, spec, sens in cases]
    assert runs[1].step(64)                      # its last feed ran the next step's RC pre-pass
    b = capi.Batch(be, capi.FP64)
    for r in runs:
        b.add(r)
    outs = _drain(b, runs)
    b.close()
    for (name, kind, spec, sens), run, got in zip(cases, runs, outs):
        run.close()
        if kind == capi.KIND_SINGLE:
            ref = capi.run_single_particle(oracle_be, spec, sens, want_fields=True)
        else:
            ref = capi.run_custom(oracle_be, spec, sens, want_fields=True)
        ef, es = rel_l2(got["fields"], ref["fields"]), spec_err(got["sensors"], ref["sensors"])
        print(f"{name}: fields {ef:.3e}, spectra {es:.3e}", flush=True)
        assert np.max(np.abs(ref["sensors"])) > 0
        assert ef <= FIELD_TOL, name
        assert es <= 1e-9, name


def test_batch_single_particle_fp32_equals_alone(engine_be):
    be = engine_be
    spec, sens = _single(8, 18, 190, capi.FP32)
    alone = capi.Run(be, capi.KIND_SINGLE, spec, sens)
    alone.step(spec.Nt)
    a = alone.finish(want_fields=True)
    alone.close()
    run = capi.Run(be, capi.KIND_SINGLE, spec, sens)
    other = capi.Run(be, capi.KIND_CUSTOM, *_custom_dispersive(100, capi.FP32))
    b = capi.Batch(be, capi.FP32)
    b.add(run)
    b.add(other)
    m, _ = _drain(b, [run, other])
    b.close()
    run.close()
    other.close()
    assert np.array_equal(m["fields"], a["fields"])
    # the batch flushes its DFT every 64 steps on the vector ALU, a run alone every
    # 256 on the matrix cores: the same sums in another order
    assert spec_err(m["sensors"], a["sensors"]) <= 1e-5
