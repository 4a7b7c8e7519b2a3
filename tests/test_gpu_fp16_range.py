"""FP16 storage range control: runs with FP16 storage check the largest
stored value of each field component every 64 steps; a component above 2^13
has its storage exponent lowered (its values rescaled by a power of two,
exactly) and a stored infinity or NaN ends the run with ERANGE."""
import ctypes as C

import numpy as np
import pytest

from aether_amd import capi
from aether_amd.capi import Material

pytestmark = pytest.mark.gpu


def _oscillator_spec(precision, Nt=900):
    """The custom 3D run of the FP16 storage test: an oscillator plane of
    amplitude 50 (sc_oscillator.cpp:37-38), a dielectric block, CPML, a planar
    sensor."""
    n = (20, 18, 16)
    mg = np.zeros(n[::-1], dtype=np.uint32)
    mg[6:10, 5:12, 7:14] = 1
    src = capi.SourceReq(orientation=capi.NORMAL_X, x1=0, x2=20, y1=0, y2=18, z1=2, z2=3,
                         lambda_min=400e-9, lambda_max=800e-9)
    sens = [capi.SensorReq(box=False, orientation=capi.NORMAL_Z, x1=0, x2=20, y1=0, y2=18, z1=15, z2=16, Nl=9)]
    spec = capi.ModeSpec(matsgrid=mg, materials=[Material(1.0), Material(4.0)], Nt=Nt, Dx=5e-9, Dy=5e-9,
                         Dz=5e-9, pml=(6,) * 6, pad=(5,) * 6, tapering=40, sources=[src], precision=precision)
    return capi.mode_finalize(capi.engine(), spec, capi.KIND_CUSTOM, sens), sens


def _scales(be, run):
    out = (C.c_int * 6)()
    assert be.fn("get_storage_scale")(run.grid(), out) == 0
    return list(out)


def test_fp16_exponent_lowered_instead_of_overflow(engine_be, oracle_be):
    """With the plane-wave storage scales (E 2^8, H 2^16) the fields near an
    oscillator source exceed the FP16 range (they overflowed to NaN before the
    range control existed); now the E exponents are lowered as the fields grow,
    the fields stay finite and the spectrum is within FP16's precision of the
    FP64 oracle."""
    be = engine_be
    spec, sens = _oscillator_spec(capi.FP16)
    run = capi.Run(be, capi.KIND_CUSTOM, spec, sens)
    exps = (C.c_int * 6)(8, 8, 8, 16, 16, 16)
    assert be.fn("set_storage_scale")(run.grid(), exps) == 0
    run.step(spec.Nt)
    after = _scales(be, run)
    got = run.finish(want_fields=True)
    run.close()
    print("storage exponents after the run:", after, flush=True)
    assert min(after[:3]) < 8
    assert np.all(np.isfinite(got["fields"]))
    r = capi.run_custom(oracle_be, capi.ModeSpec(**{**spec.__dict__, "precision": capi.FP64}), sens)
    peak = np.max(np.abs(r["sensors"]))
    err = np.max(np.abs(got["sensors"] - r["sensors"])) / peak
    print(f"FP16 with range control vs FP64 oracle: {err:.3e} of peak", flush=True)
    assert err <= 1e-2


def test_fp16_stored_infinity_ends_the_run(engine_be):
    """A value beyond the FP16 range written into the fields is stored as
    infinity; the next check ends the run with ERANGE naming the component."""
    be = engine_be
    spec, sens = _oscillator_spec(capi.FP16, Nt=200)
    run = capi.Run(be, capi.KIND_CUSTOM, spec, sens)
    g = run.grid()
    Nx, Ny, Nz = capi.custom_dims(spec)
    ey = np.zeros((Nz, Ny, Nx))
    ey[Nz // 2, Ny // 2, Nx // 2] = 1e7            # x 2^0 (the oscillator run's E scale): beyond 65504
    assert be.fn("write_field")(g, 1, ey.ctypes.data_as(C.POINTER(C.c_double))) == 0
    with pytest.raises(capi.FDTDError) as e:
        run.step(64)
    assert e.value.code == capi.ERANGE
    assert "FP16 storage overflow" in str(e.value)
    run.close()


def test_fp32_runs_are_not_checked(engine_be):
    """The range control is FP16 storage's alone: an FP32 run keeps its (zero)
    exponents and steps on."""
    spec, sens = _oscillator_spec(capi.FP32, Nt=128)
    run = capi.Run(engine_be, capi.KIND_CUSTOM, spec, sens)
    run.step(128)
    assert _scales(engine_be, run) == [0] * 6
    run.close()
