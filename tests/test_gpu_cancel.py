"""Cancellation: a run's cancel flag ends that run after the step in progress
(the end_computation poll of modes/normal_incidence.cpp:431-435,
single_particle.cpp:403, default_fdtd.cpp:201), the steps made are
post-processed, and no other run is touched.

Without tapering nothing in a run depends on its step count, so a run
cancelled during step t must give what a fixed-length run gives: the sensors
of a single-particle run that made t+1 steps, the R/T spectra of a
normal-incidence run of Nt = t (the reference sets Nt = t and breaks)."""
import ctypes as C

import numpy as np
import pytest

from aether_amd import capi, workloads

pytestmark = pytest.mark.gpu


def sp_spec(Nt):
    return workloads.spheres_single_particle(n=20, D=2e-9, count=3, rmin=3e-9, rmax=6e-9, seed=5, Nt=Nt, Nl=10,
                                             pml=5, precision=capi.FP64)


def finish_sensors(run):
    nsout = sum(q.out_len() for q in run.sensors)
    out = np.zeros(nsout)
    run.be.check(run.be.fn("run_finish")(run.h, None, out.ctypes.data_as(C.POINTER(C.c_double)), None), None)
    return out


def test_cancel_mid_run_keeps_partial_spectra(engine_be, oracle_be):
    """Cancel after 57 steps of a 150-step single-particle run: the next step is
    the last (58 steps), its box spectra equal a fixed 58-step run of the oracle;
    a second run alive at the same time is not affected."""
    print("cancel: single particle 20^3, 150 steps, cancelled during step 57", flush=True)
    spec, sens, _ = sp_spec(150)
    a = capi.Run(engine_be, 1, spec, sens)
    b = capi.Run(engine_be, 1, spec, sens)
    assert a.step(57)
    b.step(30)
    a.cancel()
    assert not a.step(10)                 # one step runs, then the flag ends the run
    assert a.progress() == (58, 58)
    assert not a.step(5)                  # a cancelled run stays ended
    assert b.step(150)                    # the other run goes on to its end
    assert b.progress() == (150, 150)
    got_a = finish_sensors(a)
    got_b = finish_sensors(b)
    spec58, _, _ = sp_spec(58)
    ref58 = capi.run_single_particle(oracle_be, spec58, sens)["sensors"]
    ref150 = capi.run_single_particle(oracle_be, spec, sens)["sensors"]
    assert np.max(np.abs(got_a - ref58)) <= 1e-9 * np.max(np.abs(ref58))
    assert np.max(np.abs(got_b - ref150)) <= 1e-9 * np.max(np.abs(ref150))


def test_one_call_driver_polls_the_callers_flag(engine_be, oracle_be):
    """A flag already set when the driver starts ends the run after its first
    step, in the engine and in the oracle alike; an unset flag changes nothing."""
    spec, sens, _ = sp_spec(80)
    flag = C.c_int(1)
    g = capi.run_single_particle(engine_be, spec, sens, end_flag=flag)
    r = capi.run_single_particle(oracle_be, spec, sens, end_flag=flag)
    spec1, _, _ = sp_spec(1)
    one = capi.run_single_particle(oracle_be, spec1, sens)["sensors"]
    assert g["cancelled"] and r["cancelled"]
    assert np.max(np.abs(g["sensors"] - one)) <= 1e-9 * max(1e-300, np.max(np.abs(one)))
    flag.value = 0
    g = capi.run_single_particle(engine_be, spec, sens, end_flag=flag)
    full = capi.run_single_particle(oracle_be, spec, sens)["sensors"]
    assert not g["cancelled"]
    assert np.max(np.abs(g["sensors"] - full)) <= 1e-9 * np.max(np.abs(full))


def test_cancelled_normal_incidence_uses_t_steps(engine_be, oracle_be):
    """Normal incidence cancelled during step 250 of 400: R/T and the complex
    amplitudes are those of a fixed run of Nt = 250 (Nt = t)."""
    spec = workloads.slab_normal_incidence(n=(8, 6, 40), slab=(12, 24), Nt=400, Nl=12, pml_z=8)
    run = capi.Run(engine_be, 0, spec)
    assert run.step(250)
    run.cancel()
    assert not run.step(1)
    Nl = spec.Nl
    bufs = {k: np.zeros(Nl) for k in ("lambda", "R", "T")}
    cplx = {k: np.zeros(2 * Nl) for k in ("refx", "refy", "transx", "transy")}
    res = capi.CNormalResult()
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    res.lambda_, res.R, res.T = dp(bufs["lambda"]), dp(bufs["R"]), dp(bufs["T"])
    res.refx, res.refy, res.transx, res.transy = (dp(cplx[k]) for k in ("refx", "refy", "transx", "transy"))
    engine_be.check(engine_be.fn("run_finish")(run.h, C.byref(res), None, None), None)
    assert res.Nt_run == 250
    spec250 = workloads.slab_normal_incidence(n=(8, 6, 40), slab=(12, 24), Nt=250, Nl=12, pml_z=8)
    r = capi.run_normal_incidence(oracle_be, spec250)
    assert r["Nt_run"] == 250
    for k in ("R", "T"):
        assert np.max(np.abs(bufs[k] - r[k])) <= 1e-9, k
    for k in ("refx", "refy", "transx", "transy"):
        got = cplx[k][0::2] + 1j * cplx[k][1::2]
        assert np.max(np.abs(got - r[k])) <= 1e-9, k
