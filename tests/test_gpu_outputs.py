"""Result files in the reference's formats (aefdtd_run_write).

The expected text is formed here from the values the run returned, with the
stream formatting the reference uses: setprecision(15) for the Poynting
sensors (sensors/s_spec_power.cpp:111,207-215), the default 6 significant
digits for spectdata_norm{,2} (modes/normal_incidence.cpp:505-557); both are
printf "%.Ng", which Python's format(x, ".Ng") reproduces digit for digit.
"""
import cmath
import os

import numpy as np
import pytest

from aether_amd import capi, workloads

pytestmark = pytest.mark.gpu


def g15(x):
    return format(float(x), ".15g")


def g6(x):
    return format(float(x), ".6g")


def lambdas(Nl, lmin, lmax):
    # lambda_grid (normal_incidence.cpp:235): lmin + (lmax - lmin) * l / (Nl - 1.0)
    return [lmin + (lmax - lmin) * l / (Nl - 1.0) for l in range(Nl)]


def read(path):
    with open(path, "rb") as f:
        return f.read().decode()


def test_normal_incidence_files(engine_be, tmp_path):
    spec = workloads.slab_normal_incidence(n=(8, 6, 40), slab=(12, 24), Nt=300, Nl=12, pml_z=8)
    sens = [capi.SensorReq(box=False, orientation=capi.NORMAL_Z, x1=0, x2=8, y1=0, y2=6, z1=30, z2=31, Nl=7,
                           lambda_min=450e-9, lambda_max=750e-9)]
    run = capi.Run(capi.engine(), capi.KIND_NORMAL, spec, sens)
    run.step(spec.Nt)
    out = run.finish()
    run.write(tmp_path, prefix="slab_", names=["tplane"])
    run.close()
    # planar sensor: lambda, flux, flux / (8 x 6 cells of Dx Dy)
    area = 8 * 6 * (spec.Dx * spec.Dy)
    L = lambdas(7, 450e-9, 750e-9)
    v = out["sensors"]
    want = "".join(f"{g15(L[l])} {g15(v[l])} {g15(v[l] / area)}\n" for l in range(7))
    assert read(tmp_path / "tplane_pspdft") == want
    # spectdata_norm: lambda R T
    want = "".join(f"{g6(out['lambda'][l])} {g6(out['R'][l])} {g6(out['T'][l])}\n" for l in range(12))
    assert read(tmp_path / "slab_spectdata_norm") == want
    # spectdata_norm2: |r| arg(r) |r|^2 for x, y, then |t| arg(t) |t|^2 n_sub/n_sup
    ratio = out["index_sub"] / out["index_sup"]
    lines = []
    for l in range(12):
        cols = [out["lambda"][l]]
        for k in ("refx", "refy"):
            z = complex(out[k][l])
            cols += [abs(z), cmath.phase(z), z.real * z.real + z.imag * z.imag]
        for k in ("transx", "transy"):
            z = complex(out[k][l])
            cols += [abs(z), cmath.phase(z), (z.real * z.real + z.imag * z.imag) * ratio]
        lines.append(" ".join(g6(c) for c in cols) + "\n")
    assert read(tmp_path / "slab_spectdata_norm2") == "".join(lines)
    # norm_octave_script (normal_incidence.cpp:24-59): loads spectdata_norm,
    # T scaled by n_sub / n_sup in the default 6 digits
    script = read(tmp_path / "slab_show_norm.m")
    assert script.startswith("function out=slab_show_norm(varargin)\n\nD=load('slab_spectdata_norm');\n")
    assert f"T={g6(out['index_sub'])}/{g6(out['index_sup'])}*t.^2;\n" in script
    assert script.endswith("if nargin>0\n    out=[L,R,T,A];\nend\n\nend\n")


def test_box_sensor_files(engine_be, tmp_path):
    spec, sens, _ = workloads.spheres_single_particle(n=24, count=2, rmin=6e-9, rmax=9e-9, Nt=60, Nl=5,
                                                     precision=capi.FP64)
    sens[0].disable = (0, 0, 1, 0, 0, 0)          # no y- face
    run = capi.Run(capi.engine(), capi.KIND_SINGLE, spec, sens)
    run.step(spec.Nt)
    out = run.finish()
    run.write(tmp_path, names=["scat"])
    run.close()
    v = out["sensors"]
    Nl = 5
    L = lambdas(Nl, 400e-9, 800e-9)
    lines = []
    for l in range(Nl):
        cols = [L[l], v[l]] + [v[(1 + f) * Nl + l] for f in range(6) if f != 2]
        lines.append(" ".join(g15(c) for c in cols) + "\n")
    assert read(tmp_path / "scat_bspdft") == "".join(lines)
    # the total is the sum of the enabled faces in the order xm xp ym yp zm zp
    tot = np.zeros(Nl)
    for f in (0, 1, 3, 4, 5):
        tot = tot + v[(1 + f) * Nl:(2 + f) * Nl]
    assert np.array_equal(tot, v[:Nl])
    # the Octave reader, with the file stem as std::filesystem::path prints it (quoted)
    assert read(tmp_path / "scat.m") == ('function [L,tot]="scat"\n   data=dlmread(\'"scat"_bspdft\');\n'
                                         "   L=data(:,1);\n   tot=data(:,2);\nend")
    assert not os.path.exists(tmp_path / "spectdata_norm")
