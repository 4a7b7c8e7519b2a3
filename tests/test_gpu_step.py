"""The one-pass step (aefdtd_step: k_step, fields of step n read from one set
of arrays and step n+1 written to the other) against the two passes
(k_update_E + k_update_H) and the oracle.  Both paths give every cell
the same inputs and operations, so the fields must agree bit for bit, in
FP64, FP32 and FP16 storage; FP64 also against the oracle at 1e-10."""
import numpy as np
import pytest

from aether_amd import capi, workloads
from aether_amd.capi import Grid, Material

pytestmark = pytest.mark.gpu

GOLD = workloads.GOLD
# 8 wavefronts per tile (6 rows owned), 16 planes, XCD fold
ONE_PASS = 1 | (1 << 2) | (1 << 4)
LORENTZ = Material(eps_inf=1.3, lorentz=[(1.5, 3e15, 1e14)])


def make_grid(be, n, pml, precision, mg, mats, Nt=40, D=(5e-9,) * 3):
    Dt = min(D) / (np.sqrt(3.0) * 299792458.0) * 0.99
    g = Grid(be, n, Nt, D, Dt, capi.MODE_CUSTOM, pml, (0,) * 6, precision)
    for f in range(6):
        g.set_pml(f)
    g.set_matsgrid(mg)
    for q, m in enumerate(mats):
        g.set_material(q, m)
    g.bootstrap()
    return g


def seeded_fields(g, seed):
    rng = np.random.default_rng(seed)
    for f in range(6):
        g.write(f, rng.standard_normal((g.Nz, g.Ny, g.Nx)) * (1.0 if f < 3 else 1 / 376.73))


CASES = [
    ("vacuum_periodic", (12, 10, 9), (0,) * 6),
    ("cpml_all_faces", (10, 11, 12), (4, 5, 3, 6, 4, 4)),
    ("cpml_z_only", (9, 8, 14), (0, 0, 0, 0, 5, 3)),
    ("kappa_quirk_xp_only", (10, 9, 8), (0, 5, 0, 0, 0, 0)),
    ("wide_x_two_wavefronts", (70, 5, 6), (3, 3, 2, 2, 2, 2)),
    ("three_wavefronts_ragged", (150, 23, 29), (6, 6, 4, 4, 6, 6)),
    ("one_row", (20, 1, 12), (3, 3, 0, 0, 3, 3)),
]


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32, capi.FP16], ids=["fp64", "fp32", "fp16"])
@pytest.mark.parametrize("name,n,pml", CASES, ids=[c[0] for c in CASES])
def test_step_bitwise_equals_two_passes(engine_be, name, n, pml, prec):
    print(f"one-pass step: {name} n={n}", flush=True)
    rng = np.random.default_rng(len(name))
    mg = rng.integers(0, 3, size=(n[2], n[1], n[0])).astype(np.uint32)
    mats = [Material(1.7), GOLD, LORENTZ]
    one = make_grid(engine_be, n, pml, prec, mg, mats)
    two = make_grid(engine_be, n, pml, prec, mg, mats)
    assert engine_be.fn("set_step_tiles")(one.h, ONE_PASS) == 0
    assert engine_be.fn("set_step_tiles")(two.h, -1) == 0
    seeded_fields(one, 3)
    seeded_fields(two, 3)
    assert engine_be.fn("step")(one.h, 9) == 0
    assert engine_be.fn("step")(two.h, 9) == 0
    for f in range(6):
        a, b = one.read(f), two.read(f)
        d = np.abs(a - b)
        if d.max() > 0:
            idx = np.argwhere(d > 0)
            print(f"field {f}: {len(idx)} cells differ, max {d.max():.3e}, rel-L2 "
                  f"{np.linalg.norm(a - b) / np.linalg.norm(b):.3e}, first (k,j,i) {idx[:6].tolist()}", flush=True)
    for f in range(6):
        assert np.array_equal(one.read(f), two.read(f)), f


def test_step_fp64_matches_oracle(engine_be, oracle_be):
    n, pml = (70, 13, 37), (3, 4, 2, 3, 5, 2)
    rng = np.random.default_rng(8)
    mg = rng.integers(0, 3, size=(n[2], n[1], n[0])).astype(np.uint32)
    mats = [Material(1.5), GOLD, LORENTZ]
    e = make_grid(engine_be, n, pml, capi.FP64, mg, mats)
    o = make_grid(oracle_be, n, pml, capi.FP64, mg, mats)
    assert engine_be.fn("set_step_tiles")(e.h, ONE_PASS) == 0
    seeded_fields(e, 5)
    seeded_fields(o, 5)
    assert engine_be.fn("step")(e.h, 12) == 0
    assert oracle_be.fn("step")(o.h, 12) == 0
    for f in range(6):
        a, b = e.read(f), o.read(f)
        assert np.linalg.norm(a - b) <= 1e-10 * np.linalg.norm(b), f


def valid_codes():
    """Step tile codes: NW index (4, 8, 16 wavefronts per tile, NW-2 rows owned)
    | KZ index (8, 16, 32, 64 planes) << 2 | XCD fold << 4; index 3 names no tile."""
    return [c for c in range(32) if c & 3 != 3]


def test_every_step_tile_code_bitwise(engine_be):
    """Every tile code (rows per tile, planes per tile, XCD fold) on a grid with
    CPML on all faces and dispersive cells."""
    n, pml = (130, 37, 41), (4, 5, 6, 3, 5, 4)
    rng = np.random.default_rng(21)
    mg = rng.integers(0, 3, size=(n[2], n[1], n[0])).astype(np.uint32)
    mats = [Material(2.0), GOLD, LORENTZ]
    ref = make_grid(engine_be, n, pml, capi.FP32, mg, mats)
    assert engine_be.fn("set_step_tiles")(ref.h, -1) == 0
    seeded_fields(ref, 1)
    ref.step(3)
    want = [ref.read(f) for f in range(6)]
    codes = valid_codes()
    print(f"{len(codes)} tile codes", flush=True)
    for c in codes:
        g = make_grid(engine_be, n, pml, capi.FP32, mg, mats)
        seeded_fields(g, 1)
        assert engine_be.fn("set_step_tiles")(g.h, c) == 0
        assert engine_be.fn("step")(g.h, 3) == 0
        for f in range(6):
            assert np.array_equal(g.read(f), want[f]), (c, f)
        g.close()


def test_drivers_on_the_one_pass_step(engine_be, oracle_be):
    """The single-particle driver (TFSF column, box sensor with strip planes,
    dispersive spheres, CPML) with the one-pass step forced, against the oracle."""
    import ctypes as C
    spec, sens, _ = workloads.spheres_single_particle(n=20, D=2e-9, count=3, rmin=3e-9, rmax=6e-9, seed=5,
                                                      Nt=150, Nl=10, pml=5, precision=capi.FP64)
    run = capi.Run(engine_be, 1, spec, sens)
    assert engine_be.fn("set_step_tiles")(run.grid(), ONE_PASS) == 0
    run.step(spec.Nt)
    got = run.finish(want_fields=True)
    ref = capi.run_single_particle(oracle_be, spec, sens, want_fields=True)
    err = np.linalg.norm(got["fields"] - ref["fields"]) / np.linalg.norm(ref["fields"])
    assert err <= 1e-10, err
    assert np.max(np.abs(got["sensors"] - ref["sensors"])) <= 1e-9 * np.max(np.abs(ref["sensors"]))
    run.close()
