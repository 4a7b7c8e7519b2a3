"""Structure discretisation (Aether's Structure::discretize,
structure/structure.cpp:94-113,223-262, with the operations of
structure/structure_primitives.cpp).

CPU tests pin the oracle (oracle/structure_oracle.cpp) with closed forms on
unit cells, where the corner points are integers: lattice-point counts of
balls (1, 7, 33, 123, 257, 515: the number of integer points with
i^2+j^2+k^2 <= r^2) and discs, half-open blocks and layers, the later-wins
rule, periodic images against a rolled grid and flips against a mirrored
grid; and the Python sphere voxeliser of the benchmark inputs against the
oracle.  GPU tests hold aefdtd_discretize to the oracle bit for bit on random
mixes of every operation with periodic and flipped axes, and on the
256^3 sphere packing of the single-particle benchmark."""
import numpy as np
import pytest

from aether_amd import capi, workloads

BALL = {1: 7, 2: 33, 3: 123, 4: 257, 5: 515}       # integer points in the closed ball
DISC = {0: 1, 1: 5, 2: 13, 3: 29, 4: 49}            # integer points in the closed disc


def grid(be, s, N, D=(1.0, 1.0, 1.0)):
    return capi.discretize(be, s, N, D)


# ---------------------------------------------------------------- CPU: the oracle

@pytest.mark.parametrize("r", sorted(BALL))
def test_ball_lattice_counts(oracle_be, r):
    g = grid(oracle_be, capi.Structure(12, 12, 12).add_sphere(6, 6, 6, r, 3), (13, 13, 13))
    assert int((g == 3).sum()) == BALL[r]
    assert int((g == 0).sum()) == 13 ** 3 - BALL[r]


def test_ellipsoid_with_exact_arms_is_the_ball(oracle_be):
    # arms of 4: the divisions are exact, so the test is the ball's
    g = grid(oracle_be, capi.Structure().add_ellipsoid(6, 6, 6, 4, 4, 4, 1), (13, 13, 13))
    assert int(g.sum()) == BALL[4]
    g = grid(oracle_be, capi.Structure().add_ellipsoid(6, 6, 6, 4, 2, 1, 1), (13, 13, 13))
    i, j, k = np.meshgrid(np.arange(13) - 6, np.arange(13) - 6, np.arange(13) - 6, indexing="ij")
    want = (i * i) / 16 + (j * j) / 4 + k * k <= 1      # quarters and halves: exact in binary
    assert np.array_equal(g.transpose(2, 1, 0) == 1, want)


def test_block_and_layers_are_half_open(oracle_be):
    g = grid(oracle_be, capi.Structure().add_block(5, 2, 1, 4, 0, 3, 2), (8, 7, 6))   # corners swapped
    want = np.zeros((6, 7, 8), dtype=np.uint32)
    want[0:3, 1:4, 2:5] = 2
    assert np.array_equal(g, want)
    for axis, sl in (("x", np.s_[:, :, 2:5]), ("Y", np.s_[:, 2:5, :]), ("z", np.s_[2:5, :, :])):
        g = grid(oracle_be, capi.Structure().add_layer(axis, 2, 5, 4), (8, 7, 6))
        want = np.zeros((6, 7, 8), dtype=np.uint32)
        want[sl] = 4
        assert np.array_equal(g, want), axis


def test_cylinder_and_cone_counts(oracle_be):
    # along z from z=1 to z=7, radius 2: 13 disc points on 7 planes (both ends closed)
    g = grid(oracle_be, capi.Structure().add_cylinder(5, 5, 1, 0, 0, 6, 2, 1), (11, 11, 10))
    assert int(g.sum()) == DISC[2] * 7
    # cone of base radius 4 at z=0 to its tip at z=4: radius 4-h at height h
    g = grid(oracle_be, capi.Structure().add_cone(5, 5, 0, 0, 0, 4, 4, 1), (11, 11, 6))
    assert int(g.sum()) == sum(DISC[4 - h] for h in range(5))
    # the same cylinder along x: the disc points of the (y, z) plane
    g = grid(oracle_be, capi.Structure().add_cylinder(1, 5, 5, 6, 0, 0, 2, 1), (10, 11, 11))
    assert int(g.sum()) == DISC[2] * 7


def test_later_operations_win_and_default_fills(oracle_be):
    s = capi.Structure().default_material(7).add_sphere(5, 5, 5, 3, 1).add_block(5, 9, 0, 11, 0, 11, 2)
    g = grid(oracle_be, s, (11, 11, 11))
    x = np.arange(11)[None, None, :]
    assert np.all(g[:, :, 5:9] == 2)
    ball = grid(oracle_be, capi.Structure().add_sphere(5, 5, 5, 3, 1), (11, 11, 11))
    rest = np.broadcast_to((x < 5) | (x >= 9), g.shape)
    assert np.array_equal(g[rest], np.where(ball == 1, 1, 7)[rest])


def test_loop_wraps_images(oracle_be):
    # a ball centred on the origin of a looped 10^3 cell is the centred ball rolled by -5
    looped = capi.Structure(10, 10, 10).loop(1, 1, 1).add_sphere(0, 0, 0, 2.5, 1)
    centred = capi.Structure(10, 10, 10).add_sphere(5, 5, 5, 2.5, 1)
    a = grid(oracle_be, looped, (10, 10, 10))
    b = grid(oracle_be, centred, (10, 10, 10))
    # integer points with i^2+j^2+k^2 <= 6: 1 + 6 + 12 + 8 + 6 + 24 + 24
    assert int(a.sum()) == 81 and np.array_equal(a, np.roll(b, (-5, -5, -5), axis=(0, 1, 2)))
    # looping one axis only: the part beyond x=10 comes back at x=0, none along y/z
    s = capi.Structure(10, 10, 10).loop(1, 0, 0).add_block(8, 12, 2, 4, 2, 4, 1)
    g = grid(oracle_be, s, (10, 10, 10))
    assert np.array_equal(np.nonzero(g.sum(axis=(0, 1)))[0], [0, 1, 8, 9])
    assert int(g.sum()) == 4 * 2 * 2


def test_flip_mirrors(oracle_be):
    s = capi.Structure(10, 10, 10).add_block(1, 4, 2, 7, 3, 5, 1).add_sphere(7, 3, 6, 2, 2)
    plain = grid(oracle_be, s, (11, 11, 11))
    s.flip(1, 0, 1)
    flipped = grid(oracle_be, s, (11, 11, 11))
    # with lx = lz = 10 the flipped corner of cell i is 10 - i: an exact mirror
    assert np.array_equal(flipped, plain[::-1, :, ::-1])


def coated_flat(theta, thickness, delta, overhang=False):
    """A substrate (layer z < 2) under a coating deposited at polar angle
    theta; with `overhang`, a slab z in [6, 7) over x < 4 that shadows the
    substrate below it.  Unit-ish cells: 8 x 8 x 40 corners, Dz = 0.25."""
    st = capi.Structure(8.0, 8.0, 10.0).add_layer("z", -1.0, 2.0, 1)
    if overhang:
        st.add_block(0.0, 4.0, -1.0, 9.0, 6.0, 7.0, 3)
    return st.add_coating(theta, 0, thickness, delta, 0, 2)


def expected_coating(n_steps, cos_t, delta, overhang=False):
    """Closed form of Add_Coating on coated_flat: air corners whose support
    march (points i delta apart along dir, up to the first one past the
    thickness, i.e. n_steps of them below the start) reaches the substrate or
    the slab top, and whose climb meets no slab."""
    z = np.arange(40) * 0.25
    x = np.arange(8)
    g = np.zeros((40, 8, 8), dtype=np.uint32)
    g[z < 2.0] = 1
    reach = n_steps * delta * cos_t
    coat = (z >= 2.0) & (z - reach < 2.0)
    g[coat] = 2
    if overhang:
        under = x < 4
        slab = (z >= 6.0) & (z < 7.0)
        g[np.ix_(coat, np.ones(8, bool), under)] = 0                  # shadowed by the slab
        g[np.ix_(slab, np.ones(8, bool), under)] = 3
        top = (z >= 7.0) & (z - reach < 7.0)
        g[np.ix_(top, np.ones(8, bool), under)] = 2
    return g


@pytest.mark.parametrize("theta,thickness,delta,n_steps", [
    (0, 1.5, 0.5, 4),       # 1.5 / 0.5 steps within the thickness, then one point past it
    (0, 1.4, 0.5, 3),
    (45, 1.5, 0.5, 4),      # reach 4 x 0.5 x cos 45 = 1.414: corners up to 3.25
])
def test_coating_flat_substrate_closed_form(oracle_be, theta, thickness, delta, n_steps):
    """add_coating (structure_primitives.cpp:45-96): on a flat substrate the
    coating is the air within n_steps x delta x cos(theta) of its top, where
    n_steps counts the support march's points, the first past `thickness`
    included (a march that stopped at the thickness would give 3 steps in the
    first case).  Angles in degrees, dir_z = -cos(theta)."""
    g = grid(oracle_be, coated_flat(theta, thickness, delta), (8, 8, 40), (1.0, 1.0, 0.25))
    want = expected_coating(n_steps, np.cos(np.radians(theta)), delta)
    assert np.array_equal(g, want)


def test_coating_is_shadowed_by_an_overhang(oracle_be):
    """The climb back up to lz: substrate below a slab is not coated, the slab's
    top is."""
    g = grid(oracle_be, coated_flat(0, 1.5, 0.5, overhang=True), (8, 8, 40), (1.0, 1.0, 0.25))
    assert np.array_equal(g, expected_coating(4, 1.0, 0.5, overhang=True))


def test_benchmark_sphere_voxeliser_matches_oracle(oracle_be):
    n, D = 48, 2e-9
    sph = workloads.random_spheres(9, 12, (n * D,) * 3, 3e-9, 9e-9)
    s = capi.Structure(n * D, n * D, n * D)
    for (x, y, z, r, m) in sph:
        s.add_sphere(x, y, z, r, m)
    assert s.cells(D, D, D) == (n, n, n)
    assert np.array_equal(grid(oracle_be, s, (n, n, n), (D, D, D)),
                          workloads.voxelize_spheres((n, n, n), (D, D, D), sph))


def test_compute_discretization_rounding():
    s = capi.Structure(2.5e-9 * 10, 1e-10, 3.5)
    assert s.cells(2.5e-9, 1.0, 1.0) == (10, 1, 4)     # 0.0001 -> max(1, 0); 3.5 rounds away from zero


# ---------------------------------------------------------------- GPU: the engine

def random_structure(seed, L, loop=(0, 0, 0), flip=(0, 0, 0), n_ops=60):
    rng = np.random.default_rng(seed)
    s = capi.Structure(*L).loop(*loop).flip(*flip).default_material(int(rng.integers(0, 3)))
    lo = -0.2 * np.asarray(L)
    hi = 1.2 * np.asarray(L)
    m = min(L)
    for _ in range(n_ops):
        kind = int(rng.integers(0, 6))
        c = rng.uniform(lo, hi)
        mat = int(rng.integers(0, 12))
        if kind == capi.OP_BLOCK:
            d = rng.uniform(lo, hi)
            s.add_block(c[0], d[0], c[1], d[1], c[2], d[2], mat)
        elif kind == capi.OP_SPHERE:
            s.add_sphere(*c, rng.uniform(0.02, 0.3) * m, mat)
        elif kind in (capi.OP_CYLINDER, capi.OP_CONE):
            v = rng.normal(size=3) * 0.3 * np.asarray(L)
            r = rng.uniform(0.02, 0.2) * m
            (s.add_cylinder if kind == capi.OP_CYLINDER else s.add_cone)(*c, *v, r, mat)
        elif kind == capi.OP_ELLIPSOID:
            s.add_ellipsoid(*c, *(rng.uniform(0.03, 0.3, size=3) * np.asarray(L)), mat)
        else:
            ax = int(rng.integers(0, 3))
            z = np.sort(rng.uniform(lo[ax], hi[ax], size=2))
            s.add_layer("xyz"[ax], z[0], z[1], mat)
    return s


@pytest.mark.gpu
@pytest.mark.parametrize("case", [
    dict(seed=1, N=(67, 45, 33), D=(3e-9, 3e-9, 3e-9)),
    dict(seed=2, N=(64, 64, 64), D=(2e-9, 2.5e-9, 1.5e-9), loop=(1, 1, 0)),
    dict(seed=3, N=(40, 50, 30), D=(5e-9, 5e-9, 5e-9), flip=(1, 0, 1)),
    dict(seed=4, N=(33, 1, 70), D=(4e-9, 4e-9, 4e-9), loop=(1, 0, 1), flip=(0, 1, 1)),
    dict(seed=5, N=(13, 13, 13), D=(1.0, 1.0, 1.0)),
])
def test_engine_matches_oracle(engine_be, oracle_be, case):
    N, D = case["N"], case["D"]
    print(f"discretize seed={case['seed']} N={N}", flush=True)
    L = tuple(n * d for n, d in zip(N, D))
    s = random_structure(case["seed"], L, case.get("loop", (0, 0, 0)), case.get("flip", (0, 0, 0)))
    e = grid(engine_be, s, N, D)
    o = grid(oracle_be, s, N, D)
    assert np.array_equal(e, o)
    assert len(np.unique(o)) > 3


@pytest.mark.gpu
def test_engine_lattice_points(engine_be):
    print("discretize lattice points", flush=True)
    for r, n in BALL.items():
        g = grid(engine_be, capi.Structure().add_sphere(6, 6, 6, r, 3), (13, 13, 13))
        assert int((g == 3).sum()) == n


@pytest.mark.gpu
def test_engine_benchmark_packing(engine_be):
    """The 256^3 packing of 40 spheres that the single-particle benchmark uses."""
    print("discretize 256^3 spheres", flush=True)
    n, D = 256, 2e-9
    sph = workloads.random_spheres(2, 40, (n * D,) * 3, 8e-9, 20e-9)
    s = capi.Structure(n * D, n * D, n * D)
    for (x, y, z, r, m) in sph:
        s.add_sphere(x, y, z, r, m)
    assert np.array_equal(grid(engine_be, s, (n, n, n), (D, D, D)),
                          workloads.voxelize_spheres((n, n, n), (D, D, D), sph))


@pytest.mark.gpu
@pytest.mark.parametrize("theta,thickness,delta,n_steps", [(0, 1.5, 0.5, 4), (0, 1.4, 0.5, 3), (45, 1.5, 0.5, 4)])
def test_engine_coating_closed_form(engine_be, theta, thickness, delta, n_steps):
    print(f"discretize coating theta={theta} thickness={thickness}", flush=True)
    g = grid(engine_be, coated_flat(theta, thickness, delta), (8, 8, 40), (1.0, 1.0, 0.25))
    assert np.array_equal(g, expected_coating(n_steps, np.cos(np.radians(theta)), delta))
    g = grid(engine_be, coated_flat(0, 1.5, 0.5, overhang=True), (8, 8, 40), (1.0, 1.0, 0.25))
    assert np.array_equal(g, expected_coating(4, 1.0, 0.5, overhang=True))


@pytest.mark.gpu
@pytest.mark.parametrize("loop,flip", [((0, 0, 0), (0, 0, 0)), ((1, 1, 0), (0, 0, 0)), ((0, 0, 0), (1, 1, 0)),
                                       ((1, 0, 1), (0, 1, 0)), ((0, 0, 0), (0, 0, 1))])
def test_engine_coating_matches_oracle(engine_be, oracle_be, loop, flip):
    """Cylinders standing on a substrate, coated at 60 degrees from the normal
    and 20 degrees of azimuth with 1 nm steps (shadowed sides and tops), on
    plain, looped and flipped axes: the engine equals the oracle bit for bit.
    With z flipped the coating is empty: Add_Coating's own Structure::index
    calls flip the already flipped points back, as in the reference."""
    print(f"discretize coating loop={loop} flip={flip}", flush=True)
    L, N, D = (200e-9, 200e-9, 250e-9), (40, 40, 50), (5e-9, 5e-9, 5e-9)
    s = capi.Structure(*L).loop(*loop).flip(*flip)
    s.add_layer("z", -50e-9, 100e-9, 1)
    s.add_cylinder(60e-9, 60e-9, 100e-9, 0, 0, 60e-9, 30e-9, 2)
    s.add_cylinder(150e-9, 130e-9, 100e-9, 0, 0, 90e-9, 25e-9, 2)
    s.add_coating(60, 20, 20e-9, 1e-9, 0, 3)
    s.add_sphere(100e-9, 170e-9, 190e-9, 15e-9, 4)
    e = grid(engine_be, s, N, D)
    o = grid(oracle_be, s, N, D)
    assert np.array_equal(e, o)
    assert (int((o == 3).sum()) > 100) == (flip[2] == 0) and int((o == 4).sum()) > 0


@pytest.mark.gpu
def test_engine_rejects_malformed(engine_be):
    with pytest.raises(capi.FDTDError):
        grid(engine_be, capi.Structure().add_cylinder(1, 1, 1, 0, 0, 0, 1, 1), (4, 4, 4))
    # a coating deposited from 90 degrees or more never ends its climb; one
    # over another coating is not provided
    with pytest.raises(capi.FDTDError):
        grid(engine_be, capi.Structure(4, 4, 4).add_layer("z", 0, 1, 1).add_coating(90, 0, 1, 0.5, 0, 2), (4, 4, 4))
    with pytest.raises(capi.FDTDError):
        grid(engine_be, capi.Structure(4, 4, 4).add_coating(0, 0, 1, 0.5, 0, 2).add_coating(0, 0, 1, 0.5, 0, 3),
             (4, 4, 4))
    with pytest.raises(capi.FDTDError):
        grid(engine_be, capi.Structure(0, 1, 1).loop(1, 0, 0), (4, 4, 4))
