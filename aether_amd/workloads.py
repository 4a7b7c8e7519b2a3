"""Synthetic inputs of the benchmark configurations.

Structures are discretised the way Aether's Structure::discretize does it
(src/structure/structure.cpp:94-113): cell (i,j,k) takes the material of the
last structure operation that contains the corner point (i*Dx, j*Dy, k*Dz),
else the default material.  Spheres follow Add_Sphere::index
(src/structure/structure_primitives.cpp:490-500, closed ball) and layers
Add_Layer::index (:323-338, half-open z1 <= z < z2).

Random numbers come from a 64-bit Mersenne twister (mt19937_64) mapped to
doubles as (x >> 11) * 2**-53, which is portable across languages.
"""
from __future__ import annotations

import math

import numpy as np

from .capi import (EX, EY, EZ, FP32, FP64, HX, HY, HZ, NORMAL_X, NORMAL_Y, NORMAL_Z, Material, ModeSpec, SensorReq,
                   SourceReq)


class MT19937_64:
    """Standard 64-bit Mersenne twister (Matsumoto & Nishimura, 2000).  The
    312-word state is regenerated with numpy in three vector pieces (words
    whose recurrence partners are all old, then those that read words already
    renewed, then the last one), so draws cost little in Python."""
    _N, _M = 312, 156
    _A = np.uint64(0xB5026F5AA96619E9)
    _UM, _LM = np.uint64(0xFFFFFFFF80000000), np.uint64(0x7FFFFFFF)
    _MASK = (1 << 64) - 1

    def __init__(self, seed):
        mt = [0] * self._N
        mt[0] = seed & self._MASK
        for i in range(1, self._N):
            prev = mt[i - 1]
            mt[i] = (6364136223846793005 * (prev ^ (prev >> 62)) + i) & self._MASK
        self.mt = np.array(mt, dtype=np.uint64)
        self.out = []
        self.idx = 0

    def _twist(self):
        N, M = self._N, self._M
        mt = self.mt
        one = np.uint64(1)

        def step(i0, i1, nxt, far):
            x = (mt[i0:i1] & self._UM) | (nxt & self._LM)
            xa = x >> one
            xa = np.where((x & one) == one, xa ^ self._A, xa)
            mt[i0:i1] = far ^ xa

        step(0, N - M, mt[1:N - M + 1], mt[M:N])                   # partners i+1, i+M not yet renewed
        step(N - M, N - 1, mt[N - M + 1:N], mt[0:M - 1].copy())     # partner i+M-N already renewed
        step(N - 1, N, mt[0:1], mt[M - 1:M])                       # wraps onto the renewed word 0
        y = mt.copy()
        y ^= (y >> np.uint64(29)) & np.uint64(0x5555555555555555)
        y ^= (y << np.uint64(17)) & np.uint64(0x71D67FFFEDA60000)
        y ^= (y << np.uint64(37)) & np.uint64(0xFFF7EEE000000000)
        y ^= y >> np.uint64(43)
        self.out = y.tolist()
        self.idx = 0

    def next_u64(self):
        if self.idx >= len(self.out):
            self._twist()
        v = self.out[self.idx]
        self.idx += 1
        return v

    def uniform(self, a=0.0, b=1.0):
        return a + (b - a) * ((self.next_u64() >> 11) * 2.0 ** -53)


# ---------------------------------------------------------------- materials
# Gold-like Drude pole of the 256^3 configuration plus the two critical-point
# poles of mat_lib/Au_400n_1000n_Vial.lua:8-9, eps_inf 1.1.
GOLD = Material(eps_inf=1.1, drude=[(1.37e16, 1.07e14)],
                critpoint=[(0.86822, 4.0812e15, -0.60756, 7.3277e14),
                           (1.3700, 6.4269e15, -0.087341, 6.7371e14)])


def voxelize_spheres(n, D, spheres, background=0):
    """spheres: list of (x0, y0, z0, r, mat) in metres; later ones win."""
    nx, ny, nz = n
    g = np.full((nz, ny, nx), background, dtype=np.uint32)
    x = np.arange(nx) * D[0]
    y = np.arange(ny) * D[1]
    z = np.arange(nz) * D[2]
    for (x0, y0, z0, r, mat) in spheres:
        i0, i1 = max(0, int(math.floor((x0 - r) / D[0])) - 1), min(nx, int(math.ceil((x0 + r) / D[0])) + 2)
        j0, j1 = max(0, int(math.floor((y0 - r) / D[1])) - 1), min(ny, int(math.ceil((y0 + r) / D[1])) + 2)
        k0, k1 = max(0, int(math.floor((z0 - r) / D[2])) - 1), min(nz, int(math.ceil((z0 + r) / D[2])) + 2)
        if i0 >= i1 or j0 >= j1 or k0 >= k1:
            continue
        dx = x[i0:i1] - x0
        dy = y[j0:j1] - y0
        dz = z[k0:k1] - z0
        inside = (dx[None, None, :] * dx[None, None, :] + dy[None, :, None] * dy[None, :, None]
                  + dz[:, None, None] * dz[:, None, None]) <= r * r
        g[k0:k1, j0:j1, i0:i1][inside] = mat
    return g


def random_spheres(seed, count, L, rmin, rmax, mat=1, max_tries=200000):
    """Non-overlapping spheres, radius U[rmin,rmax], centre U[r, L-r]^3, by rejection."""
    rng = MT19937_64(seed)
    out = []
    tries = 0
    while len(out) < count and tries < max_tries:
        tries += 1
        r = rng.uniform(rmin, rmax)
        c = [rng.uniform(r, L[a] - r) for a in range(3)]
        if all((c[0] - s[0]) ** 2 + (c[1] - s[1]) ** 2 + (c[2] - s[2]) ** 2 > (r + s[3]) ** 2 for s in out):
            out.append((c[0], c[1], c[2], r, mat))
    return out


# ---------------------------------------------------------------- configs
def slab_normal_incidence(n=(64, 64, 64), D=5e-9, slab=(22, 42), eps=2.25, Nt=2000, Nl=50,
                          pml_z=8, precision=FP64):
    """Plain 3D normal-incidence run: dielectric slab between z cells slab[0]..slab[1]
    (Add_Layer corner sampling), 8-cell CPML on +-z, periodic x/y, x-polarised
    pulse, Courant factor 0.95 of the 3D limit (time_mod = 0.95/0.99)."""
    nx, ny, nz = n
    g = np.zeros((nz, ny, nx), dtype=np.uint32)
    z = np.arange(nz) * D
    g[(z >= slab[0] * D) & (z < slab[1] * D)] = 1
    return ModeSpec(matsgrid=g, materials=[Material(1.0), Material(eps)], Nt=Nt, Dx=D, Dy=D, Dz=D,
                    time_mod=0.95 / 0.99, pml=(0, 0, 0, 0, pml_z, pml_z), Nl=Nl,
                    lambda_min=400e-9, lambda_max=800e-9, polarization=1, precision=precision)


def spheres_single_particle(n=256, D=2e-9, count=40, rmin=8e-9, rmax=20e-9, seed=2, Nt=10000, Nl=64,
                            pml=12, eps_bg=1.77, precision=FP32, time_mod=1.0):
    """Scatterer run: gold-like Drude + 2 critical-point spheres in a eps=1.77
    background, CPML on all six faces, TFSF plane wave, box spectral-Poynting
    sensor on the structure boundary."""
    L = (n * D,) * 3
    sph = random_spheres(seed, count, L, rmin, rmax)
    g = voxelize_spheres((n, n, n), (D, D, D), sph)
    spec = ModeSpec(matsgrid=g, materials=[Material(eps_bg), GOLD], Nt=Nt, Dx=D, Dy=D, Dz=D,
                    time_mod=time_mod, pml=(pml,) * 6, Nl=Nl, lambda_min=400e-9, lambda_max=800e-9,
                    polarization=1, precision=precision)
    sensor = SensorReq(box=True, x1=0, x2=n, y1=0, y2=n, z1=0, z2=n, Nl=Nl, lambda_min=400e-9, lambda_max=800e-9)
    return spec, [sensor], sph


LORENTZ_LAYER = Material(eps_inf=1.0, lorentz=[(1.5, 3e15, 1e14)])


def multilayer_normal_incidence(n=(512, 512, 384), D=4e-9, seed=3, nlayers=12, margin=16, Nt=8000, Nl=128,
                                pml_z=16, precision=FP32):
    """Precision-form stack: nlayers dielectric layers (thickness U[10,60] cells,
    eps U[1.5,6.0]) plus one Lorentz-pole layer (d_eps=1.5, w0=3e15 rad/s,
    g=1e14 1/s) at a seeded position, thicknesses rescaled to fit the structure
    height minus a vacuum margin on each side; 16-cell CPML on +-z, periodic x/y."""
    rng = MT19937_64(seed)
    nx, ny, nz = n
    thick = [rng.uniform(10, 60) for _ in range(nlayers)]
    eps = [rng.uniform(1.5, 6.0) for _ in range(nlayers)]
    lor_pos = int(rng.uniform(0, nlayers + 1))
    lor_thick = rng.uniform(10, 60)
    thick.insert(lor_pos, lor_thick)
    avail = nz - 2 * margin
    scale = avail / sum(thick)
    edges = [margin]
    acc = 0.0
    for t in thick:
        acc += t * scale
        edges.append(margin + int(round(acc)))
    mats = [Material(1.0)]
    ids = []
    li = 0
    for q in range(len(thick)):
        if q == lor_pos:
            mats.append(LORENTZ_LAYER)
        else:
            mats.append(Material(eps[li]))
            li += 1
        ids.append(len(mats) - 1)
    g = np.zeros((nz, ny, nx), dtype=np.uint32)
    for q in range(len(thick)):
        g[edges[q]:edges[q + 1]] = ids[q]
    spec = ModeSpec(matsgrid=g, materials=mats, Nt=Nt, Dx=D, Dy=D, Dz=D, pml=(0, 0, 0, 0, pml_z, pml_z), Nl=Nl,
                    lambda_min=400e-9, lambda_max=800e-9, polarization=1, precision=precision)
    return spec, {"thickness_scale": scale, "edges": edges, "lorentz_layer": lor_pos}


# ---------------------------------------------------------------- 2D sweeps
def random_cylinders(rng, n, fill, rmin=3.0, rmax=15.0, max_tries_factor=15):
    """Rejection-packed disks (cylinders along z) on an n x n cell grid: radius
    U[rmin, rmax] cells, centre U[r, n-r]^2, accepted when its centre lies
    more than r + r' + 1 cells from every earlier disk (one clear cell between
    disks); stops at the fill fraction or after max_tries_factor times the
    expected disk count of draws.  Returns the (n, n) boolean mask (corner
    sampling: cell (i,j) is inside when (i-x0)^2+(j-y0)^2 <= r^2) and the
    disk list."""
    occ = np.zeros((n, n), dtype=bool)
    target = fill * n * n
    mean_area = math.pi * (rmin * rmin + rmin * rmax + rmax * rmax) / 3.0
    max_tries = int(max_tries_factor * target / mean_area) + 100
    cell = 2.0 * rmax + 1.0                  # bucket side: neighbours lie in the 3x3 buckets around
    nb = int(n // cell) + 1
    buckets = {}
    area = 0.0
    disks = []
    for _ in range(max_tries):
        if area >= target:
            break
        r = rng.uniform(rmin, rmax)
        x0 = rng.uniform(r, n - r)
        y0 = rng.uniform(r, n - r)
        bx, by = int(x0 // cell), int(y0 // cell)
        ok = True
        for qx in (bx - 1, bx, bx + 1):
            for qy in (by - 1, by, by + 1):
                for (xa, ya, ra) in buckets.get(qx * nb + qy, ()):
                    d = r + ra + 1.0
                    if (x0 - xa) ** 2 + (y0 - ya) ** 2 <= d * d:
                        ok = False
                        break
                if not ok:
                    break
            if not ok:
                break
        if not ok:
            continue
        buckets.setdefault(bx * nb + by, []).append((x0, y0, r))
        disks.append((x0, y0, r))
        area += math.pi * r * r
    for (x0, y0, r) in disks:
        i0, i1 = max(0, int(math.floor(x0 - r))), min(n, int(math.ceil(x0 + r)) + 1)
        j0, j1 = max(0, int(math.floor(y0 - r))), min(n, int(math.ceil(y0 + r)) + 1)
        dx = np.arange(i0, i1) - x0
        dy = np.arange(j0, j1) - y0
        occ[j0:j1, i0:i1] |= (dy[:, None] ** 2 + dx[None, :] ** 2) <= r * r
    return occ, disks


def sweep_steps(side, lo=256, hi=2048):
    """Time steps of a sweep run: 4000 at side 256 rising linearly to 12000 at 2048."""
    return int(round(4000 + 8000 * (side - lo) / (hi - lo)))


def sweep_run(rng, index, D=10e-9, pml=16, Nl=64, lambda_min=500e-9, lambda_max=1000e-9, precision=FP32,
              side=None, steps=None):
    """One run of the 2D sweep: an n x n x 1 structure (n log-uniform in
    [256, 2048]), dielectric cylinders (eps U[2,12] per run, fill U[20%,50%]),
    TE (Ex,Ey,Hz) or TM (Ez,Hx,Hy) chosen at random, 16-cell CPML on the four
    in-plane faces, an oscillator line source along x at the bottom edge of
    the structure and a planar spectral-Poynting line sensor along x at its
    top edge.  The spec is not yet finalized (pads come from mode_finalize)."""
    n = side if side is not None else int(round(math.exp(rng.uniform(math.log(256), math.log(2048)))))
    eps = rng.uniform(2.0, 12.0)
    fill = rng.uniform(0.2, 0.5)
    tm = rng.uniform() < 0.5
    occ, disks = random_cylinders(rng, n, fill)
    g = occ.astype(np.uint32).reshape(1, n, n)
    nt = steps if steps is not None else sweep_steps(n)
    if tm:
        disable, orient = (EX, EY, HZ), NORMAL_Z
    else:
        disable, orient = (EZ, HX, HY), NORMAL_X
    src = SourceReq(orientation=orient, x1=0, x2=n, y1=0, y2=1, z1=0, z2=1,
                    lambda_min=lambda_min, lambda_max=lambda_max)
    sens = SensorReq(box=False, orientation=NORMAL_Y, x1=0, x2=n, y1=n - 1, y2=n, z1=0, z2=1, Nl=Nl,
                     lambda_min=lambda_min, lambda_max=lambda_max)
    spec = ModeSpec(matsgrid=g, materials=[Material(1.0), Material(eps)], Nt=nt, Dx=D, Dy=D, Dz=D,
                    pml=(pml, pml, pml, pml, 0, 0), pad=(0,) * 6, Nl=Nl, lambda_min=lambda_min,
                    lambda_max=lambda_max, disable_fields=disable, precision=precision, sources=[src])
    info = {"index": index, "side": n, "eps": eps, "fill_target": fill, "fill": float(occ.mean()),
            "cylinders": len(disks), "pol": "TM" if tm else "TE", "steps": nt}
    return spec, [sens], info


def sweep_runs(count=64, seed=4, precision=FP32, **kw):
    """The batching workload: `count` independent 2D runs drawn from one
    mt19937_64 stream (seed 4)."""
    rng = MT19937_64(seed)
    return [sweep_run(rng, q, precision=precision, **kw) for q in range(count)]


# ---------------------------------------------------------------- serving
LORENTZ_SCATTERER = Material(eps_inf=2.0, lorentz=[(1.5, 3e15, 1e14)])


def serving_requests(count=24, seed=5, rate=4.0, precision=FP32):
    """The serving stream: Poisson arrivals at `rate` per second (seed 5) of
    mixed requests, each (arrival_s, kind, spec, sensors, info):
      * 3D single-particle runs, structure side U[96,384] cells of 2 nm, as the
        256^3 scatterer run (12-cell CPML on all faces, eps 1.77 background,
        box sensor at 64 wavelengths) with round(40 (side/256)^3) spheres of
        radius 8-20 nm, either gold-like (Drude + 2 critical points) or a
        Lorentz scatterer;
      * 2D custom runs as the sweep (cylinders, TE or TM, oscillator line
        source, line sensor) with side U[512,4096];
    step counts U[2000,10000]; FP32 fields, FP64 sensor accumulators."""
    from .capi import KIND_CUSTOM, KIND_SINGLE
    rng = MT19937_64(seed)
    out = []
    t = 0.0
    for q in range(count):
        t += -math.log(1.0 - rng.uniform()) / rate
        steps = int(round(rng.uniform(2000, 10000)))
        if rng.uniform() < 0.5:
            side = int(round(rng.uniform(96, 384)))
            nsph = max(1, int(round(40 * (side / 256.0) ** 3)))
            gold = rng.uniform() < 0.5
            D = 2e-9
            L = (side * D,) * 3
            sph = random_spheres(int(rng.next_u64() & 0x7FFFFFFF), nsph, L, 8e-9, 20e-9)
            g = voxelize_spheres((side,) * 3, (D,) * 3, sph)
            spec = ModeSpec(matsgrid=g, materials=[Material(1.77), GOLD if gold else LORENTZ_SCATTERER], Nt=steps,
                            Dx=D, Dy=D, Dz=D, pml=(12,) * 6, Nl=64, lambda_min=400e-9, lambda_max=800e-9,
                            polarization=1, precision=precision)
            sens = [SensorReq(box=True, x1=0, x2=side, y1=0, y2=side, z1=0, z2=side, Nl=64,
                              lambda_min=400e-9, lambda_max=800e-9)]
            info = {"index": q, "kind": "3D", "side": side, "spheres": len(sph), "steps": steps,
                    "material": "gold" if gold else "lorentz"}
            out.append((t, KIND_SINGLE, spec, sens, info))
        else:
            side = int(round(rng.uniform(512, 4096)))
            spec, sens, info = sweep_run(rng, q, side=side, steps=steps, precision=precision)
            info["kind"] = "2D"
            out.append((t, KIND_CUSTOM, spec, sens, info))
    return out
