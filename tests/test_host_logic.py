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
    engine_only = {"aefdtd_profile_enable", "aefdtd_profile_read", "aefdtd_device_bytes", "aefdtd_stats",
                   "aefdtd_cancel", "aefdtd_autotune", "aefdtd_set_tiles", "aefdtd_synchronize",
                   "aefdtd_set_storage_scale"}
    missing = [s for s in declared_symbols()
               if s not in engine_only and not s.startswith("aefdtd_run_") or s in
               ("aefdtd_run_normal_incidence", "aefdtd_run_single_particle")]
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
