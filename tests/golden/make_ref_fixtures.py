"""Writes tests/golden/ref_materials.json from Aether's own compiled
dispersive-model and math code (oracle/_ref/libaeref.so, built by
oracle/Makefile from src/materials/dielec_models.cpp and src/common/mathUT.cpp
where they lie in the reference tree).

The fixture holds inputs and the reference's outputs only:
  - Dielec_model pole time exponentials (a, b) and Drude sigma
    (dielec_models.cpp:222-226,281-290,350-358,419-423);
  - Dielec_model::eval(w) (dielec_models.cpp:52-64);
  - s_curve and nearest_integer (mathUT.cpp:1169-1193).
Run:  python tests/golden/make_ref_fixtures.py
"""
import ctypes as C
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

# (kind, params): 0 Debye(ds,t0) 1 Drude(wd,g) 2 Lorentz(A,O,G) 3 CritPoint(A,O,P,G)
MODELS = [
    (0, [1.2, 3e-15]),
    (0, [5.0, 1e-13]),
    (1, [1.37e16, 1.07e14]),
    (1, [1.3e16, 5e13]),
    (2, [1.5, 3e15, 1e14]),
    (2, [0.7, 5e15, 2e15]),
    (3, [0.86822, 4.0812e15, -0.60756, 7.3277e14]),
    (3, [1.3700, 6.4269e15, -0.087341, 6.7371e14]),
]
OMEGAS = [1e14, 2.35e15, 3.1e15, 4.7e15, 1e16]
SCURVE = [(x, 9.0, 3.0) for x in (0.0, 2.9, 3.0, 3.7, 5.5, 8.999, 9.0, 12.0)] + [(t, 1999.0, 1900.0) for t in (1899, 1900, 1950, 1999, 2000)]
NEAREST = [0.0, 0.49999, 0.5, 1.5, 2.5, -0.5, -1.5, -2.4999, 12.5, 99.49, -7.5]


def main():
    ref = oracle.load_ref()
    if ref is None:
        raise SystemExit("oracle/_ref/libaeref.so not built (needs the reference tree)")
    dp = C.POINTER(C.c_double)
    out = {"models": [], "s_curve": [], "nearest_integer": []}
    for kind, p in MODELS:
        pa = (C.c_double * 4)(*p)
        te = (C.c_double * 5)()
        assert ref.aeref_time_exp(kind, pa, te) == 0
        ev = []
        for w in OMEGAS:
            r = (C.c_double * 2)()
            assert ref.aeref_eval(kind, pa, 1.3, w, r) == 0
            ev.append([w, r[0], r[1]])
        out["models"].append({"kind": kind, "params": p, "a": [te[0], te[1]], "b": [te[2], te[3]],
                              "sigma": te[4], "eps_inf": 1.3, "eval": ev})
    for (x, a, b) in SCURVE:
        out["s_curve"].append([x, a, b, ref.aeref_s_curve(x, a, b)])
    for x in NEAREST:
        out["nearest_integer"].append([x, ref.aeref_nearest_integer(x)])
    with open(os.path.join(HERE, "ref_materials.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "ref_materials.json"))


if __name__ == "__main__":
    main()
