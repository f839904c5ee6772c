"""Write the golden vectors under tests/golden/ from the reference's own host
code, compiled from /root/reference by oracle/Makefile into
oracle/_ref/libfcref.so (units.h + airfoil.cpp). Run in the build container:

    make -C oracle ref && python tests/golden/make_golden.py

Outputs (data only): units_ref.json, naca_ref.npz.
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as orc  # noqa: E402

FIELDS = ["chord_m", "airspeed_ms", "nu_m2s", "chord_cells", "u_lat", "dx", "dt",
          "nu_lat_target", "nu_lat", "tau", "tau_clamped", "re_target", "re_effective"]

# (chord_m, airspeed, nu, Nc, u_lat) cases: the reference's own test_units inputs
# plus every workload's scaling.
SCALING_CASES = [
    (1.2, 30.0, 1.48e-5, 256, 0.08),
    (1.0, 1000.0 * 1.48e-5, 1.48e-5, 128, 0.08),
    (1.0, 500.0 * 1.48e-5, 1.48e-5, 128, 0.08),
    (1.2, 30.0, 1.48e-5, 256, 0.5),
    (1.2, 30.0, 1.48e-5, 256, 0.05),
    (1.2, 30.0, 1.48e-5, 128, 0.08),
    (1.2, 30.0, 1.48e-5, 96, 0.08),
    (1.2, 30.0, 1.48e-5, 192, 0.08),
    (1.0, 1.48e-5 * 1e-3 * 64 / 0.01, 1.48e-5, 64, 1e-3),
]


def bits(x):
    return int(np.array([x], dtype=np.float32).view(np.uint32)[0])


def main():
    R = orc.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libfcref.so not built (needs /root/reference)")
    out = {"scaling": [], "refined": [], "ramp": [], "max_re": [], "coef": []}
    buf = (C.c_float * 13)()
    for case in SCALING_CASES:
        R.ref_compute_scaling(*case, buf)
        out["scaling"].append({"in": list(case), "bits": [bits(v) for v in buf[:]]})
        for m in (2, 4):
            R.ref_refined_scaling(*case, m, buf)
            out["refined"].append({"in": list(case), "factor": m,
                                   "bits": [bits(v) for v in buf[:]]})
        for nx in (384, 512, 768):
            for step in (0, 1, 17, 100, nx // 2, nx - 1, nx, nx + 3, 2 * nx - 1, 2 * nx, 5 * nx):
                t = R.ref_ramped_tau(*case, step, nx)
                u = R.ref_ramped_u(*case, step, nx)
                out["ramp"].append({"in": list(case), "nx": nx, "step": step,
                                    "tau_bits": bits(t), "u_bits": bits(u)})
        for force, span in ((1.0, 96), (123.25, 240), (-7.5, 64), (0.0, 96), (1.0, 0)):
            c = R.ref_force_to_coefficient(*case, force, span)
            out["coef"].append({"in": list(case), "force": force, "span": span,
                                "bits": bits(c)})
    for nc in (64, 96, 128, 192, 256, 320, 384):
        for ul in (0.05, 0.08, 0.2):
            out["max_re"].append({"nc": nc, "u_lat": ul, "bits": bits(R.ref_max_effective_re(nc, ul))})
    out["fields"] = FIELDS
    with open(os.path.join(HERE, "units_ref.json"), "w") as fh:
        json.dump(out, fh, indent=0)

    arrays = {}
    for digits in ("2412", "0012", "4412", "2415"):
        for pps in (200, 64, 20):
            cap = 2 * max(pps, 20)
            xy = np.zeros((cap, 2), dtype=np.float32)
            le = C.c_int(0)
            n = R.ref_generate_naca4(digits.encode(), pps, xy.ctypes.data_as(C.POINTER(C.c_float)),
                                     cap, C.byref(le))
            arrays[f"naca{digits}_{pps}"] = xy[:n].copy()
            arrays[f"naca{digits}_{pps}_le"] = np.array([le.value], dtype=np.int32)
    np.savez_compressed(os.path.join(HERE, "naca_ref.npz"), **arrays)
    print("wrote", os.path.join(HERE, "units_ref.json"), os.path.join(HERE, "naca_ref.npz"))


if __name__ == "__main__":
    main()
