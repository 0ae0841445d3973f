"""Write golden input/output vectors of the reference's closed-form cooling
module (Particles_in_BH_Tree_III/hcool_with_unit_tests_Here/photolibs3.py).

Run once in a container that has the reference checked out at
/root/reference; the output JSON is committed and the tests read only it.
photolibs3 needs numpy only.  The functions it covers (ionisation
equilibrium by damped fixed-point iteration, photo-rate integrals, the
composite cooling/heating rates, the implicit u update) are not pinned by a
closed form, so their values are taken from the reference itself.
"""
import json
import os
import sys

import numpy as np

REF = "/root/reference/Particles_in_BH_Tree_III/hcool_with_unit_tests_Here"


def main(out=os.path.join(os.path.dirname(os.path.abspath(__file__)), "cooling_photolibs3.json")):
    sys.path.insert(0, REF)
    import photolibs3 as P  # noqa: E402

    res = {"source": "Particles_in_BH_Tree_III/hcool_with_unit_tests_Here/photolibs3.py", "XH": 0.76}
    Ts = [1.0e3, 5.0e3, 1.0e4, 2.0e4, 5.0e4, 1.0e5, 3.0e5, 1.0e6, 1.0e7, 1.0e8]
    res["RandCIRates"] = [{"T": T, "out": [float(v) for v in P.RandCIRates(T)]} for T in Ts]
    rf = P.RadiationField()
    res["RadiationField"] = [float(v) for v in rf]
    gJH0, gJHe0, gJHep = rf[0], rf[1], rf[2]
    ab = []
    for T in [1.0e4, 3.0e4, 1.0e5, 1.0e6]:
        for nH in [1e-4, 1e-2, 1.0, 1e2]:
            ab.append({"T": T, "nH": nH, "out": [float(v) for v in P.Abundance_hX(T, nH, gJH0, gJHe0, gJHep)]})
    res["Abundance_hX"] = ab
    ch = []
    for T in [1.0e4, 1.5e4, 2.0e4, 5.0e4, 1.0e5, 2.0e5, 5.0e5, 1.0e6, 1.0e7, 1.0e8]:
        for nH in [1e-4, 1e-2, 1.0, 1e2]:
            h, l = P.coolingHeatingRates(T, nH)
            ch.append({"T": T, "nH": nH, "heat": float(h), "lambda": float(l)})
    res["coolingHeatingRates"] = ch
    ut = []
    for u in [1e11, 1e12, 1e13, 1e14, 1e15]:
        for nH in [1e-2, 1.0]:
            ut.append({"u": u, "nH": nH, "T": float(P.convert_u_to_temp_h(u, nH, 0.76)),
                       "rate": float(P.coolingRateFromU_h(u, nH, 0.76))})
    res["u_to_T_and_rate"] = ut
    dc = []
    for rho in [1e-26, 1e-24, 1e-22]:
        for u in [5e12, 2e13, 1e14]:
            for dt in [3.15e11, 3.15e13]:
                try:
                    dc.append({"rho": rho, "u": u, "dt": dt, "u_new": float(P.DoCooling_h(rho, u, dt, 0.76))})
                except (OverflowError, ZeroDivisionError) as e:
                    # the reference's bracket search divides u by 1.1 without a floor; strong cooling
                    # drives T to 0 and T**-1.5 overflows.  Recorded as such.
                    dc.append({"rho": rho, "u": u, "dt": dt, "error": type(e).__name__})
    res["DoCooling_h"] = dc
    with open(out, "w") as f:
        json.dump(res, f, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main()
