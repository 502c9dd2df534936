"""Write tests/golden/dbyd2_ref.npz: outputs of the reference DBYD2
(dbyd2.f + mtridag.f compiled from the reference sources by oracle/Makefile
into oracle/_ref) on seeded inputs, for every boundary mode used by the 2D
path and a few column lengths.  Run in the build container:

    make -C oracle ref && python tests/golden/make_dbyd2_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref not built")
    rng = np.random.default_rng(20261202)
    out = {}
    for n2 in (5, 7, 20, 150, 300):
        for ibc in (0, 1, 2, 3):
            B = rng.standard_normal((n2, 3))
            A = rng.standard_normal((n2, 3))
            out[f"B_{n2}_{ibc}"] = B
            out[f"A_{n2}_{ibc}"] = A
            out[f"R_{n2}_{ibc}"] = oracle.ref_dbyd2(B, ibc, A)
    # unstretch of the 300-point solar_deep-like z column and a polytrope column
    from sparc_flows_amd import background as bg
    z = bg.polytrope(300)[:, 0].copy()
    out["z_poly300"] = z
    out["dz_poly300"] = oracle.ref_dbyd2(z, 1)[:, 0]
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "dbyd2_ref.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
