"""One small occluded solve on cuda:0 through the C ABI, checked against the
oracle (test infrastructure, imported only here, in tests and in bench.py's
CPU baseline)."""
import os
import sys

import numpy as np


def run():
    from . import api
    from . import mesh as M
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, os.path.join(root, "oracle"))
    import oracle as orc

    m = M.occluded_plates(n=6, subdiv=2)
    for prec in ("fp64", "mixed"):
        res = api.compute_view_factors(m, nquad=3, obstruction_groups=[3],
                                       backend=api.MI355XBackend(precision=prec), verbose=False)
        r = orc.compute(m, nquad=3, obstruction_groups=[3])
        F = r["raw"] / r["A"][:, None]
        d = float(np.abs(res.F_elem - F).max())
        tol = 0.0 if prec == "fp64" else 1e-6
        if d > tol:
            raise AssertionError(f"smoke: {prec} max |dF| = {d} > {tol}")
        print(f"smoke {prec}: N={m.n_elems} max|dF|={d:.3e} F_group[1,2]={res.F_group[0, 1]:.6f}")
