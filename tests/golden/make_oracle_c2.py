"""Summary of the CPU oracle's FP64 BiCG solve of the 128x128x64 workload
(steady DR advection-diffusion, lambda = 1e-3, velocity seed 0, Gaussian
source, inflow Dirichlet), written to oracle_c2.json.  Takes about a minute:
    python tests/golden/make_oracle_c2.py
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import __graft_entry__ as ge  # noqa: E402

pkg = ge.load_package()
po = ge.load_oracle()
p = pkg.problem.config_problem("c2")
bb, b, d = po.assemble(p)
b, d, x0 = po.impose_mask(p.mask, b, d, np.zeros(p.N))
t = time.time()
x, it, err, rc = po.dbcg_simple(b, x0, 1e-8, 20000, bb, p.iss, d, p.nax, p.nay, p.naz, p.mask)
sec = time.time() - t
r = b - po.matvec_0(x, p.iss, bb, p.nax, p.nay, p.naz, p.mask)
out = {"workload": "unit box 128x128x64, config_problem('c2')", "rtol": 1e-8,
       "iterations": it, "recursive_err": err,
       "true_rel_residual": float(np.linalg.norm(r) / np.linalg.norm(b)),
       "x_norm2": float(np.linalg.norm(x)), "x_sum": float(x.sum()), "x_max": float(x.max()),
       "mass": po.mass(p.iss, p.xx, p.yy, p.zz, x, p.E), "b_norm2": float(np.linalg.norm(b)),
       "oracle_seconds": sec}
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_c2.json"), "w") as f:
    json.dump(out, f, indent=1)
print(out)
