"""Oracle trajectory of the 2-D single-block parity configuration (128 x 128
cells on 1 m x 1 m, k = 50, rho = 7800, cp = 460, 302 at 500 K / 300 K on
faces 1 / 2, adiabatic 301 on faces 3 / 4, seeded +-1 K noise on 300 K,
RK3 with local time steps, CFL (vnn) 0.8, no IRS, no multigrid) run to the
reference's residual test 1e-10 (Mod_Explicit.f90:65-90).

The run takes minutes on one core, so its outcome is stored as a fixture
(plate2d_128_oracle.npz: stop iteration, final T, the residual every 500
iterations and at the stop) that the GPU test reads.  Everything here is
the oracle (test infrastructure); the case is fuss_amd.cases.plate2d(128).

Usage: python tests/golden/make_plate128.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

from fuss_amd import cases  # noqa: E402
from oracle import pyoracle  # noqa: E402


def main():
    c = cases.plate2d(128)
    o = pyoracle.Oracle(c)
    t0 = time.perf_counter()
    n, hist, _ = o.run()
    dt = time.perf_counter() - t0
    res = hist[:, 2]
    every = np.arange(499, n, 500)
    np.savez_compressed(os.path.join(HERE, "plate2d_128_oracle.npz"), stop_iter=np.int64(n), T=o.get("T"),
                        res_iters=(every + 1).astype(np.int64), res=res[every], res_last=res[-1],
                        oracle_seconds=dt)
    print(f"plate2d_128: stop iteration {n}, final residual {res[-1]:.6e}, {dt:.1f} s")


if __name__ == "__main__":
    main()
