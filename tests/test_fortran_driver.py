"""The Fortran iso_c_binding module drives the library like the reference's
driver would: the demo program (fuss_amd/fortran/fuss_gpu_demo.f90) builds a
plate in Fortran arrays and steps it through fuss_step / fuss_run, then
solves the steady state with fuss_refine; the Python binding on the same case
must print the same residuals, refine outcome and state sum."""
import os
import subprocess

import numpy as np
import pytest

from fuss_amd import build, cases
from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu


def test_fortran_driver_matches_python():
    if not gpu_available():
        pytest.skip("no GPU")
    if not os.path.isfile(build.DEMO):
        pytest.skip("Fortran driver not built (no Fortran compiler)")
    out = subprocess.run([build.DEMO], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = [ln.split() for ln in out.stdout.splitlines() if ln.strip()]
    fres = [float(x[2]) for x in lines if x[0] == "step"]
    frun = [x for x in lines if x[0] == "run"][0]
    fref = [x for x in lines if x[0] == "refine"][0]
    fsum = float([x for x in lines if x[0] == "sumT"][0][1])

    from fuss_amd import gpu

    c = cases.plate2d(n=32, noise=0.0)
    s = gpu.Solver(c)
    pres = [np.sqrt(s.explicit_step(1).res_sq) for _ in range(10)]
    k, hr, _, r = s.run_level(1, 200, 200, 1e-10, 1e30)
    np.testing.assert_allclose(fres, pres, rtol=1e-15)
    assert int(frun[1]) == r.iter
    np.testing.assert_allclose(float(frun[2]), np.sqrt(r.res_sq), rtol=1e-15)
    rr, _ = s.refine(gpu.refine_params(res_threshold=1e-10))
    assert int(fref[1]) == rr.outer and int(fref[2]) == 1 == rr.converged
    np.testing.assert_allclose(float(fref[3]), rr.res, rtol=1e-15)
    # the checksum is summed in a different order by Fortran's sum and numpy (~1e-13)
    np.testing.assert_allclose(fsum, float(np.sum(s.get("T"))), rtol=1e-12)
    s.close()
