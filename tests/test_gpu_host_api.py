"""The C++11 host API end to end on the GPU: problem classes written in the
reference's style (CRTP equation -> boundary -> problem layering) run
through Problem::iterateOverTime with the update phase on the device.

The programs print the reference's "Local Linf" lines; for the
TimeLinSpaceLin problems (exact discrete solution) they must be round-off,
and the final field must equal the oracle's bit for bit.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
BIN = os.path.join(O.ROOT, "masterthesis_amd", "bin")


def run(prog, args, dump):
    env = dict(os.environ, SCAFES_DUMP=dump)
    p = subprocess.run([os.path.join(BIN, prog)] + args, capture_output=True, text=True, env=env, timeout=120)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    return p.stdout


@pytest.mark.parametrize("prog,n,pennes", [("heat_timelin_1d", (17,), False), ("heat_timelin_2d", (12, 9), False),
                                           ("heat_timelin_3d", (10, 9, 8), False),
                                           ("pennes_timelin_3d", (9, 10, 11), True)])
def test_timelin_through_problem(tmp_path, hip, prog, n, pennes):
    nsteps = 40
    g = O.make_grid(n)
    h2 = sum(1.0 / O.lib().oracle_gridsize(g, p) ** 2 for p in range(len(n)))
    tau = 0.8 / (2.0 * h2 + (1.0 if pennes else 0.0))
    dump = str(tmp_path / "u.bin")
    args = [f"--dim={len(n)}", "--nNodes=" + "x".join(map(str, n)), f"--nTimesteps={nsteps}", f"--tau={tau!r}"]
    out = run(prog, args, dump)
    linf = [float(x) for x in re.findall(r"Local Linf error\(U, USolDom\): (\S+)", out)]
    assert len(linf) == nsteps + 1
    assert max(linf) < 1e-12
    coef = O.heat7_coef(tau, perf_c=1.0, perf_w=1.0, rhob_cb_w=1.0, has_perf=pennes, has_source=True)
    ref, rlinf, _ = O.heat7_run(g, np.zeros(O.memall_shape(g)), nsteps, 0.0, coef,
                                O.KNOWN_PENNES_TIMELIN if pennes else O.KNOWN_HEAT_TIMELIN)
    got = np.fromfile(dump, dtype=np.float64)
    np.testing.assert_array_equal(got, O.nodes_view(g, ref).ravel())
    # the "Local Linf" values are those of the CPU path
    np.testing.assert_allclose(linf, rlinf, rtol=0, atol=1e-15)
