"""Every planted mistake in the oracle fails some pin.

The oracle is built once per mistake with -DKP_PLANT=k (oracle/kp_oracle.c),
into a temporary library, and the pin files are run against it in a child
pytest.  A pin set that no plausible slip in the oracle's formulas can pass
is what makes the oracle a checker: the GPU kernels are compared with it bit
for bit, so a mistake the pins let through would pass on both sides.

The mistakes are the ones a restatement of the phase matrices, the Stokes
algebra, the march and the tallies could plausibly make: a dropped factor, a
wrong sign or exponent, a transposed image, a changed tie rule.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PLANTS = {
    1: "HG-Rayleigh degree p_max(1-mu^2) without the 1/(1+mu^2)",
    2: "HG-Rayleigh z33 = Z11 2mu, without the 1/(1+mu^2)",
    3: "HG-Rayleigh z33 = Z11 mu",
    4: "scatter a33 scaled by p_max",
    5: "scatter a12 with the wrong sign",
    6: "HG phase function with exponent 1 instead of 3/2",
    7: "peel-off U with the wrong sign",
    8: "image axes swapped",
    9: "Stokes rotation by -2 phi in the scatter",
    10: "get_dl ties to the lowest axis",
    11: "J without the cell volume",
    12: "albedo weighting dropped",
    13: "Thomson z12 with the wrong sign",
    14: "peel-off without the attenuation e^-tau",
}

PIN_FILES = ["tests/test_oracle_pins.py", "tests/test_oracle_hg_pins.py"]


@pytest.mark.parametrize("k", sorted(PLANTS), ids=lambda k: f"plant{k}")
def test_planted_mistake_fails_a_pin(k, tmp_path):
    env = dict(os.environ, KP_ORACLE_PLANT=str(k),
               KP_ORACLE_OUT=str(tmp_path / f"liboracle_plant{k}.so"))
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-m", "not gpu",
                        "-p", "no:cacheprovider", "-p", "no:randomly", *PIN_FILES],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = r.stdout[-3000:] + r.stderr[-2000:]
    assert os.path.exists(env["KP_ORACLE_OUT"]), "planted oracle was not built:\n" + out
    # 1: tests ran and some failed; 0 would mean every pin passed the mistake
    assert r.returncode == 1, f"plant {k} ({PLANTS[k]}) passed every pin:\n" + out
    assert " failed" in r.stdout, out
