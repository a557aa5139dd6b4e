"""Mutation test of the oracle's pinning: each planted mistake of
oracle/oracle.cpp (ORACLE_PLANT = k: a swapped z parity of the 3D prediction,
detail and on-level prediction coefficients, the 3D upwind sum without
+(-F_F), a corner ghost copied from the wrong cell, the Dirichlet<1> ghost with
the wrong sign, the wrong child in the 3D edge/vertex cascade, a swapped x
parity of the 2D prediction, a refinement of several adapted fields that
needs every field, the polynomial extrapolation of the far ghost layers,
the WENO5 weights and boundary stencil) is built into a scratch library, and the
known-answer checks (golden fixtures, pinning and invariant tests) run on it
must fail.  A check that passes on a mutant pins nothing about that part."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "oracle.cpp")
CHECKS = ["tests/test_oracle_pinning.py", "tests/test_oracle_invariants.py", "tests/test_oracle_golden.py", "tests/test_oracle_weno5.py"]
PLANTS = {1: "3D prediction z parity", 2: "3D detail z parity", 3: "3D upwind without +(-F_F)", 4: "corner ghost source",
          5: "Dirichlet sign", 6: "3D cascade child offset", 7: "3D on-level prediction z parity", 8: "2D prediction x parity",
          9: "several fields: refinement only when every field asks", 10: "stencil-6 extrapolation coefficient",
          11: "stencil-4 extrapolation sources swapped", 12: "WENO5 linear weights swapped", 13: "corner block copied from the layer below",
          14: "WENO5 boundary stencil in the opposite direction", 15: "relative detail multiplied by the max",
          16: "refine-boundary without the low faces"}


@pytest.fixture(scope="module")
def mutants(tmp_path_factory):
    """Build every mutant and run the checks on each, the mutants side by side
    (one process each); returns {plant: completed process}."""
    d = tmp_path_factory.mktemp("mutants")

    def run(k):
        out = str(d / f"liboracle_plant{k}.so")
        subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", f"-DORACLE_PLANT={k}", "-o", out, SRC],
                       check=True)
        env = dict(os.environ, SMR_ORACLE_LIB=out, OMP_NUM_THREADS="1")
        r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu", *CHECKS],
                           cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
        return k, r

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        return dict(ex.map(run, PLANTS))


@pytest.mark.parametrize("k", sorted(PLANTS))
def test_planted_mistake_is_caught(mutants, k):
    r = mutants[k]
    # exit code 1: tests ran and at least one failed (2+ would be a usage or collection error)
    assert r.returncode == 1, f"plant {k} ({PLANTS[k]}) not caught:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}"
