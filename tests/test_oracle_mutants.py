"""Mutation test of the oracle's pinning: each planted mistake of
oracle/oracle.cpp (ORACLE_PLANT = k: a swapped z parity of the 3D prediction,
detail and on-level prediction coefficients, the 3D upwind sum without
+(-F_F), a corner ghost copied from the wrong cell, the Dirichlet<1> ghost with
the wrong sign, the wrong child in the 3D edge/vertex cascade, a swapped x
parity of the 2D prediction, a refinement of several adapted fields that
needs every field) is built into a scratch library, and the
known-answer checks (golden fixtures, pinning and invariant tests) run on it
must fail.  A check that passes on a mutant pins nothing about that part."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "oracle.cpp")
CHECKS = ["tests/test_oracle_pinning.py", "tests/test_oracle_invariants.py", "tests/test_oracle_golden.py"]
PLANTS = {1: "3D prediction z parity", 2: "3D detail z parity", 3: "3D upwind without +(-F_F)", 4: "corner ghost source",
          5: "Dirichlet sign", 6: "3D cascade child offset", 7: "3D on-level prediction z parity", 8: "2D prediction x parity",
          9: "several fields: refinement only when every field asks"}


@pytest.fixture(scope="module")
def mutants(tmp_path_factory):
    d = tmp_path_factory.mktemp("mutants")

    def build(k):
        out = str(d / f"liboracle_plant{k}.so")
        subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", f"-DORACLE_PLANT={k}", "-o", out, SRC],
                       check=True)
        return k, out

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        return dict(ex.map(build, PLANTS))


@pytest.mark.parametrize("k", sorted(PLANTS))
def test_planted_mistake_is_caught(mutants, k):
    env = dict(os.environ, SMR_ORACLE_LIB=mutants[k])
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu", *CHECKS],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    # exit code 1: tests ran and at least one failed (2+ would be a usage or collection error)
    assert r.returncode == 1, f"plant {k} ({PLANTS[k]}) not caught:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}"
