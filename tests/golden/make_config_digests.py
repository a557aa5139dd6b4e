"""Writes tests/golden/config_digests.json: SHA-256 digests of the oracle's
leaves and cell values along the large 2D workload (L4..12, seeded 3 boxes +
2 Gaussians) and the 3D sphere workload (L2..8), at their full sizes, and
after the 100 frozen-mesh steps of each of the 32 adjoint instances (L2..10).

The oracle needs minutes for these (the first adaptation starts from every
cell at the finest level), too long for a GPU test, so it runs here once and
the GPU tests compare digests of the product's output with these.  Inputs are
bit-identical on every host: the fields use only IEEE basic operations
(initial_conditions.det_exp for the Gaussians).

Run from the repository root:  python tests/golden/make_config_digests.py [name ...]
"""
import hashlib
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from initial_conditions import seeded_2d, sphere_3d  # noqa: E402
from oracle_lib import OracleSim  # noqa: E402

OUT = os.path.join(HERE, "config_digests.json")

# name -> (dim, min_level, max_level, initial field, velocity, dt, eps, steps)
WORKLOADS = {
    "advection_2d_L4_12_seed0": (2, 4, 12, lambda c: seeded_2d(c, 0), [math.cos(math.pi / 5), math.sin(math.pi / 5)],
                                 0.5 * 2.0 ** -12, 1e-4, 3),
    "advection_3d_L2_8_sphere": (3, 2, 8, lambda c: sphere_3d(c), [1.0, 1.0, 1.0], 0.4 * 2.0 ** -8 / 3, 1e-3, 2),
}
REGULARITY = 2.0


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def record(cells, values):
    return {"leaves": int(len(cells)), "cells_sha256": digest(cells.astype(np.int64)),
            "values_sha256": digest(values.astype(np.float64))}


def run(name):
    dim, lmin, lmax, ic, a, dt, eps, steps = WORKLOADS[name]
    t0 = time.time()
    o = OracleSim(dim, lmin, lmax, pred_s=1)
    o.set_poison(False)
    v0 = ic(o.cells())
    o.set_values(v0)
    o.set_dirichlet(0.0)
    out = {"dim": dim, "min_level": lmin, "max_level": lmax, "velocity": a, "dt": dt, "eps": eps,
           "regularity": REGULARITY, "prediction_radius": 1, "initial_values_sha256": digest(v0), "steps": []}
    for k in range(steps):
        harten = o.adapt(eps, REGULARITY)
        adapted = record(o.cells(), o.values())
        o.upwind_step(np.array(a), dt)
        stepped = record(o.cells(), o.values())
        out["steps"].append({"harten_calls": int(harten), "adapted": adapted, "stepped": stepped})
        print(f"{name}: step {k} ({adapted['leaves']} leaves) {time.time() - t0:.0f} s", flush=True)
    return out


# the frozen-mesh adjoint workload: 32 instances L2..10 (seeds 2000 + k),
# adapted once at t = 0, then 100 upwind steps without adaptation
ADJOINT = {"instances": 32, "seed0": 2000, "min_level": 2, "max_level": 10, "eps": 1e-4, "steps": 100,
           "velocity": [math.cos(math.pi / 5), math.sin(math.pi / 5)]}


def adjoint_instance(k):
    p = ADJOINT
    o = OracleSim(2, p["min_level"], p["max_level"], pred_s=1)
    o.set_poison(False)
    o.set_values(seeded_2d(o.cells(), p["seed0"] + k))
    o.set_dirichlet(0.0)
    o.adapt(p["eps"], REGULARITY)
    adapted = record(o.cells(), o.values())
    dt = 0.5 * 2.0 ** -p["max_level"]
    for _ in range(p["steps"]):
        o.upwind_step(np.array(p["velocity"]), dt)
    return {"adapted": adapted, "forward": record(o.cells(), o.values())}


def run_adjoint():
    from multiprocessing import Pool
    with Pool(min(8, os.cpu_count() or 1)) as pool:
        inst = pool.map(adjoint_instance, range(ADJOINT["instances"]))
    return dict(ADJOINT, regularity=REGULARITY, prediction_radius=1, per_instance=inst)


def main():
    names = sys.argv[1:] or list(WORKLOADS) + ["adjoint_forward_32x_L2_10"]
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for name in names:
        data[name] = run_adjoint() if name == "adjoint_forward_32x_L2_10" else run(name)
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
