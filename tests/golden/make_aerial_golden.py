"""Writes tests/golden/aerial_oracle.json: the FP64 oracle's result on the full aerial workload
(1000 cameras, 200k points, ~1.5M observations, seed 2, Cauchy 1 px, shared focal + Brown refined,
20 iterations) together with a fingerprint of the generated problem.

Run from the repository root: python tests/golden/make_aerial_golden.py
"""
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from oracle import oracle as O  # noqa: E402
from plabundle_amd import options as PO  # noqa: E402
from plabundle_amd import synthetic as S  # noqa: E402


def fingerprint(p) -> str:
    h = hashlib.sha256()
    for a in (p.cam_rotation, p.cam_center, p.cam_intrinsics, p.track_point, p.track_offset, p.obs_camera, p.obs_uv):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main():
    t = time.time()
    p = S.aerial_problem()
    gen = time.time() - t
    opts = PO.to_c(S.aerial_options(PO.SolveOptions()))
    orc = O.Oracle(p)
    t = time.time()
    rep = orc.solve(opts)
    wall = time.time() - t
    out = {
        "workload": "aerial block: 1000 cameras in 3 strips, 200000 points, seed 2, Cauchy 1 px, "
                    "shared focal + Brown refined, 20 iterations",
        "fingerprint": fingerprint(p),
        "num_observations": p.num_observations,
        "oracle_threads": O.lib().oracle_max_threads(),
        "generation_seconds": gen,
        "solve_wall_seconds": wall,
    }
    for k in ("solve_status", "converged", "iterations", "initial_cost", "final_cost", "accepted_steps",
              "rejected_steps", "linearizations", "objective_evaluations", "observation_count",
              "active_track_count", "rejected_initial_tracks", "primary_block_count", "track_block_count",
              "solve_seconds", "assembly_seconds", "linear_solve_seconds"):
        out[k] = rep[k]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "aerial_oracle.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
