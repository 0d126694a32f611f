"""Writes the FP64 oracle's results on the full aerial workload (1000 cameras, 200k points, ~1.5M observations,
seed 2, Cauchy 1 px, shared focal + Brown refined) together with a fingerprint of the generated problem:

  tests/golden/aerial_oracle.json       scalars of two oracle runs: "converged_run", to the reference's convergence
                                        stop (step RMS <= 1e-6, BundleAdjustPlaMatrix.cpp:446-466,524-526; budget
                                        400 iterations), and "budget20_run", at the reference's default budget of 20
                                        iterations (whose fields are also kept at the top level)
  tests/golden/aerial_oracle_state.npz  the converged refined state: every camera's rotation and centre, the four
                                        groups' refined intrinsics (one camera per group), every 100th track's point
                                        and valid flag, and the per-axis sums of all valid points

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

HERE = os.path.dirname(os.path.abspath(__file__))
POINT_STRIDE = 100


def fingerprint(p) -> str:
    h = hashlib.sha256()
    for a in (p.cam_rotation, p.cam_center, p.cam_intrinsics, p.track_point, p.track_offset, p.obs_camera, p.obs_uv):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


KEYS = ("solve_status", "converged", "iterations", "initial_cost", "final_cost", "accepted_steps", "rejected_steps",
        "linearizations", "objective_evaluations", "observation_count", "active_track_count", "rejected_initial_tracks",
        "primary_block_count", "track_block_count", "solve_seconds", "assembly_seconds", "linear_solve_seconds")


def run(p, max_iterations):
    opts = PO.to_c(S.aerial_options(PO.SolveOptions()))
    opts.max_iterations = max_iterations
    orc = O.Oracle(p)
    t = time.time()
    rep = orc.solve(opts)
    wall = time.time() - t
    out = {k: rep[k] for k in KEYS}
    out["solve_wall_seconds"] = wall
    return out, orc.state()


def main():
    t = time.time()
    p = S.aerial_problem()
    gen = time.time() - t
    conv, st = run(p, 400)
    budget, _ = run(p, 20)
    out = {
        "workload": "aerial block: 1000 cameras in 3 strips, 200000 points, seed 2, Cauchy 1 px, "
                    "shared focal + Brown refined",
        "fingerprint": fingerprint(p),
        "num_observations": p.num_observations,
        "oracle_threads": O.lib().oracle_max_threads(),
        "generation_seconds": gen,
        "converged_run": conv,
        "budget20_run": budget,
        # the fields of the previous fixture layout: the 20-iteration run
        **{k: budget[k] for k in KEYS},
    }
    valid = st["track_valid"] != 0
    group_cams = [int(np.nonzero(p.cam_group == g)[0][0]) for g in range(int(p.cam_group.max()) + 1)]
    np.savez_compressed(
        os.path.join(HERE, "aerial_oracle_state.npz"),
        cam_rotation=st["cam_rotation"], cam_center=st["cam_center"],
        group_cameras=np.array(group_cams, dtype=np.int32), group_intrinsics=st["cam_intrinsics"][group_cams],
        point_index=np.arange(0, p.num_tracks, POINT_STRIDE, dtype=np.int64),
        track_point=st["track_point"][::POINT_STRIDE], track_valid=st["track_valid"][::POINT_STRIDE],
        valid_point_sum=st["track_point"][valid].sum(axis=0), valid_count=np.array([int(valid.sum())]))
    with open(os.path.join(HERE, "aerial_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
