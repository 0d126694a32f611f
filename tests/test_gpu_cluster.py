"""GPU tests of the cluster block-Jacobi PCG (plaMatrixPreconditionerClusterSize, options.h:164; 1..16,
options.cpp:393-394), through the C ABI.

The preconditioner changes how fast PCG converges, not what it converges to: with the reference's stop
(relative residual 2e-10 on the online Schur path, 1e-12 on the assembled path, BundleAdjustPlaMatrix.cpp:326)
every cluster size must follow the FP64 oracle's exact-solve trajectory (same iterations and accepted steps,
final cost within 1e-9). A cluster that covers every primary block is the exact inverse, so each step takes one
PCG iteration (between, a larger block-Jacobi is not always the better one: on the ring, clusters of 4 took 303
iterations against 289 for single blocks). A 13-slot group spans three 6-wide units; its primary block is one
preconditioner block even at cluster size 1, which takes the cluster PCG.
"""
import sys

import numpy as np
import pytest

from plabundle_amd import _abi
from plabundle_amd import options as PO
from plabundle_amd import synthetic as S
from plabundle_amd.solver import Solver, SolverWorkspace

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ws():
    lib = _abi.load_library()
    assert lib.pb_hip_available(0), "no gfx950 device"
    w = SolverWorkspace(0)
    yield w
    w.close()


def _ring(metashape=False, seed=7):
    p = S.ring_problem(seed=seed, num_cameras=14, num_points=300, max_views=6, metashape=metashape)
    o = S.ring_options(PO.SolveOptions(), max_iterations=10)
    o.calibration.refineSharedFocalLength = True
    o.calibration.refineSharedRadialDistortion = True
    if metashape:
        c = o.calibration
        c.refineSharedFocalAspectRatio = c.refineSharedPrincipalPoint = True
        c.refineSharedHighOrderDistortion = c.refineSharedMetashapeParameters = True
    o.backend.reducedSolver = PO.ReducedSolver.Pcg
    return p, o


@pytest.mark.parametrize("online", [True, False])
@pytest.mark.parametrize("cluster", [1, 2, 5, 16])
def test_cluster_sizes_follow_the_oracle(ws, oracle_mod, cluster, online):
    _say(f"case cluster {cluster} online={online}")
    p, o = _ring()
    o.solver.useReferenceOnlineSchur = online
    o.backend.plaMatrixPreconditionerClusterSize = cluster
    ref = oracle_mod.Oracle(p, threads=4).solve(PO.to_c(o))
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    assert r.usable(), r.backendMessage
    rep = r.report
    assert rep["linear_solver_name"] == "pcg"
    assert rep["preconditioner_name"] == ("block-jacobi-6" if cluster == 1 else f"cluster-jacobi-{cluster}")
    assert rep["iterations"] == ref["iterations"]
    assert rep["accepted_steps"] == ref["accepted_steps"]
    assert abs(r.finalCost - ref["final_cost"]) <= 1e-9 * ref["final_cost"]


def test_a_cluster_of_every_primary_block_is_the_exact_inverse(ws):
    _say("case cluster iteration counts")
    p, o = _ring()
    o.solver.useReferenceOnlineSchur = True
    counts = {}
    for cluster in (1, 4, 16):
        o.backend.plaMatrixPreconditionerClusterSize = cluster
        r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
        counts[cluster] = r.report["linear_iterations"]
    _say(f"  PCG iterations by cluster size: {counts}")
    # 12 free cameras and one group fit one cluster of 16 primary blocks: the preconditioner is the whole
    # system's inverse, so each step takes one PCG iteration
    assert counts[16] <= r.report["iterations"]
    assert counts[16] < counts[4] and counts[16] < counts[1]


def test_thirteen_slot_group_is_one_preconditioner_block(ws, oracle_mod):
    _say("case 13-slot group, cluster 1")
    p, o = _ring(metashape=True)
    o.solver.useReferenceOnlineSchur = True
    o.backend.plaMatrixPreconditionerClusterSize = 1
    ref = oracle_mod.Oracle(p, threads=4).solve(PO.to_c(o))
    r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
    assert r.usable(), r.backendMessage
    assert r.report["preconditioner_name"] == "cluster-jacobi-1"
    assert r.report["iterations"] == ref["iterations"]
    assert abs(r.finalCost - ref["final_cost"]) <= 1e-9 * ref["final_cost"]


def test_cluster_pcg_mixed_and_repeatable(ws, oracle_mod):
    _say("case cluster mixed")
    p, o = _ring()
    o.solver.useReferenceOnlineSchur = True
    o.solver.maxIterations = 30
    o.backend.plaMatrixPreconditionerClusterSize = 3
    ref = oracle_mod.Oracle(p, threads=4).solve(PO.to_c(o))
    a = Solver().solve(p, o, ws, precision=PO.Precision.Mixed)
    b = Solver().solve(p, o, ws, precision=PO.Precision.Mixed)
    assert a.usable(), a.backendMessage
    assert a.finalCost == b.finalCost and np.array_equal(a.cam_center, b.cam_center)
    assert abs(a.finalCost - ref["final_cost"]) <= 1e-6 * ref["final_cost"]


def test_cluster_size_out_of_range_is_rejected(ws):
    p, o = _ring()
    for bad in (0, 17):
        o.backend.plaMatrixPreconditionerClusterSize = bad
        r = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
        assert r.abi_status == _abi.PB_ERR_INVALID
