"""Controller invariants of the oracle (BundleAdjustPlaMatrix.cpp:162-631, online Schur branch)."""
import numpy as np

from plabundle_amd import options as PO
from plabundle_amd import synthetic as S


def test_reference_native_scene(oracle_mod):
    # test_native_problem.cpp:51-94: online Schur used, usable, final cost below the initial cost
    p = S.native_three_cameras()
    orc = oracle_mod.Oracle(p, threads=1)
    o = oracle_mod.default_options()
    o.max_iterations = 8
    o.use_online_schur = 1  # the reference scene opts in (test_native_problem.cpp:84)
    rep = orc.solve(o)
    assert rep["online_schur_used"] == 1
    assert rep["solution_usable"] == 1 and rep["solve_status"] in (1, 2)
    assert rep["final_cost"] < rep["initial_cost"]
    # noise-free data: the perturbed camera returns to the truth
    st = orc.state()
    np.testing.assert_allclose(st["cam_center"][2], [2.0, 0.0, 0.0], atol=1e-6)
    np.testing.assert_allclose(st["cam_rotation"][2], np.eye(3).reshape(9), atol=1e-6)


def test_first_step_is_gauss_newton_and_costs_decrease(oracle_mod):
    p = S.ring_problem(seed=3, num_cameras=12, num_points=200)
    orc = oracle_mod.Oracle(p, threads=2)
    o = PO.to_c(S.ring_options(PO.SolveOptions(), max_iterations=1))
    prep = orc.prepare(o)
    P, Ne = prep["primary_block_count"], prep["track_block_count"]
    primary, elim, dd, _, _ = orc.compute_step(P, Ne, 0.0)  # lambda_0 = 0 (BundleAdjustPlaMatrix.cpp:209)
    rep = orc.solve(o)
    assert rep["accepted_steps"] == 1
    assert abs(rep["final_cost"] - orc.cost_after_step(primary, elim, 1.0)) <= 1e-9 * rep["final_cost"]
    costs = []
    for it in (1, 2, 3, 4, 6):
        o.max_iterations = it
        costs.append(oracle_mod.Oracle(p, threads=2).solve(o)["final_cost"])
    assert all(b <= a for a, b in zip(costs, costs[1:]))


def test_cancel_publishes_nothing(oracle_mod):
    p = S.ring_problem(seed=3, num_cameras=10, num_points=100)
    orc = oracle_mod.Oracle(p, threads=1)
    rep = orc.solve(PO.to_c(S.ring_options()), cancel=True)
    assert rep["solve_status"] == 3 and rep["solution_usable"] == 0
    st = orc.state()
    np.testing.assert_array_equal(st["cam_center"], p.cam_center)
    np.testing.assert_array_equal(st["track_point"], p.track_point)


def test_report_counters_are_consistent(oracle_mod):
    p = S.ring_problem(seed=4, num_cameras=16, num_points=300)
    orc = oracle_mod.Oracle(p, threads=4)
    rep = orc.solve(PO.to_c(S.ring_options(max_iterations=10)))
    assert rep["iterations"] == rep["accepted_steps"] + rep["rejected_steps"]
    assert rep["linearizations"] == rep["iterations"]
    assert rep["objective_evaluations"] >= rep["linearizations"] + rep["accepted_steps"]
    assert rep["converged"] == 1 and rep["solve_status"] == 1


def test_thread_count_does_not_change_the_answer(oracle_mod):
    p = S.ring_problem(seed=5, num_cameras=12, num_points=200)
    o = PO.to_c(S.ring_options(max_iterations=5))
    r1 = oracle_mod.Oracle(p, threads=1).solve(o)
    r4 = oracle_mod.Oracle(p, threads=4).solve(o)
    assert abs(r1["final_cost"] - r4["final_cost"]) <= 1e-10 * r1["final_cost"]
