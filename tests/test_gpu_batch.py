"""GPU parity of the batched-window entry (pb_hip_solve_batch) against the FP64 oracle, through the C ABI.

Each window of a batch must come out as the oracle solves that window alone (same options): the oracle's
reduced solve is a dense Cholesky, as the batched path's, so the FP64 bar is the single-problem one (final
cost within 1e-9 relative, same iteration and accepted-step counts, refined state within 1e-7 m / 1e-6 m);
the mixed bar is 1e-6 relative final cost.
"""
import sys

import numpy as np
import pytest

from plabundle_amd import _abi
from plabundle_amd import options as PO
from plabundle_amd import synthetic as S
from plabundle_amd.problem import GaugePolicy
from plabundle_amd.solver import Solver, SolverWorkspace, SolveStatus

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


def _windows(n=6, seed=11):
    return S.batch_windows(seed=seed, num_windows=n, cameras=(8, 20), points=(150, 500))


@pytest.mark.parametrize("intrinsics", [True, False])
def test_batch_fp64_matches_oracle_per_window(ws, oracle_mod, intrinsics):
    _say(f"case batch fp64 intrinsics={intrinsics}")
    wins = _windows()
    o = S.window_options(PO.SolveOptions(), max_iterations=10)
    if not intrinsics:
        o.calibration.refineSharedFocalLength = False
        o.calibration.refineSharedRadialDistortion = False
    res, summary = Solver().solve_batch(wins, o, ws, precision=PO.Precision.Fp64)
    assert summary["abi_status"] == _abi.PB_OK, summary["message"]
    co = PO.to_c(o, PO.Precision.Fp64)
    for w, (p, r) in enumerate(zip(wins, res)):
        orc = oracle_mod.Oracle(p, threads=2)
        ref = orc.solve(co)
        st = orc.state()
        rel = abs(r.finalCost - ref["final_cost"]) / ref["final_cost"]
        assert r.report["fp64_iterations"] == r.report["iterations"]
        _say(f"window {w}: {p.num_cameras} cameras {p.num_observations} obs, {r.report['iterations']} it "
             f"(oracle {ref['iterations']}), rel cost {rel:.2e}, {r.report['device_seconds'] * 1e3:.2f} ms")
        assert r.status == SolveStatus(ref["solve_status"])
        assert r.report["iterations"] == ref["iterations"]
        assert r.report["accepted_steps"] == ref["accepted_steps"]
        assert abs(r.report["initial_cost"] - ref["initial_cost"]) <= 1e-12 * ref["initial_cost"]
        assert rel <= 1e-9
        np.testing.assert_allclose(r.cam_center, st["cam_center"], atol=1e-7)
        np.testing.assert_allclose(r.track_point, st["track_point"], atol=1e-6)
        np.testing.assert_allclose(r.cam_intrinsics, st["cam_intrinsics"], rtol=1e-7, atol=1e-9)
        np.testing.assert_array_equal(r.track_valid, st["track_valid"])


def test_batch_mixed_within_north_star_tolerance(ws, oracle_mod):
    # Mixed windows assemble in FP32 while their steps still lower the objective by 1e-2 or more and finish in
    # FP64 (a window's stop is only taken in FP64). Windows that both the batch and the oracle end as converged
    # must agree to the north-star bar; a window that the reference itself ends by an exhausted line search
    # (NoConvergence after a few iterations) stops wherever its trajectory stalls, so it is held to 1e-4.
    _say("case batch mixed")
    wins = _windows(seed=12) + _windows(n=6, seed=17)
    o = S.window_options(PO.SolveOptions(), max_iterations=20)
    res, summary = Solver().solve_batch(wins, o, ws, precision=PO.Precision.Mixed)
    assert summary["abi_status"] == _abi.PB_OK, summary["message"]
    co = PO.to_c(o)
    both = 0
    for w, (p, r) in enumerate(zip(wins, res)):
        ref = oracle_mod.Oracle(p, threads=2).solve(co)
        rel = abs(r.finalCost - ref["final_cost"]) / ref["final_cost"]
        _say(f"window {w}: mixed rel cost {rel:.2e}, {r.report['iterations']} it ({r.report['fp64_iterations']} "
             f"in FP64) status {r.status.name}; oracle {ref['iterations']} it status {ref['solve_status']}")
        assert r.usable()
        assert r.report["fp64_iterations"] >= 1
        if r.status == SolveStatus.Success and ref["solve_status"] == int(SolveStatus.Success):
            both += 1
            assert rel <= 1e-6
        else:
            assert rel <= 1e-4
    assert both >= len(wins) // 2


def test_batch_auto_anchor_gauge_matches_single_solves(ws):
    # AutoAnchor: each window fixes its camera 0 and moves its farthest camera on the baseline sphere (the
    # per-window gauge scale camera); the batch must match the single-problem path window by window
    _say("case batch auto-anchor vs single")
    wins = []
    for p in _windows(n=4, seed=13):
        q = p.copy()
        q.fixed_cameras = None
        q.gauge_policy = GaugePolicy.AutoAnchor
        wins.append(q)
    o = S.window_options(PO.SolveOptions(), max_iterations=8)
    res, summary = Solver().solve_batch(wins, o, ws, precision=PO.Precision.Fp64)
    assert summary["abi_status"] == _abi.PB_OK, summary["message"]
    for w, (p, r) in enumerate(zip(wins, res)):
        single = Solver().solve(p, o, ws, precision=PO.Precision.Fp64)
        rel = abs(r.finalCost - single.finalCost) / single.finalCost
        _say(f"window {w}: batch {r.finalCost:.10g} single {single.finalCost:.10g} rel {rel:.2e}")
        assert single.report["gauge_scale_camera"] >= 0
        assert r.report["iterations"] == single.report["iterations"]
        assert rel <= 1e-8
        np.testing.assert_allclose(r.cam_center, single.cam_center, atol=1e-6)


def test_batch_window_without_usable_tracks(ws, oracle_mod):
    _say("case batch empty window")
    wins = _windows(n=3, seed=14)
    bad = wins[1].copy()
    bad.obs_uv = np.full_like(bad.obs_uv, np.nan)
    wins[1] = bad
    o = S.window_options(PO.SolveOptions(), max_iterations=10)
    res, summary = Solver().solve_batch(wins, o, ws, precision=PO.Precision.Fp64)
    assert summary["abi_status"] == _abi.PB_OK, summary["message"]
    assert res[1].status == SolveStatus.InvalidInput and not res[1].usable()
    co = PO.to_c(o, PO.Precision.Fp64)
    for w in (0, 2):
        ref = oracle_mod.Oracle(wins[w], threads=2).solve(co)
        assert abs(res[w].finalCost - ref["final_cost"]) <= 1e-9 * ref["final_cost"]


def test_batch_rejects_oversized_and_invalid_windows(ws):
    _say("case batch rejects")
    wins = _windows(n=2, seed=15)
    big = S.window_problem(np.random.Generator(np.random.PCG64(3)), 90, 300)
    o = S.window_options(PO.SolveOptions(), max_iterations=5)
    res, summary = Solver().solve_batch([wins[0], big], o, ws, precision=PO.Precision.Fp64)
    assert summary["abi_status"] == _abi.PB_ERR_UNSUPPORTED
    assert "window 1" in res[0].backendMessage
    bad = wins[1].copy()
    bad.track_offset = bad.track_offset.copy()
    bad.track_offset[0] = 1
    res, summary = Solver().solve_batch([wins[0], bad], o, ws, precision=PO.Precision.Fp64)
    assert summary["abi_status"] == _abi.PB_ERR_INVALID
    assert "window 1" in res[0].backendMessage


def test_batch_cancel_publishes_nothing(ws):
    _say("case batch cancel")
    wins = _windows(n=2, seed=16)
    o = S.window_options(PO.SolveOptions(), max_iterations=5)
    res, summary = Solver().solve_batch(wins, o, ws, precision=PO.Precision.Fp64, cancel=True)
    assert summary["solve_status"] == int(SolveStatus.Cancelled)
    assert all(not r.usable() for r in res)


@pytest.mark.parametrize("precision", [PO.Precision.Fp64, PO.Precision.Mixed])
def test_batch_graph_replay_matches_direct_launches(ws, precision):
    # every round replays the captured head / trial-pass / tail graphs, with the host's polls and syncs between
    # replays; the outcome of every window must be bitwise that of direct launches, and the same on a second
    # batch solve (new captures)
    _say(f"case batch graphs vs direct launches precision={precision.name}")
    wins = _windows(n=8, seed=13)
    o = S.window_options(PO.SolveOptions(), max_iterations=12)
    o.backend.batchGraphs = False
    direct, sd = Solver().solve_batch(wins, o, ws, precision=precision)
    o.backend.batchGraphs = True
    for rep in range(2):
        graph, sg = Solver().solve_batch(wins, o, ws, precision=precision)
        assert sd["abi_status"] == sg["abi_status"] == _abi.PB_OK
        assert sd["iterations"] == sg["iterations"]
        for a, b in zip(direct, graph):
            assert a.report["iterations"] == b.report["iterations"]
            assert a.finalCost == b.finalCost
            np.testing.assert_array_equal(a.cam_center, b.cam_center)
            np.testing.assert_array_equal(a.track_point, b.track_point)


@pytest.mark.parametrize("precision", [PO.Precision.Fp64, PO.Precision.Mixed])
def test_batch_factor_shapes_are_bitwise_equal_and_timed(ws, precision):
    # every launch shape of the dense window factorisation updates each value in the same order, so fixed
    # shapes and the timed choice give bitwise the same windows; the windows span 8 to 64 cameras (systems of
    # up to 402 unknowns: every row slot of the 128-thread shapes is used)
    _say(f"case batch factorisation shapes precision={precision.name}")
    rng = np.random.Generator(np.random.PCG64(21))
    wins = _windows(n=4, seed=17) + [S.window_problem(rng, 64, 1500), S.window_problem(rng, 41, 900)]
    o = S.window_options(PO.SolveOptions(), max_iterations=6)
    o.backend.batchFactorTiles = False
    runs = []
    for shape in range(_abi.PB_BATCH_FACTOR_SHAPES):
        o.backend.batchFactorShape = shape
        res, summ = Solver().solve_batch(wins, o, ws, precision=precision)
        assert summ["abi_status"] == _abi.PB_OK, summ["message"]
        runs.append(res)
    for shape, res in enumerate(runs[1:], start=1):
        for w, (a, b) in enumerate(zip(runs[0], res)):
            assert a.report["iterations"] == b.report["iterations"], (shape, w)
            assert a.finalCost == b.finalCost, (shape, w)
            np.testing.assert_array_equal(a.cam_center, b.cam_center)
            np.testing.assert_array_equal(a.cam_intrinsics, b.cam_intrinsics)
            np.testing.assert_array_equal(a.track_point, b.track_point)
    o.backend.batchFactorShape = -1
    ws2 = SolverWorkspace(0)  # a fresh context times the shapes on its first batch
    try:
        for rep in range(2):
            res, summ = Solver().solve_batch(wins, o, ws2, precision=precision)
            assert summ["abi_status"] == _abi.PB_OK
            tuning = summ["factor_tuning"]
            _say(f"  solve {rep}: {tuning}")
            names = ["fp64"] + (["fp32"] if precision == PO.Precision.Mixed else [])
            for name in names:
                t = tuning[name]
                assert 0 <= t["shape"] < _abi.PB_BATCH_FACTOR_SHAPES
                assert len(t["ms"]) == _abi.PB_BATCH_FACTOR_SHAPES and min(t["ms"].values()) > 0.0
                assert t["ms"][f"{t['threads_panel'][0]}x{t['threads_panel'][1]}"] == min(t["ms"].values())
            if rep == 0:
                first = tuning
            else:
                assert tuning == first  # same batch class: the second batch keeps the timed shapes
            for a, b in zip(runs[0], res):
                assert a.finalCost == b.finalCost
                np.testing.assert_array_equal(a.track_point, b.track_point)
        # a batch of another class (2 windows, not 6) times the shapes again
        res, summ = Solver().solve_batch(wins[-2:], o, ws2, precision=precision)
        assert summ["abi_status"] == _abi.PB_OK
        assert summ["factor_tuning"]["fp64"]["ms"] != first["fp64"]["ms"]
    finally:
        ws2.close()
    o.backend.batchFactorShape = _abi.PB_BATCH_FACTOR_SHAPES
    res, summ = Solver().solve_batch(wins, o, ws, precision=precision)
    assert summ["abi_status"] == _abi.PB_ERR_INVALID


@pytest.mark.parametrize("precision", [PO.Precision.Fp64, PO.Precision.Mixed])
def test_tile_factorisation_matches_panels_and_the_oracle(ws, oracle_mod, precision):
    # the MFMA tile factorisation (default) and the panel factorisation order their sums differently: windows of 8
    # to 64 cameras (systems of 48 to 402 unknowns: one to seven 64-row tiles, padded) must reach the same
    # trajectory and final costs within rounding, the FP64 ones the oracle's; repeated tile batches are bitwise
    # identical
    _say(f"case batch tile factorisation precision={precision.name}")
    rng = np.random.Generator(np.random.PCG64(23))
    wins = _windows(n=4, seed=19) + [S.window_problem(rng, 64, 1500), S.window_problem(rng, 41, 900),
                                      S.window_problem(rng, 8, 500), S.window_problem(rng, 11, 600)]
    o = S.window_options(PO.SolveOptions(), max_iterations=12)
    o.backend.batchFactorTiles = True
    tiles, st = Solver().solve_batch(wins, o, ws, precision=precision)
    again, _ = Solver().solve_batch(wins, o, ws, precision=precision)
    o.backend.batchFactorTiles = False
    o.backend.batchFactorShape = 0
    panels, sp = Solver().solve_batch(wins, o, ws, precision=precision)
    assert st["abi_status"] == sp["abi_status"] == _abi.PB_OK, st["message"]
    assert st["factor_tuning"]["fp64"] is None  # nothing to time
    for w, (a, b, c) in enumerate(zip(tiles, again, panels)):
        assert a.finalCost == b.finalCost and np.array_equal(a.track_point, b.track_point), w
        assert a.report["iterations"] == c.report["iterations"], w
        assert abs(a.finalCost - c.finalCost) <= 1e-10 * c.finalCost, w
        np.testing.assert_allclose(a.cam_center, c.cam_center, atol=1e-8)
        if precision == PO.Precision.Fp64:
            ref = oracle_mod.Oracle(wins[w], threads=4).solve(PO.to_c(S.window_options(PO.SolveOptions(),
                                                                                       max_iterations=12)))
            assert a.report["iterations"] == ref["iterations"], w
            assert abs(a.finalCost - ref["final_cost"]) <= 1e-9 * ref["final_cost"], w
