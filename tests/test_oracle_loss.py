"""Pins the robust losses (BundleAdjustPlaMatrixProjection.cpp:75-109) by closed forms and by
the identity weight = 2 d(rho)/ds (the IRLS weight is the loss derivative)."""
import math

import numpy as np
import pytest

LS, HUBER, CAUCHY = 0, 1, 2


@pytest.mark.parametrize("loss", [LS, HUBER, CAUCHY])
def test_weight_is_twice_the_loss_derivative(oracle_mod, loss):
    for scale in (0.5, 1.0, 3.0):
        for s in (1e-3, 0.1, 0.9, 2.0, 7.5, 100.0):
            h = 1e-6 * s
            cp = oracle_mod.robust_loss(s + h, loss, scale)[0]
            cm = oracle_mod.robust_loss(s - h, loss, scale)[0]
            w = oracle_mod.robust_loss(s, loss, scale)[1]
            if loss == HUBER and abs(s - scale * scale) < 2 * h:
                continue
            assert abs(w - (cp - cm) / h) < 1e-6 * max(1.0, w)


def test_least_squares_and_huber_inlier_closed_form(oracle_mod):
    assert oracle_mod.robust_loss(3.0, LS, 1.0) == (1.5, 1.0)
    assert oracle_mod.robust_loss(0.64, HUBER, 1.0) == (0.32, 1.0)  # s <= delta^2 -> least squares
    c, w = oracle_mod.robust_loss(9.0, HUBER, 1.0)                 # |r| = 3, delta = 1
    assert abs(c - (3.0 - 0.5)) < 1e-15 and abs(w - 1.0 / 3.0) < 1e-15


def test_cauchy_closed_form_and_large_scale_limit(oracle_mod):
    c, w = oracle_mod.robust_loss(3.0, CAUCHY, 1.0)
    assert abs(c - 0.5 * math.log(4.0)) < 1e-15 and abs(w - 0.25) < 1e-15
    c, w = oracle_mod.robust_loss(2.0, CAUCHY, 1e6)                # scale -> infinity -> least squares
    assert abs(c - 1.0) < 1e-9 and abs(w - 1.0) < 1e-9


def test_invalid_inputs_rejected(oracle_mod):
    assert oracle_mod.robust_loss(-1.0, LS, 1.0) is None
    assert oracle_mod.robust_loss(float("nan"), LS, 1.0) is None
    assert oracle_mod.robust_loss(1.0, HUBER, 0.0) is None
    assert oracle_mod.robust_loss(1.0, 7, 1.0) is None


def test_observation_weight_whitening(oracle_mod):
    # normalWeight = weight / ms^2 * robust weight; cost uses s = weight / ms^2 * |r|^2
    eye = np.eye(3).reshape(9)
    s = np.array([100.0, 0.0, 50.0, 50.0] + [0.0] * 9)
    lin = oracle_mod.linearize(eye, [0, 0, 0], s, 0, [0.1, 0.0, 5.0], 50.0, 50.0, weight=4.0, scale=2.0,
                               loss=LS, reference_param=False)
    r2 = float(lin["residual"] @ lin["residual"])
    assert abs(lin["weight"] - 1.0) < 1e-15
    assert abs(lin["cost"] - 0.5 * r2) < 1e-12
    bad = oracle_mod.linearize(eye, [0, 0, 0], s, 0, [0.1, 0.0, 5.0], 50.0, 50.0, weight=-1.0)
    assert bad is not None and bad["weight"] == 0.0
