"""CPU tests: the oracle's convergence scaling (src/base/scaling.rs:88-135) with user reference
scales (Variables::set_scldom_scale / set_vecdom_scale, vars.rs:203-213, lower bounds of the field
scales) and Variables::scale_off (the unscaled residual |x_new - x_old| / (|x_old| + 1e-10)),
on hand-computed cases."""
import math

import numpy as np

import vfvm_oracle as O


def _flow_case():
    # N = 4 vertices: velocity 0.5 in every component, pressure 1..4; one velocity entry
    # changes by 0.1 and one pressure entry by 0.2
    N = 4
    xo = np.r_[np.full(12, 0.5), [1.0, 2.0, 3.0, 4.0]]
    xn = xo.copy()
    xn[0] += 0.1
    xn[13] += 0.2
    return N, xo, xn


def test_flow_residual_field_scales():
    N, xo, xn = _flow_case()
    vmax = math.sqrt(0.6 ** 2 + 0.5 ** 2 + 0.5 ** 2)          # largest |u| over both states
    r_v = math.sqrt(0.1 ** 2 / N) / vmax                      # 0.0539
    r_p = math.sqrt(0.2 ** 2 / N) / 3.0                       # range 3 > rho U^2 = 0.86
    assert abs(O.scaled_residual_flow(N, xo, xn, np.ones(N)) - max(r_v, r_p)) <= 1e-15


def test_flow_residual_user_scales_are_lower_bounds():
    N, xo, xn = _flow_case()
    got = O.scaled_residual_flow(N, xo, xn, np.ones(N), scale_vel=10.0, scale_pres=50.0)
    assert abs(got - max(0.05 / 10.0, 0.1 / 50.0)) <= 1e-15
    # a user scale below the field's own scale changes nothing
    base = O.scaled_residual_flow(N, xo, xn, np.ones(N))
    assert O.scaled_residual_flow(N, xo, xn, np.ones(N), scale_vel=1e-3, scale_pres=1e-3) == base


def test_scale_off_is_the_unscaled_norm():
    N, xo, xn = _flow_case()
    expect = math.sqrt(0.1 ** 2 + 0.2 ** 2) / (math.sqrt(12 * 0.25 + 1 + 4 + 9 + 16) + 1e-10)
    assert abs(O.scaled_residual_flow(N, xo, xn, np.ones(N), scale_off=True) - expect) <= 1e-15
    assert abs(O.scaled_residual_scalar(xo, xn, scale_off=True) - expect) <= 1e-15
    bad = xn.copy()
    bad[3] = np.nan
    assert math.isnan(O.scaled_residual_flow(N, xo, bad, np.ones(N), scale_off=True))


def test_scalar_residual_user_scale():
    xo = np.array([1.0, -2.0, 0.5])
    xn = np.array([1.3, -2.0, 0.5])
    assert abs(O.scaled_residual_scalar(xo, xn) - math.sqrt(0.09 / 3) / 2.0) <= 1e-15
    assert abs(O.scaled_residual_scalar(xo, xn, scale=8.0) - math.sqrt(0.09 / 3) / 8.0) <= 1e-15
