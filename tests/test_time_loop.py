"""The explicit time loop (solve_explicit.c:92-116, check_exit :192-222) restated in
dpgsolver_amd.solver.explicit_time_loop, on the CPU with a recording stepper.

Known answers traced by hand from the reference's loop:
- dt 0.3 to time 1.0: steps 0.3, 0.3, 0.3, then 0.9 + 0.3 > 1 - 1e-12 clamps dt to 0.1 (times 0.3, 0.6,
  0.9, 1.0), and time == time_final ends the loop;
- a step that lands within 1e3 EPS short of time_final is stretched to reach it (the reference's
  clamp test, not min(dt, remaining));
- once clamped, dt stays clamped;
- max_rhs < exit_tol stops at once; "EI" also stops on max_rhs / max_rhs0 < exit_ratio, "E" does not.
"""
import math

import pytest

from dpgsolver_amd.solver import explicit_time_loop


class Rec:
    def __init__(self, rhs=None):
        self.dts = []
        self.rhs = rhs or (lambda k: 1.0)

    def __call__(self, dt):
        self.dts.append(dt)
        return self.rhs(len(self.dts) - 1)


def test_clamp_and_final_time():
    r = Rec()
    t, mx, n = explicit_time_loop(r, 0.3, 1.0)
    assert n == 4 and t == 1.0
    assert r.dts[:3] == [0.3, 0.3, 0.3]
    assert math.isclose(r.dts[3], 1.0 - 0.8999999999999999, rel_tol=0, abs_tol=0)


def test_stretch_within_1e3_eps():
    # time_final a hair (5e-13 < 1e3 EPS = 1e-12) above 4 steps of 0.25: the 4th step is stretched
    r = Rec()
    tf = 1.0 + 5e-13
    t, _, n = explicit_time_loop(r, 0.25, tf)
    assert n == 4 and t == tf
    assert r.dts[3] == tf - 0.75
    # a hair more than 1e3 EPS above: a fifth, tiny step is taken instead
    r2 = Rec()
    tf2 = 1.0 + 4e-12
    t2, _, n2 = explicit_time_loop(r2, 0.25, tf2)
    assert n2 == 5 and t2 == tf2 and r2.dts[4] < 1e-11


def test_exit_tolerances():
    r = Rec(lambda k: 10.0 ** (-k))           # 1, 0.1, 0.01, 1e-3 (not below 1e-3), 1e-4
    t, mx, n = explicit_time_loop(r, 0.1, 100.0, exit_tol=1e-3)
    assert n == 5 and mx == 1e-4 and r.dts == [0.1] * 5
    # EI: ratio to the first step's max_rhs
    r = Rec(lambda k: 2.0 * 0.5 ** k)
    _, mx, n = explicit_time_loop(r, 0.1, 100.0, solver_proc="EI", exit_ratio=0.1)
    assert n == 5 and mx / 2.0 < 0.1
    # E ignores the ratio and runs to the final time
    r = Rec(lambda k: 2.0 * 0.5 ** k)
    t, _, n = explicit_time_loop(r, 0.5, 2.0, solver_proc="E", exit_ratio=0.1)
    assert n == 4 and t == 2.0


def test_bad_arguments():
    with pytest.raises(ValueError):
        explicit_time_loop(Rec(), 0.1, -1.0)
    with pytest.raises(ValueError):
        explicit_time_loop(Rec(), 0.1, 1.0, solver_proc="I")
