"""Host logic of the drivers (CPU): the reference's stop test restated in
drivers.ConvergeCriteria (snippet::converge_criteria, utils/snippet.h:156-213), checked
on hand-worked sequences."""
import math

from topoptspline3d_amd.drivers import ConvergeCriteria


def test_constant_history_stops_at_the_ninth_iteration():
    # iterations 1-3 see NaN history (never converged); from the 4th on fch = 0, the
    # counter climbs 1..6 and the test passes once it exceeds 5: at the 9th update
    cc = ConvergeCriteria(1, 5, 1e-3)
    out = [cc.update(2.0, [0.4]) for _ in range(9)]
    assert out == [False] * 8 + [True]
    assert cc.counter == 6 and cc.fch == 0.0


def test_change_is_the_four_point_difference():
    cc = ConvergeCriteria(1, 5, 1e-3)
    for c in (10.0, 9.0, 8.5):
        cc.update(c, [0.4])
    cc.update(8.4, [0.4])
    assert math.isclose(cc.fch, abs(8.4 + 8.5 - 9.0 - 10.0) / (9.0 + 10.0))
    assert cc.counter == 0


def test_counter_falls_on_unconverged_iterations_and_constraints_count():
    cc = ConvergeCriteria(1, 5, 1e-3)
    for _ in range(6):
        cc.update(1.0, [0.5])
    assert cc.counter == 3
    cc.update(1.5, [0.5])          # objective jumps: counter 3 -> 2
    assert cc.counter == 2
    cc.update(1.5, [0.5])          # c: 1.5 + 1.5 - 1 - 1 over 2: still large -> 1
    cc.update(1.5, [0.5])          # 1.5+1.5-1.5-1 = 0.5 over 2.5 -> 0
    assert cc.counter == 0
    cc.update(1.5, [0.5])          # fch 0 -> 1
    assert cc.counter == 1
    cc.update(1.5, [0.6])          # constraint change (0.6+0.5-0.5-0.5)/1 = 0.1 > 1e-3 -> 0
    assert cc.counter == 0


def test_history_wraps_after_twenty_iterations():
    # 30 iterations of a geometric decay: every update uses the last four values,
    # also after the 20-entry circular history wraps
    cc = ConvergeCriteria(1, 100, 1e-9)
    vals = [100.0 * 0.9 ** k for k in range(30)]
    for k, v in enumerate(vals):
        cc.update(v, [0.4])
        if k >= 3:
            want = abs(vals[k] + vals[k - 1] - vals[k - 2] - vals[k - 3]) / (vals[k - 2] + vals[k - 3])
            assert math.isclose(cc.fch, want, rel_tol=1e-12)
