"""Known-answer tests that pin the oracle's building blocks independently of
its own code: textbook values of the WENO candidate stencils and of the
6th-order derivative, the reference's branch rule on hand-built windows,
and exact symmetries."""
import math

import numpy as np
import pytest

ULP = 2.0 ** -52


def test_weno_constant_data(orc):
    for c in (1.0, -3.5, 1e5, 4.0e18):
        for d in (1, -1):
            r = orc.weno([c] * 5, d)
            assert abs(r - c) <= 4 * ULP * abs(c)


@pytest.mark.parametrize("a,b", [(0.0, 1.0), (7.0, -2.0), (1e5, 3.0), (-2.5, 0.125)])
def test_weno_linear_data_gives_exact_face_value(orc, a, b):
    # Q_k = a + b k, k = 1..5 centred on k = 3: all three smoothness
    # indicators equal b^2, so the weights are the linear ones and the
    # 5th-order reconstruction is exact: a + 3.5 b at the right face of the
    # centre cell (dir > 0) and a + 2.5 b at its left face (dir < 0).
    q = [a + b * k for k in range(1, 6)]
    scale = abs(a) + 5 * abs(b)
    assert abs(orc.weno(q, 1) - (a + 3.5 * b)) <= 8 * ULP * scale
    assert abs(orc.weno(q, -1) - (a + 2.5 * b)) <= 8 * ULP * scale


def test_weno_mirror_identity(orc):
    # dir < 0 on a window is dir > 0 on the reversed window, bit for bit
    rng = np.random.default_rng(0)
    for _ in range(2000):
        q = rng.normal(size=5) * 10.0 ** rng.integers(-3, 6)
        assert orc.weno(q, -1) == orc.weno(q[::-1].copy(), 1)


def test_weno_branch_one_drops_third_stencil(orc):
    # Q = (1,2,3,4,8): beta = (1, 1, 10) -> tau40 = 0 <= min beta = 1 and
    # tau41 = 9 > 1, so omega = (1/3, 2/3, 0) and the value is
    # (q0 + 2 q1)/3 with the Jiang-Shu candidates q0 = Q1/3 - 7Q2/6 + 11Q3/6 = 3.5,
    # q1 = -Q2/6 + 5Q3/6 + Q4/3 = 3.5.  The WENO-Z weights would mix in
    # q2 = Q3/3 + 5Q4/6 - Q5/6 = 3 and give a different value.
    assert abs(orc.weno([1.0, 2.0, 3.0, 4.0, 8.0], 1) - 3.5) <= 4 * ULP * 3.5


def test_weno_branch_two_drops_first_stencil(orc):
    # Q = (3,1,2,3,4): beta = (16, 1, 1) -> tau40 = 15 > 1, tau41 = 0 < 1, so
    # omega = (0, 2/3, 1/3): value (2 q1 + q2)/3 with q1 = 2.5, q2 = 2.5.
    assert abs(orc.weno([3.0, 1.0, 2.0, 3.0, 4.0], 1) - 2.5) <= 4 * ULP * 2.5


def test_weno_general_branch_is_convex_combination(orc):
    rng = np.random.default_rng(1)
    for _ in range(500):
        q = rng.normal(size=5)
        s = [q[0] / 3 - 7 * q[1] / 6 + 11 * q[2] / 6, -q[1] / 6 + 5 * q[2] / 6 + q[3] / 3,
             q[2] / 3 + 5 * q[3] / 6 - q[4] / 6]
        r = orc.weno(q, 1)
        assert min(s) - 1e-12 <= r <= max(s) + 1e-12


def test_d6_exact_on_polynomials(orc):
    # the 6th-order centred first derivative is exact for degree <= 6
    x0 = 2.0
    for k in range(0, 7):
        f = [(x0 + m) ** k for m in range(-3, 4)]
        exact = k * x0 ** (k - 1) if k else 0.0
        assert abs(orc.d6(f, 1.0) - exact) <= 1e-12 * max(1.0, abs(exact))
    f7 = [(x0 + m) ** 7 for m in range(-3, 4)]
    assert abs(orc.d6(f7, 1.0) - 7 * x0 ** 6) > 1.0  # degree 7 is not exact


def test_d6_scales_with_spacing(orc):
    # truncation error of the 6th-order stencil: h^6 |f'''''''| / 140
    f = [math.sin(0.1 * m) for m in range(-3, 4)]
    err = abs(orc.d6(f, 0.1) - 1.0)
    assert 1e-9 < err < 0.1 ** 6 / 140 * 1.01
