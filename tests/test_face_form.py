"""The viscous divergence in face form (CPU, exact arithmetic).

The fused stage's marches do not difference the viscous node fluxes with the
reference's 4-point stencil (src/viscous_contribution_kernel.cu:44-67,
coefficients 1/12, -2/3, 0, 2/3, -1/12 over nodes m-2..m+2); they subtract
the face value h(m+1/2) = 7/12 (v[m] + v[m+1]) - 1/12 (v[m-1] + v[m+2]) from
the face fluxes they already difference (k_fused.hip, visc_face).  The two
agree identically: checked here in rational arithmetic on random integers,
together with the reasons a mistake would show (a face value built from the
wrong neighbours or with a wrong weight fails).
"""
import random
from fractions import Fraction as Fr

C0, C1 = Fr(1, 12), Fr(-2, 3)


def central(v, m):
    """The reference's difference at node m (its sign convention: the sum
    is subtracted from the right-hand side)."""
    return C0 * v[m - 2] + C1 * v[m - 1] - C1 * v[m + 1] - C0 * v[m + 2]


def face(v, m, a=Fr(7, 12), b=Fr(-1, 12)):
    """h(m+1/2) as the marches form it."""
    return a * (v[m] + v[m + 1]) + b * (v[m - 1] + v[m + 2])


def test_face_difference_equals_central_difference():
    rng = random.Random(7)
    for _ in range(200):
        v = [Fr(rng.randint(-10**6, 10**6), rng.randint(1, 1000)) for _ in range(12)]
        for m in range(3, 9):
            assert face(v, m) - face(v, m - 1) == central(v, m)


def test_face_form_is_fourth_order_and_pins_its_weights():
    # on polynomials of degree <= 4 the face difference is exact: for
    # v[k] = k^p it equals the central difference, which is exact for the
    # derivative of degree <= 4 polynomials at spacing 1
    for p in range(5):
        v = [Fr(k) ** p for k in range(12)]
        m = 6
        exact = p * Fr(m) ** (p - 1) if p else Fr(0)
        assert central(v, m) == exact
        assert face(v, m) - face(v, m - 1) == exact
    # the weights are forced: a second-order face value (1/2, 0) or a
    # misplaced pair fails the identity on a cubic
    v = [Fr(k) ** 3 for k in range(12)]
    assert face(v, 6, Fr(1, 2), Fr(0)) - face(v, 5, Fr(1, 2), Fr(0)) != central(v, 6)

    def shifted(v, m):  # outer pair taken one node to the right
        return Fr(7, 12) * (v[m] + v[m + 1]) - Fr(1, 12) * (v[m] + v[m + 3])
    assert shifted(v, 6) - shifted(v, 5) != central(v, 6)
