"""The oracle's anisotropic and layered-media paths against closed forms that do
not go through its own formulation (tests/closed_forms.py): bulk-Maxwell null
vectors of every Berreman eigenpair (gyrotropic, tilted uniaxial, biaxial,
general passive eps and mu), the gyrotropic normal-incidence modes with their
handedness, the point-source seed against a bulk 3x3 Maxwell solve, the layer-1
dyadic against the angular spectrum with TE/TM Airy reflections (C1 and a
hyperbolic 5-layer stack, all nine entries), the C2 stack's Faraday reflection at
normal incidence, the reference's 3-layer semi-analytical model at C1's
observation heights, and the reference's logged s/p reflection norms."""
import pytest

import closed_forms as cf

CHECKS = dict(cf.all_checks())


@pytest.mark.parametrize("name", list(CHECKS))
def test_closed_form(name):
    CHECKS[name]()


def test_semi_analytical_model_reports_the_reference_tail():
    # the reference's layer-1 frame drops nodes with 2 kappa1 d1 > 700; at C1's
    # closest plane (20 nm) they carry ~1.7e-6 of G, and the check's two sums
    # (reflected, direct) agree to rounding on the nodes it keeps
    refl, direct, dropped = cf.check_semi_analytical_c1()
    assert refl < 1e-13 and direct < 1e-13
    assert 1e-7 < dropped < 1e-5
