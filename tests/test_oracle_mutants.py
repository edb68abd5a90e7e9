"""Each closed-form check must be able to fail: builds of the oracle with one
planted mistake (ORACLE_MUTANT in oracle/ytf_oracle.cpp, built by
`make -C oracle mutants`) are run through every check of tests/closed_forms.py,
and each mutant must be rejected by the checks aimed at it.

  1  eps_12 and eps_21 exchanged in the Berreman matrix (eps -> eps^T, g -> -g)
  2  the eps_tz term of the point-source seed dropped
  3  eps_t and eps_z exchanged in the uniaxial extraordinary k_z
"""
import os
import subprocess

import pytest

import closed_forms as cf
from helpers import po, ROOT

ORACLE = os.path.join(ROOT, "oracle")
MUTANTS = {
    1: {"null vectors, gyrotropic", "null vectors, general eps and mu", "gyrotropic normal-incidence modes",
        "gyrotropic normal-incidence reflection, C2"},
    2: {"particular solution, tilted uniaxial", "particular solution, general"},
    3: {"layer-1 dyadic, C1", "layer-1 dyadic, hyperbolic 5-layer", "3-layer semi-analytical model, C1",
        "logged s/p reflections"},
}


@pytest.fixture(scope="module")
def mutant_libs():
    p = subprocess.run(["make", "-C", ORACLE, "-j3", "mutants"], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    return {m: os.path.join(ORACLE, "_mutants", f"liboracle_m{m}.so") for m in MUTANTS}


def failing_checks():
    failed = set()
    for name, fn in cf.all_checks():
        try:
            fn()
        except (AssertionError, RuntimeError, ValueError, FloatingPointError):
            failed.add(name)
    return failed


def test_the_default_oracle_passes_every_check():
    assert failing_checks() == set()


@pytest.mark.parametrize("m", sorted(MUTANTS))
def test_each_planted_mistake_is_caught(mutant_libs, m):
    with po.using(mutant_libs[m]):
        failed = failing_checks()
    missed = MUTANTS[m] - failed
    assert not missed, f"mutant {m} passed checks aimed at it: {sorted(missed)}"
