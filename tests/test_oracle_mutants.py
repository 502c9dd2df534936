"""Each pin of tests/oracle_pins.py fails on the mistake it is meant to
catch.  Every case compiles a copy of oracle/sparc_oracle.c with one planted
mistake (ORC_MUTANT=n, on the CPU, into a temporary directory) and runs the
named pins on it; each must raise.  The committed oracle passes all of them
(tests/test_oracle_physics.py)."""
from concurrent.futures import ThreadPoolExecutor

import pytest

import oracle
from tests import oracle_pins as P

# mutant -> (what it plants, the pins that must fail on it)
MUTANTS = {
    1: ("sign of rho' g in RHS v_z (physics2d.f90:1044-1046)", [P.force_free, P.horizontal_mode, P.rho0_symmetry]),
    2: ("sign of xi_z dp0/dz in p (:998)", [P.force_free, P.rho0_symmetry]),
    3: ("flow terms without the factor 2 (:1065-1079)", [P.flow_increment]),
    4: ("v0x and v0z exchanged in the flow terms", [P.flow_increment]),
    5: ("x-PML auxiliary from the uncorrected derivative (:970-986)", [P.pml_auxiliaries]),
    6: ("z-PML auxiliary from the corrected derivative (:957-967)", [P.pml_auxiliaries]),
    7: ("sign of xi_z drho0/dz in rho' (:1043)", [P.force_free]),
    8: ("K_psi with one factor dimen (kernels.f90:511-512)", [P.kernels_restatement, P.kernels_chain_rule]),
    9: ("d K_x/dz of the curl with ibc 1 instead of bcz = 0 (:510)", [P.kernels_restatement]),
    10: ("ddzkern without FILTER_VAR_Z (:648-661)", [P.kernels_restatement]),
    11: ("K_c without its factor 2 (:579)", [P.kernels_restatement, P.kernels_finite_differences]),
    12: ("K_v with the opposite sign (:520-522)", [P.kernels_restatement, P.kernels_chain_rule,
                                                  P.kernels_finite_differences]),
    13: ("Hessian without the acc_x term (:293-294)", [P.kernels_restatement]),
    14: ("top closure d xi_z/dz = +d xi_x/dx (physics2d.f90:954)", [P.horizontal_mode]),
}


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_mutants")
    P.born_data()                       # finite differences with the committed oracle
    with ThreadPoolExecutor(8) as ex:
        paths = list(ex.map(lambda m: oracle.build_variant(m, d), [0, *MUTANTS]))
    return dict(zip([0, *MUTANTS], paths))


ALL_PINS = [P.force_free, P.horizontal_mode, P.flow_increment, P.rho0_symmetry, P.pml_auxiliaries,
            P.kernels_restatement, P.kernels_chain_rule, P.kernels_finite_differences]


def test_unmutated_copy_passes_every_pin(mutant_libs):
    with oracle.using(mutant_libs[0]):
        for pin in ALL_PINS:
            pin()


@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_pins_catch_planted_mistake(mutant, mutant_libs):
    what, pins = MUTANTS[mutant]
    with oracle.using(mutant_libs[mutant]):
        for pin in pins:
            with pytest.raises(AssertionError):
                pin()
