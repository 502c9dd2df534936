"""Pins of the oracle's RHS and PRODUCE_KERNELS against what the
mathematics fixes (tests/oracle_pins.py holds the checks and their
derivations): a force-free hydrostatic translation, a horizontal xi_x mode
in closed form, the flow increments, the PML auxiliary equations, a numpy
restatement of the kernel sums on the reference-built dbyd2, the
psi <-> v0 chain rule of the kernels, and central finite differences of a
linear functional of the traces against the kernels."""
import pytest

from tests import oracle_pins as P


def test_rhs_hydrostatic_translation_is_force_free():
    P.force_free()


def test_rhs_horizontal_mode_closed_form():
    P.horizontal_mode()


def test_rhs_flow_increment_closed_form():
    P.flow_increment()


def test_rhs_rho0_weighted_symmetry():
    P.rho0_symmetry()


def test_rhs_pml_auxiliary_equations():
    P.pml_auxiliaries()


def test_kernels_numpy_restatement():
    P.kernels_restatement()


def test_kernels_psi_chain_rule():
    P.kernels_chain_rule()


def test_kernels_against_finite_differences():
    ec, ev = P.kernels_finite_differences()
    # measured: sound speed 0.2 %, flow 2 % (the recipe's continuous-adjoint accuracy)
    assert ec < 0.02 and ev < 0.05
