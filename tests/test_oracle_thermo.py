"""Runs the independent SI pins of the oracle's thermodynamic equations, terminal velocities, State
update and qcond recompute (tests/oracle_pins.py) against the checker build of the oracle."""
import pytest

from tests import oracle_pins


@pytest.mark.parametrize("pin", oracle_pins.PINS, ids=lambda f: f.__name__)
def test_oracle_pin(oracle, pin):
    pin(oracle)
