"""The oracle's conductor Fresnel, Trowbridge-Reitz distribution and sampling
(both semantics) and ray spawning against the closed forms, invariants and
distribution tests of tests/oracle_pins.py."""
import ctypes as C

import pytest

import cpu_oracle as O
import oracle_pins as PINS


@pytest.fixture(scope="module")
def lib():
    return PINS.bind(C.CDLL(str(O.lib()._name)))


@pytest.mark.parametrize("name", sorted(PINS.PINS))
def test_pin(lib, name):
    PINS.PINS[name](lib)
