"""Each plausible mistake planted in the oracle fails at least one pin.

The mutations live in the oracle (oracle/sph_oracle.hpp, enum Mutation) and are
switched on only inside this test, on the CPU.
"""
import pytest

from oracle import oracle as O
from tests import pins


@pytest.mark.parametrize("name", sorted(O.MUTATIONS))
def test_mutation_is_caught(oracle_lib, name):
    failed = []
    with O.set_mutation(name):
        for pin in pins.ALL_PINS:
            for prec in ((64, 32) if "precision" in pin.__code__.co_varnames[:pin.__code__.co_argcount] else (None,)):
                try:
                    pin() if prec is None else pin(prec)
                except AssertionError:
                    failed.append(pin.__name__)
                    break
    assert failed, f"mutation {name} passes every pin"


def test_faithful_oracle_passes_every_pin(oracle_lib):
    assert O.lib().orc_get_mutation() == 0
    for pin in pins.ALL_PINS:
        pin()
