"""The oracle against closed forms, golden vectors, brute force and invariants."""
import inspect

import pytest

from tests import pins


def _cases():
    out = []
    for f in pins.ALL_PINS:
        if "precision" in inspect.signature(f).parameters:
            out += [pytest.param(f, 32, id=f"{f.__name__}-fp32"), pytest.param(f, 64, id=f"{f.__name__}-fp64")]
        else:
            out.append(pytest.param(f, None, id=f.__name__))
    return out


@pytest.mark.parametrize("pin,precision", _cases())
def test_pin(oracle_lib, pin, precision):
    if precision is None:
        pin()
    else:
        pin(precision)
