"""The native host case builder (Metric, BaseState, initial conditions,
Davies weights and targets) reproduces, bit for bit, the arrays recorded in
tests/golden/host_builder_digests.json (SHA-256 of each FP64 array) for the
idealized case files, the all-switches test grid and three regional cases."""
import json

import pytest

from golden import make_host_digests as M

WANT = json.loads(M.OUT.read_text())


@pytest.fixture(scope="module")
def got():
    return M.compute()


@pytest.mark.parametrize("name", sorted(WANT))
def test_host_builder_arrays_are_bit_identical(name, got):
    assert name in got
    diff = sorted(k for k in WANT[name] if WANT[name][k] != got[name].get(k))
    assert not diff, f"{name}: arrays differ: {diff}"
    assert set(got[name]) == set(WANT[name])
