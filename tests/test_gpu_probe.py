"""The byte-gather probe that bounds the walk from above: it runs, returns a
rate, and one row per wave-load is faster than 64 lines per wave-load."""
import ctypes as ct

import pytest

pytestmark = pytest.mark.gpu


def test_gather_probe_rates(gpu):
    from anacor_public_amd import _lib
    lib = _lib.load()
    rates = {}
    for name, pattern in (("lanes", _lib.GATHER_LANES), ("row", _lib.GATHER_ROW)):
        print(f"case: gather probe, 16 MiB working set, pattern {name}", flush=True)
        r = ct.c_double()
        _lib.check(lib.anacor_probe_gather(0, 24, pattern, 1 << 30, ct.byref(r)), lib)
        rates[name] = r.value
    assert rates["lanes"] > 1e9
    assert rates["row"] > rates["lanes"]
