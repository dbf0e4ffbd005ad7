"""The C-ABI library loads without a GPU and exports every symbol the public
header declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

from dpnegf_amd import _native as nat

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "dpnegf_amd.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(dpnegf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_bound_symbols():
    assert set(_declared()) == set(nat.EXPORTED)


@pytest.mark.skipif(not os.path.exists(nat.lib_path()), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(nat.lib_path())
    for name in _declared():
        assert hasattr(lib, name), name
    nat.load()
    assert nat.load().dpnegf_abi_version() == 1


@pytest.mark.skipif(not os.path.exists(nat.lib_path()), reason="library not built")
def test_argument_errors_are_reported_without_a_device():
    # invalid sizes are rejected on the host before anything touches the GPU
    lib = nat.load()
    bsz = nat.host_i32([80])
    rc = lib.dpnegf_rgf_transmission(None, 0, 1, bsz, None, None, 0, 1, None, 1, None, None, 1,
                                     None, 1, 1, None, None)
    assert rc == -2 and b"exceeds" in lib.dpnegf_last_error()
    rc = lib.dpnegf_lead_self_energy(None, 0, 4, 0, None, 0, 1, None, 1, None, 1e-5, 1e-40,
                                     None, None, None, 1, None, None, None)
    assert rc == -1
    with pytest.raises(ValueError):
        nat.call("dpnegf_rgf_full", None, 0, None, None, 1, None, None, 0, None, 0, 1,
                 *([None] * 17))
