"""The C ABI: the built libraries load here (no GPU needed) and export every function the
headers under include/ declare; a call without a GPU fails with a message, not a crash."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

from autan_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared(header: str) -> list[str]:
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(mx\w+)\s*\(", text, flags=re.M)))


def _mlp():
    from autan_amd import mlp

    return mlp.lib()


@pytest.mark.parametrize("header,loader", [("mxmesh.h", _lib.mesh), ("mxeuler.h", _lib.euler), ("mxmlp.h", _mlp)])
def test_every_declared_symbol_is_exported(header, loader):
    names = declared(header)
    assert len(names) >= 8
    lib = loader()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string():
    lib = _lib.euler()
    lib.mxeuler_version.restype = ctypes.c_char_p
    assert lib.mxeuler_version().decode().startswith("mxeuler")


def test_create_validates_before_touching_a_device():
    """Bad input is refused on the host with the reference's exception kinds' messages, before
    any allocation (physics/src/device_stepper.cuh:184-193)."""
    import numpy as np

    import cases
    from autan_amd.stepper import DeviceStepper

    case = cases.fixture_case("hex", "roe")
    case.scheme.planar = True
    poly = cases.fixture_case("poly", "roe")
    poly.scheme.planar = True
    with pytest.raises(RuntimeError, match="planar"):
        DeviceStepper(poly.mesh, poly.scheme, "fp64", poly.bc_kind, poly.bc)
    with pytest.raises(ValueError):
        DeviceStepper(case.mesh, case.scheme, "fp64", case.bc_kind[:-1], case.bc)
    bad = cases.fixture_case("hex", "roe")
    bad.mesh.arrays["owner"] = bad.mesh.owner.copy()
    bad.mesh.arrays["owner"][3] = 10**6
    with pytest.raises(RuntimeError, match="owner out of range"):
        DeviceStepper(bad.mesh, bad.scheme, "fp64", bad.bc_kind, bad.bc)
    assert np.isfinite(case.u).all()


def test_the_status_tells_refused_input_from_a_device_failure():
    """mxeuler_create returns 1 for input it refuses (the reference's std::invalid_argument) and 2
    for a device failure (std::runtime_error, E1101), here a host without a GPU."""
    import os

    import cases
    from autan_amd import stepper as st

    lib = _lib.euler()
    bad = cases.fixture_case("hex", "roe")
    bad.mesh.arrays["owner"] = bad.mesh.owner.copy()
    bad.mesh.arrays["owner"][3] = 10**6
    calls = []
    real = lib.mxeuler_create

    def spy(*args):
        rc = real(*args)
        calls.append(rc)
        return rc

    lib.mxeuler_create = spy
    try:
        with pytest.raises(RuntimeError, match="owner out of range"):
            st.DeviceStepper(bad.mesh, bad.scheme, "fp64", bad.bc_kind, bad.bc)
        if not os.path.exists("/dev/kfd"):
            good = cases.tets_case(3)
            with pytest.raises(RuntimeError, match="E1101"):
                st.DeviceStepper(good.mesh, good.scheme, "fp64", good.bc_kind, good.bc, good.source)
    finally:
        lib.mxeuler_create = real
    assert calls[0] == 1
    if len(calls) > 1:
        assert calls[1] == 2
