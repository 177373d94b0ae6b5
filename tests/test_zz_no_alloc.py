"""mxeuler_face_force_noalloc (the face force staged in the library's static device array, no
device allocation) against mxeuler_face_force, through the C ABI: the same bits on two calls in a
row, the second with fewer faces and another state, then on a second, smaller stepper on the same
GPU (the array is shared); more faces than the array holds are refused before any device work."""

from __future__ import annotations

import numpy as np
import pytest

import cases
from autan_amd.stepper import STATIC_FORCE_FACES, DeviceStepper

pytestmark = pytest.mark.gpu


def _header_limit() -> int:
    from pathlib import Path

    text = (Path(__file__).resolve().parent.parent / "include" / "mxeuler.h").read_text()
    for line in text.splitlines():
        if line.startswith("#define MXEULER_STATIC_FORCE_FACES"):
            return int(line.split()[2])
    raise AssertionError("MXEULER_STATIC_FORCE_FACES is not defined")


def test_the_binding_knows_the_headers_limit():
    assert STATIC_FORCE_FACES == _header_limit()


@pytest.mark.parametrize("policy", ["fp32+acc64", "fp64"])
def test_noalloc_face_force_gives_the_allocating_calls_bits(policy):
    big = cases.bump_case(16, 12, 8, hybrid=True)
    print(f"no-alloc face force: bump 16x12x8 {policy}", flush=True)
    dev = DeviceStepper(big.mesh, big.scheme, policy, big.bc_kind, big.bc, big.source)
    try:
        dev.upload(big.u)
        faces = np.arange(big.mesh.n_interior, big.mesh.n_faces, dtype=np.int32)
        assert faces.size > 100
        first = dev.face_force_noalloc(faces)
        np.testing.assert_array_equal(first, dev.face_force(faces))
        assert np.all(np.isfinite(first)) and np.any(first != 0.0)
        # Second call: another state and fewer faces, in another order; what the first call left in
        # the array must not show through.
        dev.step(cfl=0.5)
        fewer = faces[::-3].copy()
        second = dev.face_force_noalloc(fewer)
        assert second.shape == (3, fewer.size)
        np.testing.assert_array_equal(second, dev.face_force(fewer))
        assert not np.array_equal(second, first[:, ::-3])
        # No faces: nothing to do.
        assert dev.face_force_noalloc(np.empty(0, dtype=np.int32)).shape == (3, 0)
    finally:
        dev.close()
    small = cases.tets_case(4)
    print(f"no-alloc face force: tets 4 {policy}", flush=True)
    dev = DeviceStepper(small.mesh, small.scheme, policy, small.bc_kind, small.bc, small.source)
    try:
        dev.upload(small.u)
        faces = np.arange(small.mesh.n_interior, small.mesh.n_faces, dtype=np.int32)[:37]
        np.testing.assert_array_equal(dev.face_force_noalloc(faces), dev.face_force(faces))
    finally:
        dev.close()


def test_noalloc_face_force_refuses_more_faces_than_its_array_holds():
    c = cases.tets_case(4)
    print("no-alloc face force: refusals", flush=True)
    dev = DeviceStepper(c.mesh, c.scheme, "fp32+acc64", c.bc_kind, c.bc, c.source)
    try:
        dev.upload(c.u)
        b = np.arange(c.mesh.n_interior, c.mesh.n_faces, dtype=np.int32)
        too_many = np.resize(b, STATIC_FORCE_FACES + 1)
        with pytest.raises(RuntimeError, match="static staging holds"):
            dev.face_force_noalloc(too_many)
        with pytest.raises(RuntimeError, match="not a boundary face"):
            dev.face_force_noalloc(np.array([0], dtype=np.int32))
        # The refusals left the stepper usable.
        np.testing.assert_array_equal(dev.face_force_noalloc(b[:5]), dev.face_force(b[:5]))
    finally:
        dev.close()
