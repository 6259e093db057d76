"""The product's scene builder (host code in libvxhip) against the records the
reference World built from the same scene descriptions (golden fixtures).

Spheres, boxes, capsules, planes and heightfield meshes use the reference's
closed forms and layout, so their records match bit for bit.  Hull mass
properties come from an independent integration and eigen-solve, so they are
compared physically: centre, volume-derived mass, and the world-frame inverse
inertia tensor R diag(I^-1) R^T (invariant to axis order and sign).
"""
import os

import numpy as np
import pytest

from velox_amd import device
from vxfix import as_bodies, load_case

pytestmark = pytest.mark.skipif(not os.path.exists(device.LIBRARY), reason="libvxhip.so not built")

def raw(a):
    return np.ascontiguousarray(a).view(np.uint8)


DATA_FIELDS = [n for n in as_bodies(np.zeros((1, 320), np.uint8)).dtype.names if not n.startswith("_pad")]


def qrot(q):
    x, y, z, w = [float(v) for v in q]
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def world_inv_inertia(b):
    R = qrot(b["orientation"]) @ qrot(b["inertiaOrientation"])
    return R @ np.diag(b["invInertia"].astype(np.float64)) @ R.T


def test_lattice_records_match_reference_bitwise():
    got = device.Scene("lattice", 4).bodies
    want = as_bodies(load_case("lattice64")["bodies_in"])
    assert len(got) == len(want)
    for f in DATA_FIELDS:
        assert np.array_equal(raw(got[f]), raw(want[f])), f


@pytest.mark.parametrize("name,case", [("fixture-pile", "pile"), ("fixture-terrain", "terrain")])
def test_mixed_records_match_reference(name, case):
    sc = device.Scene(name)
    ref = load_case(case)
    want = as_bodies(ref["bodies_in"])
    got = sc.bodies
    assert len(got) == len(want)
    hull = want["shape"] == 5
    exact = [f for f in DATA_FIELDS if f not in ("position", "invInertia", "inertiaOrientation", "radius",
                                                    "hullFaceCount", "hullFaceFirst")]
    for f in exact:
        assert np.array_equal(raw(got[f]), raw(want[f])), f
    for f in ("position", "invInertia", "inertiaOrientation", "radius"):
        assert np.array_equal(raw(got[f][~hull]), raw(want[f][~hull])), f
    np.testing.assert_allclose(got["position"][hull], want["position"][hull], rtol=0, atol=2e-6)
    np.testing.assert_allclose(got["radius"][hull], want["radius"][hull], rtol=2e-6)
    for g, w in zip(got[hull], want[hull]):
        np.testing.assert_allclose(world_inv_inertia(g), world_inv_inertia(w), rtol=0,
                                   atol=2e-5 * np.abs(world_inv_inertia(w)).max())
    # Hull geometry: same recentred points and the same facet set.
    np.testing.assert_allclose(sc.soup.hullpts, ref["soup"].hullpts, rtol=0, atol=2e-6)
    for g, w in zip(got[hull], want[hull]):
        def tris(faces, b):
            t = faces[b["hullFaceFirst"] * 3:(b["hullFaceFirst"] + b["hullFaceCount"]) * 3].reshape(-1, 3)
            return sorted(tuple(np.roll(x, -int(np.argmin(x)))) for x in t)
        assert tris(sc.soup.hullfaces, g) == tris(ref["soup"].hullfaces, w)
    # Heightfield meshes: vertices, indices and BVH bit for bit.
    for slot in ("vertices", "indices", "meshes", "nodes", "trirefs"):
        assert np.array_equal(getattr(sc.soup, slot), getattr(ref["soup"], slot)), slot
