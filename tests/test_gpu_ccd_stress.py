"""CCD stress (physical checks): 4,096 spheres at 2 km/s into a 100 x 100 wall of
stacked unit boxes, 256 boxes spinning at 1.5 km/s tip speed above a 256 x 256
heightfield.  The conservative-advancement net must keep every bullet on the
near side of the wall while it stands (steps 1-5), and after 10 steps no
state may be non-finite and no spinning box may be under the terrain.  (The
full 120 steps take minutes: the struck wall becomes one island of about
165,000 contacts.)  Needs an MI355X."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WALL_BACK = 21.0      # the wall occupies x in [20, 21]
BULLET_R = 0.05


def _terrain_floor(scene, x, z):
    """Lowest corner height of the heightfield cell under (x, z); the
    triangulated surface of the cell never dips below it."""
    verts = scene.soup.vertices  # the heightfield's vertices, row by row (z, then x)
    side = 256
    cell = 0.5
    ox = -0.5 * cell * (side - 1)
    heights = verts[:, 1].reshape(side, side)  # row z, column x
    ix = np.clip(((x - ox) / cell).astype(int), 0, side - 2)
    iz = np.clip(((z - ox) / cell).astype(int), 0, side - 2)
    return np.minimum(np.minimum(heights[iz, ix], heights[iz, ix + 1]),
                      np.minimum(heights[iz + 1, ix], heights[iz + 1, ix + 1]))


def test_ccd_stress_no_tunnelling():
    from velox_amd.device import Context, Scene
    sc = Scene("ccdstress")
    bullets = sc.bodies["shape"] == 0
    spinners = (sc.bodies["shape"] == 2) & (sc.bodies["position"][:, 0] < 0.0)
    assert bullets.sum() == 4096 and spinners.sum() == 256
    c = Context()
    try:
        c.load_scene(sc)
        for step in range(1, 6):  # the wall is still standing
            c.step(sc.dt, 1)
            b = c.bodies()
            x = b["position"][bullets, 0]
            print(f"step {step}: bullet x max {x.max():.3f}, stats {c.stats()['ccdRescued']} rescues", flush=True)
            assert np.isfinite(x).all()
            assert (x <= WALL_BACK + BULLET_R).all(), f"step {step}: {(x > WALL_BACK + BULLET_R).sum()} bullets passed the wall"
        for step in range(6, 11):  # about 1.5 s of device time per step once the wall is hit
            c.step(sc.dt, 1)
            print(f"step {step}: {c.stats()['contacts']} contacts", flush=True)
        b = c.bodies()
        state = np.concatenate([b["position"], b["velocity"], b["angularVelocity"]], axis=1)
        assert np.isfinite(state).all()
        p = b["position"][spinners]
        inside = (np.abs(p[:, 0]) < 63.0) & (np.abs(p[:, 2]) < 63.0)
        floor = _terrain_floor(sc, p[inside, 0], p[inside, 2])
        below = p[inside, 1] < floor
        assert not below.any(), f"{below.sum()} spinning boxes under the terrain"
    finally:
        c.close()
