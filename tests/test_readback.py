"""Readback formats and the qualitative measures (CPU): PLY and PNG round trips,
the splat image's nearest-particle rule, and the pile-angle and crater measures
on heaps whose answer the geometry fixes."""
import math

import numpy as np
import pytest

from xpbi import readback


def _cloud(n, seed):
    g = np.random.default_rng(seed)
    pos = np.empty((n, 4), np.float32)
    pos[:, :3] = g.random((n, 3), dtype=np.float32)
    pos[:, 3] = g.integers(0, 8, n) + g.random(n, dtype=np.float32) * 0.49
    vel = g.standard_normal((n, 4)).astype(np.float32)
    return pos, vel


def test_ply_round_trip(tmp_path):
    pos, vel = _cloud(1000, 1)
    col = np.random.default_rng(2).random((1000, 3))
    path = str(tmp_path / "p.ply")
    readback.write_ply(path, pos, vel, col)
    r = readback.read_ply(path)
    assert len(r) == 1000
    for k, a in enumerate("xyz"):
        assert np.array_equal(r[a], pos[:, k])
        assert np.array_equal(r["v" + a], vel[:, k])
    assert np.array_equal(r["material"], np.trunc(pos[:, 3]).astype(np.uint8))
    assert np.array_equal(r["red"], np.rint(col[:, 0] * 255).astype(np.uint8))


def test_png_round_trip(tmp_path):
    img = np.random.default_rng(3).integers(0, 256, (37, 53, 3), dtype=np.uint8)
    path = str(tmp_path / "i.png")
    readback.write_png(path, img)
    assert np.array_equal(readback.read_png(path), img)
    data = bytearray(open(path, "rb").read())
    data[40] ^= 1  # inside IDAT: its CRC no longer matches
    open(path, "wb").write(bytes(data))
    with pytest.raises(ValueError):
        readback.read_png(path)


def test_colors_follow_render_variation():
    mats = np.zeros((1, 8, 16), np.float32)
    mats[0, 2, 8:11] = [0.5, 0.25, 1.0]
    pos = np.array([[0, 0, 0, 2.0], [0, 0, 0, 2.25]], np.float32)
    c = readback.particle_colors(pos, mats)
    assert np.allclose(c[0], [0.44, 0.22, 0.88])
    f = 0.88 + 0.24 * 0.25 * 2.04
    assert np.allclose(c[1], np.clip(np.array([0.5, 0.25, 1.0]) * f, 0, 1))


def test_splat_nearest_wins():
    pos = np.array([[0.5, 0.5, 0.2, 0], [0.5, 0.5, 0.8, 0], [0.1, 0.1, 0.5, 0]], np.float32)
    col = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    img = readback.splat_image(pos, col, [0, 0, 0], [1, 1, 1], view="front", width=101, radius_px=2)
    assert img.shape == (101, 101, 3)
    assert tuple(img[50, 50]) == (0, 255, 0)  # larger z is nearer in the front view
    img = readback.splat_image(pos, col, [0, 0, 0], [1, 1, 1], view="side", width=101, radius_px=2)
    # the side view's horizontal axis is z
    assert tuple(img[50, 20]) == (255, 0, 0)
    assert tuple(img[50, 80]) == (0, 255, 0)
    assert tuple(img[90, 50]) == (0, 0, 255)
    # a tie in depth: the lower index wins
    tie = readback.splat_image(pos[[1, 1]], col[[2, 0]], [0, 0, 0], [1, 1, 1], width=101, radius_px=1)
    assert tuple(tie[50, 50]) == (0, 0, 255)


@pytest.mark.parametrize("deg", [25.0, 35.0])
def test_pile_angle_of_a_cone(deg):
    g = np.random.default_rng(4)
    R = 0.3
    H = R * math.tan(math.radians(deg))
    n = 400_000
    p = g.random((n, 3)) * [2 * R, H, 2 * R] - [R, 0, R]
    r = np.hypot(p[:, 0], p[:, 2])
    p = p[(r < R) & (p[:, 1] < H * (1 - r / R))] + [0.7, 0.0, 0.7]
    got = readback.pile_angle(p, floor=0.0, bins=30)
    assert abs(got - deg) < 1.5, got


def test_crater_depth_of_a_dented_slab():
    s = 0.005
    ax = np.arange(0, 1, s) + s / 2
    x, y, z = np.meshgrid(ax, np.arange(0, 0.1, s) + s / 2, ax, indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    R, d = 0.06, 0.03
    c = np.array([0.5, 0.1 + R - d, 0.5])
    p = p[np.linalg.norm(p - c, axis=1) >= R]
    a = math.sqrt(R * R - (R - d) ** 2)
    got = readback.crater_depth(p, (0.5, 0.5), a)
    assert abs(got - d) < 1.5 * s, got
