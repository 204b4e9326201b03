"""Readback and render interop, and a headless dump for qualitative checks.

The reference's renderer binds ``posBuf``, ``matsBuf`` and ``velA`` (main.js:189-199)
and draws one sphere impostor per particle, coloured by its material record's
``color`` with a small per-particle albedo variation seeded from the fraction of
``pos.w`` (render.js:53-60); its validation is visual: pile angles, craters, spreading
(docs/IMPLEMENTATION.md:111-121).  Here the same buffers are

* exported zero-copy as device pointers (``Simulation.render_view()``, the C ABI's
  ``xpbi_get_render_view``) for native consumers on the sim's stream, or
* read back (``render_arrays``) and written as a binary PLY point cloud or a PNG
  splat image, with the pile-angle and crater measures those checks read off
  the particles.

Only numpy, zlib and struct are used: no image or mesh library.
"""
import math
import struct
import zlib

import numpy as np

# ---------------------------------------------------------------- buffers


def mat_of_tag(tag):
    """matOf (simkernels.js:41-44): the material index is the truncated tag."""
    return np.trunc(np.asarray(tag, np.float32)).astype(np.int64)


def render_arrays(sim, scene=None):
    """Host copies of what the renderer binds: pos (N, 4), velA (N, 4) and the
    material records (scenes, 8, 16) in the MaterialR layout (render.js:16-21:
    mu, lam, density, kind, p0..p3, color[4], boundFriction, mass, 2 pads).
    With ``scene`` set, only that scene's rows of pos and vel."""
    pos, vel, mats = sim.read("POS"), sim.read("VEL"), sim.read("MATS")
    if scene is not None:
        rows = sim.scenes_info[scene].rows
        pos, vel = pos[rows], vel[rows]
    return pos, vel, mats


def particle_colors(pos, mats, scene_of=None):
    """Per-particle RGB in [0, 1]: the material colour times the albedo variation
    0.88 + 0.24 (2.04 fract(tag)) of render.js:53-60."""
    tag = pos[:, 3].astype(np.float32)
    m = np.clip(mat_of_tag(tag), 0, 7)
    q = np.zeros(len(pos), np.int64) if scene_of is None else np.asarray(scene_of, np.int64)
    col = mats[q, m, 8:11].astype(np.float32)
    hv = (tag - np.floor(tag)) * np.float32(2.04)
    return np.clip(col * (np.float32(0.88) + np.float32(0.24) * hv)[:, None], 0.0, 1.0)


# ---------------------------------------------------------------- PLY

_PLY_PROPS = [("x", "<f4", "float"), ("y", "<f4", "float"), ("z", "<f4", "float"),
              ("vx", "<f4", "float"), ("vy", "<f4", "float"), ("vz", "<f4", "float"),
              ("material", "u1", "uchar"), ("red", "u1", "uchar"), ("green", "u1", "uchar"),
              ("blue", "u1", "uchar")]


def write_ply(path, pos, vel=None, colors=None):
    """Binary little-endian PLY of the particles: position, velocity, material
    index and 8-bit colour per vertex."""
    n = len(pos)
    rec = np.zeros(n, dtype=[(a, t) for a, t, _ in _PLY_PROPS])
    for k, a in enumerate("xyz"):
        rec[a] = pos[:, k]
        if vel is not None:
            rec["v" + a] = vel[:, k]
    rec["material"] = np.clip(mat_of_tag(pos[:, 3]), 0, 255)
    if colors is not None:
        c8 = np.clip(np.rint(np.asarray(colors) * 255.0), 0, 255).astype(np.uint8)
        rec["red"], rec["green"], rec["blue"] = c8[:, 0], c8[:, 1], c8[:, 2]
    head = ["ply", "format binary_little_endian 1.0", "comment xpbi particles", f"element vertex {n}"]
    head += [f"property {p} {a}" for a, _, p in _PLY_PROPS]
    head.append("end_header")
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode("ascii"))
        f.write(rec.tobytes())


def read_ply(path):
    """Parse a PLY written by write_ply (binary little endian, one vertex element)."""
    types = {"float": "<f4", "uchar": "u1", "int": "<i4", "uint": "<u4"}
    with open(path, "rb") as f:
        data = f.read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    lines = data[:end].decode("ascii").split("\n")
    if lines[0] != "ply" or lines[1] != "format binary_little_endian 1.0":
        raise ValueError("not a binary little-endian PLY")
    n, props = None, []
    for ln in lines[2:]:
        w = ln.split()
        if w[:2] == ["element", "vertex"]:
            n = int(w[2])
        elif w[:1] == ["property"]:
            props.append((w[2], types[w[1]]))
    return np.frombuffer(data, dtype=props, count=n, offset=end)


# ---------------------------------------------------------------- PNG


def _png_chunk(kind, body):
    return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)


def write_png(path, rgb):
    """8-bit RGB PNG (filter 0 on every row) of an (H, W, 3) uint8 array."""
    rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
    H, W, _ = rgb.shape
    raw = np.concatenate([np.zeros((H, 1), np.uint8), rgb.reshape(H, W * 3)], axis=1).tobytes()
    png = b"\x89PNG\r\n\x1a\n" + _png_chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, 8, 2, 0, 0, 0))
    png += _png_chunk(b"IDAT", zlib.compress(raw, 6)) + _png_chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(png)


def read_png(path):
    """Decode a PNG written by write_png (checks every chunk's CRC)."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG")
    at, idat, W = 8, b"", None
    while at < len(data):
        (n,) = struct.unpack(">I", data[at:at + 4])
        kind, body = data[at + 4:at + 8], data[at + 8:at + 8 + n]
        (crc,) = struct.unpack(">I", data[at + 8 + n:at + 12 + n])
        if crc != zlib.crc32(kind + body) & 0xFFFFFFFF:
            raise ValueError(f"bad CRC in {kind!r}")
        if kind == b"IHDR":
            W, H, depth, ctype = struct.unpack(">IIBB", body[:10])
            if (depth, ctype) != (8, 2):
                raise ValueError("only 8-bit RGB")
        elif kind == b"IDAT":
            idat += body
        at += 12 + n
    rows = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(H, 1 + 3 * W)
    if np.any(rows[:, 0] != 0):
        raise ValueError("only filter 0")
    return rows[:, 1:].reshape(H, W, 3)


# view: (horizontal axis, vertical axis, depth axis, depth sign: +1 = larger is nearer);
# front looks along -z, side along -x, top down along -y
VIEWS = {"front": (0, 1, 2, 1.0), "side": (2, 1, 0, 1.0), "top": (0, 2, 1, 1.0)}


def splat_image(pos, colors, lo, hi, view="front", width=640, radius_px=1, background=(18, 20, 26)):
    """Orthographic splat of the particles over the box [lo, hi]: each particle a
    disc of radius_px pixels, nearest particle per pixel (ties: lower index),
    shaded by the disc normal against a light along the view axis."""
    a, b, d, sgn = VIEWS[view]
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    ext_a, ext_b = hi[a] - lo[a], hi[b] - lo[b]
    W = int(width)
    H = max(1, int(round(W * ext_b / ext_a)))
    s = (W - 1) / ext_a
    u = np.rint((pos[:, a] - lo[a]) * s).astype(np.int64)
    v = np.rint((hi[b] - pos[:, b]) * s).astype(np.int64)  # image rows grow downwards
    depth = -sgn * pos[:, d].astype(np.float64)  # smaller is nearer
    img = np.empty((H, W, 3), np.float32)
    img[:] = np.asarray(background, np.float32) / 255.0
    pix_all, key_all, col_all = [], [], []
    idx = np.arange(len(pos))
    for dy in range(-radius_px, radius_px + 1):
        for dx in range(-radius_px, radius_px + 1):
            r2 = (dx * dx + dy * dy) / max(radius_px, 1) ** 2
            if r2 > 1.0:
                continue
            uu, vv = u + dx, v + dy
            ok = (uu >= 0) & (uu < W) & (vv >= 0) & (vv < H)
            shade = 0.35 + 0.65 * math.sqrt(1.0 - r2)
            pix_all.append(vv[ok] * W + uu[ok])
            # the disc's front surface: nearer at its centre
            key_all.append(depth[ok] - math.sqrt(1.0 - r2) * (radius_px / s))
            col_all.append(colors[idx[ok]] * shade)
    if pix_all:
        pix, key, col = np.concatenate(pix_all), np.concatenate(key_all), np.concatenate(col_all)
        order = np.lexsort((key, pix))
        first = np.unique(pix[order], return_index=True)[1]
        win = order[first]
        img.reshape(-1, 3)[pix[win]] = col[win]
    return np.clip(np.rint(img * 255.0), 0, 255).astype(np.uint8)


def dump(sim, prefix, scene=None, view="front", width=640):
    """Write ``prefix.ply`` and ``prefix.png`` of the sim's current state (one
    scene of a batch with ``scene``); returns the two paths."""
    pos, vel, mats = render_arrays(sim, scene)
    q = 0 if scene is None else scene
    if scene is None and sim.num_scenes > 1:
        scene_of = np.concatenate([np.full(si.N, k) for k, si in enumerate(sim.scenes_info)])
    else:
        scene_of = np.full(len(pos), q)
    col = particle_colors(pos, mats, scene_of)
    sc = sim.scenes_info[q].scene
    write_ply(prefix + ".ply", pos, vel, col)
    write_png(prefix + ".png", splat_image(pos, col, sc["domainMin"], sc["domainMax"], view, width))
    return prefix + ".ply", prefix + ".png"


# ---------------------------------------------------------------- measures


def surface_profile(pos, center=None, bins=40, r_max=None):
    """Top surface of a heap against the horizontal distance from its centre:
    per annulus of the (x, z) plane, the 95th percentile of y (y is up).
    Returns (annulus mid radii, heights); empty annuli are NaN."""
    p = np.asarray(pos, np.float64)
    c = p[:, [0, 2]].mean(axis=0) if center is None else np.asarray(center, np.float64)
    r = np.hypot(p[:, 0] - c[0], p[:, 2] - c[1])
    r_max = float(r.max()) if r_max is None else float(r_max)
    edges = np.linspace(0.0, r_max, bins + 1)
    k = np.clip(np.searchsorted(edges, r, side="right") - 1, 0, bins - 1)
    h = np.full(bins, np.nan)
    for j in range(bins):
        y = p[k == j, 1]
        if len(y):
            h[j] = np.percentile(y, 95)
    return 0.5 * (edges[1:] + edges[:-1]), h


def pile_angle(pos, floor=0.0, center=None, bins=40, lo=0.2, hi=0.8):
    """Angle of repose in degrees: the least-squares slope of the surface profile
    over the flank where the height above the floor lies between lo and hi of the
    heap's height (the qualitative check of docs/IMPLEMENTATION.md:113-117)."""
    r, h = surface_profile(pos, center, bins)
    h = h - floor
    ok = np.isfinite(h)
    top = np.nanmax(h)
    flank = ok & (h >= lo * top) & (h <= hi * top)
    if flank.sum() < 3:
        raise ValueError("heap profile has fewer than 3 flank points")
    slope = np.polyfit(r[flank], h[flank], 1)[0]
    return math.degrees(math.atan(-slope))


def crater_depth(pos, center, radius, ring=(2.0, 3.0)):
    """Depth of a crater at (x, z) = center: the median top surface (95th
    percentile of y per particle column) in the ring [ring0, ring1] x radius
    minus the lowest column top inside the radius."""
    p = np.asarray(pos, np.float64)
    d = np.hypot(p[:, 0] - center[0], p[:, 2] - center[1])
    cell = radius / 4.0
    ij = np.floor(p[:, [0, 2]] / cell).astype(np.int64)
    key = ij[:, 0] * 1_000_003 + ij[:, 1]
    order = np.argsort(key, kind="stable")
    uk, start = np.unique(key[order], return_index=True)
    tops, dist = [], []
    for a, b in zip(start, list(start[1:]) + [len(order)]):
        rows = order[a:b]
        tops.append(np.percentile(p[rows, 1], 95))
        dist.append(d[rows].mean())
    tops, dist = np.asarray(tops), np.asarray(dist)
    inner = tops[dist <= radius]
    outer = tops[(dist >= ring[0] * radius) & (dist <= ring[1] * radius)]
    if len(inner) == 0 or len(outer) == 0:
        raise ValueError("no particles inside the crater or its ring")
    return float(np.median(outer) - inner.min())
