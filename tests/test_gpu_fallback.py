"""The sweeps' fallback to cell-row walks.

The support-neighbour list holds up to 256 16-bit offsets per particle (the
default of xpbi_set_neighbour_slots).  A particle with more neighbours, or whose
candidates lie more than 65,535 particles past its first one, is flagged and its
sweeps walk the 9 cell rows instead (with the reference's 384-candidate cap per
row, simkernels.js:73).  The list and the walk visit the same neighbours in the
same order, so any capacity must give the same state bit for bit; and the scenes
below, which force each cause, hold one full substep to the bar of the other
parity tests against the oracle.  Each of them first checks, on the host, that
its scene really takes the path it is meant to exercise.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from xpbi import scenes
from helpers import F3
from test_gpu_parity import make_sim, one_substep_parity

pytestmark = pytest.mark.gpu

LIST_SLOTS = 256
OFFSET_MAX = 0xFFFF
ROW_CAP = 384


def _grid(sim, pos):
    """Cell coordinates (x, y, z) of each particle and the grid dimensions."""
    origin, cell, dims = sim.gridOrigin, sim.cellSize, sim.gridDims
    c = np.floor((pos[:, :3].astype(np.float32) - np.asarray(origin, np.float32)) / np.float32(cell)).astype(np.int64)
    return np.clip(c, 0, np.asarray(dims) - 1), np.asarray(dims, np.int64)


def test_sand_in_dense_water_walks_rows(oracle):
    """A 14x14x14 blob of water at spacing 0.35 r sunk into the top of a DP column.
    Solids never come closer than the 0.75 h contact distance (simkernels.js:333-350),
    but water does, and a sand particle's list holds its water neighbours too
    (simkernels.js:199-210): the sand inside the blob has several hundred of them, more
    than the list holds, and cell rows of the blob pass the 384 cap.  The blob's PBF
    is made inert (CFM epsilon 1e12, no s_corr) so that it stays put: at 23 times its
    rest density it would otherwise burst within a substep, and a state that violent
    pins little."""
    base = scenes.sand_column(16, 16, 16, 0.01, [0.5, 0.5, 0.5], seed=11, iterations=10)
    g = np.random.default_rng(4)
    s = 0.35 * 0.01
    ax = (np.arange(14) - 6.5) * s
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    blob = np.stack([X.ravel() + 0.25, Y.ravel() + 0.15, Z.ravel() + 0.25], 1)
    blob += g.uniform(-0.05 * s, 0.05 * s, blob.shape)
    n0, nb = base["positions"].shape[0], blob.shape[0]
    scene = dict(base, positions=np.concatenate([base["positions"], blob.astype(np.float32)]),
                 matIds=np.concatenate([base["matIds"], np.ones(nb, np.uint32)]),
                 tags=np.concatenate([base["tags"], np.full(nb, 1.25, np.float32)]),
                 states=np.concatenate([base["states"], np.zeros(nb, np.float32)]),
                 materials=[scenes.DP_SAND_STIFF(30.0), scenes.WATER(1000.0)], hasWater=True, pbfEps=1e12,
                 sCorrK=0.0)
    # a shear of the sand (v_x = 2 (y - 0.08) m/s) deforms it within two substeps,
    # before the blob spreads
    vel = np.zeros((n0 + nb, 3), np.float32)
    vel[:n0, 0] = 2.0 * (base["positions"][:, 1] - 0.08)
    scene["velocities"] = vel
    sim = make_sim(scene)
    sim.step(2, 10, True)
    pos = sim.read("POS")
    sand_now = np.floor(pos[:, 3]).astype(int) == 0
    dev = np.abs(F3(sim.read("F"))[sand_now] - np.array([1, 0, 0, 0, 1, 0, 0, 0, 1], np.float32)).max()
    print(f"sand |F - I| max {dev:.2e}", flush=True)
    h = scene["h"]
    pts = pos[:, :3].astype(np.float64)
    cnt = np.array([len(l) - 1 for l in cKDTree(pts).query_ball_point(pts, 2 * h * (1 - 1e-6))])
    sand = np.floor(pos[:, 3]).astype(int) == 0
    c, dims = _grid(sim, pos)
    cs = np.append(sim.read("CELLSTART").astype(np.int64), sim.N)  # cellStart[C] = N
    row = c[:, 1] * dims[0] + c[:, 2] * dims[0] * dims[1]
    span = cs[row + np.minimum(c[:, 0] + 1, dims[0] - 1) + 1] - cs[row + np.maximum(c[:, 0] - 1, 0)]
    print(f"sand in dense water: {sim.N} particles, neighbours max {cnt.max()}, sand past the list "
          f"{(cnt[sand] > LIST_SLOTS).sum()}, own row span max {span.max()}", flush=True)
    assert (cnt[sand] > LIST_SLOTS).sum() > 20, "no sand particle overflows the list"
    assert span.max() > ROW_CAP, "no cell row reaches the 384 cap"
    one_substep_parity(sim, scene, oracle, "sand in dense water", iters=10, water=True)
    sim.close()


def test_wide_wall_offsets_past_16_bits(oracle):
    """A 160 x 110 x 6 DP wall (105,600 particles at r = 0.01): a z-layer of cells holds
    35,200 particles, so a middle-layer particle's candidates span more than 65,535
    particles and its offsets do not fit 16 bits."""
    scene = scenes.sand_column(160, 110, 6, 0.01, [2.0, 1.2, 0.2], seed=12, iterations=10)
    sim = make_sim(scene)
    sim.step(200, 10, False)  # the wall settles on the floor and deforms
    pos = sim.read("POS")
    c, dims = _grid(sim, pos)
    cs = np.append(sim.read("CELLSTART").astype(np.int64), sim.N)  # cellStart[C] = N
    gx, gxy = dims[0], dims[0] * dims[1]
    first = cs[np.maximum(c[:, 0] - 1, 0) + np.maximum(c[:, 1] - 1, 0) * gx + np.maximum(c[:, 2] - 1, 0) * gxy]
    last = cs[np.minimum(c[:, 0] + 1, dims[0] - 1) + np.minimum(c[:, 1] + 1, dims[1] - 1) * gx +
              np.minimum(c[:, 2] + 1, dims[2] - 1) * gxy + 1] - 1
    span = last - first
    print(f"wide wall: {sim.N} particles, candidate span max {span.max()}, past 16 bits {(span > OFFSET_MAX).sum()}",
          flush=True)
    assert (span > OFFSET_MAX).sum() > 1000, "no particle's offsets overflow 16 bits"
    one_substep_parity(sim, scene, oracle, "wide DP wall", iters=10)
    sim.close()


def test_list_capacity_changes_nothing_dam_breach_1m():
    """The 1M dam breach (600k PBF water, 400k DP sand) at full size: 2 substeps from
    the state after 60, once through the list (256 slots) and once walking every row
    (0 slots), give the same state bit for bit.  A whole-state property: this scene
    drives its water to the velocity clamp within a substep (the CPU oracle does the
    same), so a 1-substep tolerance against the oracle would pin little here."""
    scene = scenes.dam_breach_1m()
    sim = make_sim(scene)
    print("dam breach 1M: settling 60 substeps", flush=True)
    sim.step(60, scene["iterations"], True)
    snap = {k: sim.read(k) for k in ("POS", "VEL", "F", "AUX", "ID")}
    out = {}
    for slots in (256, 0):
        for k, v in snap.items():
            sim.write(k, v)
        sim.set_neighbour_slots(slots)
        sim.step(2, scene["iterations"], True)
        out[slots] = {k: sim.read(k) for k in ("POS", "VEL", "F", "AUX")}
        print(f"dam breach 1M: {slots} slots done", flush=True)
    assert np.abs(out[256]["POS"] - snap["POS"]).max() > 0, "the state did not move"
    for k in out[256]:
        assert np.array_equal(out[0][k], out[256][k]), f"dam breach 1M: {k} differs with 0 slots"
    sim.close()


@pytest.mark.parametrize("name", ["sand column", "Dam breach: water + sand"])
def test_list_capacity_changes_nothing(name):
    """8 substeps with the default 256 slots, with 16 (many particles past the list,
    waves mixing list and walk) and with 0 (every sweep walks its rows): identical
    positions, velocities, F and aux, bit for bit."""
    if name == "sand column":
        scene = scenes.sand_column(16, 32, 16, 0.01, [0.6, 0.5, 0.6], seed=5, iterations=10)
        settle, water = 150, False
    else:
        scene = scenes.SCENES[name](seed=1)
        settle, water = 20 * scene["substeps"], True
    sim = make_sim(scene)
    sim.step(settle, scene["iterations"], water)
    snap = {k: sim.read(k) for k in ("POS", "VEL", "F", "AUX", "ID")}
    out = {}
    for slots in (256, 16, 0):
        for k, v in snap.items():
            sim.write(k, v)
        sim.set_neighbour_slots(slots)
        sim.step(8, scene["iterations"], water)
        out[slots] = {k: sim.read(k) for k in ("POS", "VEL", "F", "AUX")}
        print(f"{name}: {slots} slots done", flush=True)
    assert np.abs(out[256]["POS"] - snap["POS"]).max() > 0, "the state did not move"
    for slots in (16, 0):
        for k in out[256]:
            assert np.array_equal(out[slots][k], out[256][k]), f"{name}: {k} differs with {slots} slots"
    sim.close()
