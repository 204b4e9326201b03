"""Headless frame driver: the step semantics of the reference's frame loop
(web/js/main.js:383-415) without the UI, as js/frame.js has them.

One frame writes Params once -- dt = (1/60) timeScale / substeps unless the scene
fixes dt, the collider velocity from the finite difference of its position over
the frame's wall time, clamped to +-1.5 m/s -- and runs the frame's substeps in
the reference's chunks of 6 command encoders.  Here each chunk is a replay of the
simulation's captured HIP graph (xpbi_graph_step), whose results equal eager
substeps.
"""
from .sim import params_from_scene

CHUNK = 6


class FrameDriver:
    def __init__(self, sim, cfg, graph=True):
        self.sim, self.cfg, self.graph = sim, cfg, graph
        self.sim_time = 0.0
        self.collider_pos = [0.5, 0.3, 0.5]
        self.collider_prev = [0.5, 0.3, 0.5]
        self.collider_r = 0.08
        self.poke_active = False
        col = cfg.get("collider")
        if col:
            self.collider_pos = list(col["pos"])
            self.collider_prev = list(col["pos"])
            self.collider_r = col["R"]
            self.poke_active = True

    def frame_dt(self):
        """Simulated seconds per frame (main.js:386-390)."""
        cfg = self.cfg
        return (cfg.get("dt") or (1.0 / 60.0) * cfg["timeScale"] / cfg["substeps"]) * cfg["substeps"]

    def params(self, dt_wall):
        cfg = self.cfg
        cv = [(self.collider_pos[k] - self.collider_prev[k]) / max(dt_wall, 1e-3) for k in range(3)]
        self.collider_prev = list(self.collider_pos)
        p = params_from_scene(cfg)
        p.update(colliderPos=list(self.collider_pos), colliderR=self.collider_r,
                 colliderVel=[max(-1.5, min(1.5, v)) for v in cv], colliderActive=self.poke_active,
                 time=self.sim_time)
        return p

    def frame(self, dt_wall=1.0 / 60.0):
        """One frame; dt_wall = wall seconds since the previous frame (clamped to 1/30, main.js:369-371)."""
        cfg = self.cfg
        dt_wall = min(dt_wall, 1.0 / 30.0)
        self.sim.writeParams(self.params(dt_wall))
        I, water, n = cfg["iterations"], cfg["hasWater"], cfg["substeps"]
        for s in range(0, n, CHUNK):
            k = min(CHUNK, n - s)
            if self.graph:
                self.sim.graph_step(k, I, water)
            else:
                self.sim.step(k, I, water)
        self.sim_time += (1.0 / 60.0) * cfg["timeScale"]

    def move_collider(self, delta):
        self.collider_pos = [self.collider_pos[k] + delta[k] for k in range(3)]
