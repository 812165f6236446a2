"""Batched and served dam-break scenes on one GPU.

Many independent scenes of uneven size advance together: each admitted scene
owns a context (its own HIP stream and device state, resident until the scene
finishes) and a host thread that drives its dual-time loop through the C ABI.
The C calls release the interpreter lock, so the scenes' kernels are in
flight on their streams at the same time and small scenes fill the GPU
together.  A request loop admits scenes at their arrival time when a slot and
enough of the HBM budget are free (continuous batching: requests that arrive
while others run are admitted as slots and memory free up), and records each
scene's time to solution.

A scene's result does not depend on what else runs: its state is private and
every kernel is deterministic, so the same request alone and in a mix gives
bit-identical fields.
"""
import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _capi as capi
from . import cases


@dataclass
class SceneSpec:
    """One dam-break request: tank and water block sizes, particle spacing,
    end time and arrival time (seconds after the stream starts)."""
    ident: int
    dim: int
    dp: float
    tank: tuple           # (DL, DH) or (DL, DW, DH)
    water: tuple          # (LL, LH) or (LL, LW, LH)
    t_end: float = 0.0    # physical end time (0: run `steps` advection steps)
    steps: int = 0
    arrival: float = 0.0
    precision: int = capi.MIXED
    geom: object = None   # the particle set, once generated

    def geometry(self):
        if self.geom is None:
            self.geom = self._generate()
        return self.geom

    def _generate(self):
        if self.dim == 2:
            (DL, DH), (LL, LH) = self.tank, self.water
            return cases.dambreak_2d(dp=self.dp, DL=DL, DH=DH, LL=LL, LH=LH)
        (DL, DW, DH), (LL, LW, LH) = self.tank, self.water
        return cases.dambreak_3d(dp=self.dp, DL=DL, DW=DW, DH=DH, LL=LL, LW=LW, LH=LH,
                                 probe=(DL, 0.5 * DW, 0.1 * DH))

    def fluid_count(self):
        return int(np.prod([round(w / self.dp) for w in self.water]))

    def footprint_bytes(self):
        """Device memory estimate: state plus neighbour slices of the fluid,
        state of the (three or four layers of) wall particles."""
        n = self.fluid_count()
        kmax = 96 if self.dim == 3 else 32
        per_fluid = 220 + kmax * 6
        surface = sum(w / self.dp for w in self.tank) * (self.dim - 1)
        n_wall = int(4 * surface ** (self.dim - 1)) if self.dim == 3 else int(8 * surface)
        return n * per_fluid + n_wall * 160 + (64 << 20)


def batched_2d_specs(count=256, seed=20261202, n_lo=1000, n_hi=64000, steps=40):
    """Independent 2D dam-break/sloshing tanks: water column width and height
    each U(0.5, 2.0), tank aspect ratio (length / height) U(1, 3), spacing
    chosen so the fluid count is log-uniform in [n_lo, n_hi]."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(count):
        LL, LH = rng.uniform(0.5, 2.0, size=2)
        aspect = rng.uniform(1.0, 3.0)
        n = math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))
        dp = math.sqrt(LL * LH / n)
        DH = 2.0 * LH
        DL = max(aspect * DH, LL + 4 * dp)
        specs.append(SceneSpec(k, 2, dp, (DL, DH), (LL, LH), steps=steps))
    return specs


def served_3d_specs(count=16, rate=0.5, seed=20261202, n_lo=20000, n_hi=2000000, t_end=1.0):
    """A Poisson stream (`rate` requests per second) of 3D dam breaks: tank
    U(1,4) x U(0.5,1.5) x U(1,2), water block over a fill fraction U(0.2,0.5)
    of the tank length, full width and half the height, spacing giving a
    log-uniform fluid count in [n_lo, n_hi], run to t_end."""
    rng = np.random.default_rng(seed)
    t = 0.0
    specs = []
    for k in range(count):
        DL, DW, DH = rng.uniform(1, 4), rng.uniform(0.5, 1.5), rng.uniform(1, 2)
        fill = rng.uniform(0.2, 0.5)
        LL, LW, LH = fill * DL, DW, 0.5 * DH
        n = math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))
        dp = (LL * LW * LH / n) ** (1.0 / 3.0)
        specs.append(SceneSpec(k, 3, dp, (DL, DW, DH), (LL, LW, LH), t_end=t_end, arrival=t))
        t += rng.exponential(1.0 / rate)
    return specs


@dataclass
class SceneResult:
    ident: int
    n_fluid: int
    acoustic_steps: int = 0
    advection_steps: int = 0
    physical_time: float = 0.0
    arrival: float = 0.0
    start: float = 0.0
    finish: float = 0.0
    fields: dict = field(default_factory=dict)

    @property
    def latency(self):
        return self.finish - self.arrival


def run_scene(lib, spec, result, keep_fields=False):
    """Build, run and release one scene on its own context; fills `result`."""
    geom = spec.geometry()
    sc = cases.DamBreak(lib, geom, spec.precision)
    try:
        sc.initialize()
        result.n_fluid = sc.fluid.n
        while (spec.steps and sc.iterations < spec.steps) or (spec.t_end and sc.time < spec.t_end):
            sc.advection_step()
        sc.ctx.synchronize()
        result.acoustic_steps = sc.acoustic_steps
        result.advection_steps = sc.iterations
        result.physical_time = sc.time
        if keep_fields:
            result.fields = {k: sc.fluid.read(k) for k in ("Position", "Velocity", "Compression")}
    finally:
        sc.close()


class Server:
    """Admits scene requests as slots and HBM budget free up and runs each on
    its own stream; `run` returns the results in request order."""

    def __init__(self, lib, max_active=16, hbm_budget=200 << 30):
        self.lib = lib
        self.max_active = max_active
        self.hbm_budget = hbm_budget

    def run(self, specs, keep_fields=False, realtime=True):
        results = [SceneResult(s.ident, 0, arrival=s.arrival) for s in specs]
        order = sorted(range(len(specs)), key=lambda k: specs[k].arrival)
        cond = threading.Condition()
        state = {"active": 0, "bytes": 0, "error": None}
        threads = []
        t0 = time.perf_counter()

        def worker(k):
            spec, res = specs[k], results[k]
            try:
                res.start = time.perf_counter() - t0
                run_scene(self.lib, spec, res, keep_fields)
            except Exception as e:  # reported after the stream drains
                with cond:
                    state["error"] = state["error"] or e
            finally:
                res.finish = time.perf_counter() - t0
                with cond:
                    state["active"] -= 1
                    state["bytes"] -= spec.footprint_bytes()
                    cond.notify_all()

        for k in order:
            spec = specs[k]
            need = spec.footprint_bytes()
            if need > self.hbm_budget:
                raise ValueError(f"scene {spec.ident} needs {need} B, more than the HBM budget")
            if realtime:
                wait = spec.arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
            else:
                results[k].arrival = time.perf_counter() - t0
            with cond:
                while state["active"] >= self.max_active or state["bytes"] + need > self.hbm_budget:
                    cond.wait()
                state["active"] += 1
                state["bytes"] += need
            th = threading.Thread(target=worker, args=(k,), daemon=True)
            th.start()
            threads.append(th)
        for th in threads:
            th.join()
        if state["error"] is not None:
            raise state["error"]
        return results, time.perf_counter() - t0


def summarize(results, wall):
    lat = np.array([r.latency for r in results])
    work = sum(r.n_fluid * r.acoustic_steps for r in results)
    return {"scenes": len(results), "wall_s": wall, "particle_steps": work,
            "particle_steps_per_s": work / wall,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_max_s": float(lat.max())}
