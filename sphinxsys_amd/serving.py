"""Batched and served dam-break scenes on one GPU.

Many independent scenes of uneven size advance together.  `BatchServer`
runs them on the batch path (sphx_batch_*): every stage of an advection step
is one launch over all admitted scenes, the step is one graph launch, and a
request loop admits arriving scenes when a slot and enough of the HBM budget
are free (continuous batching: requests that arrive while others run join at
the next step; finished scenes leave) and records each scene's time to
solution.  `Server` is the stream-per-scene form: each admitted scene owns a
context and a host thread that drives its own loop through the C ABI; with
one slot it is the serial baseline (one plain run after another).

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
    its own stream; `run` returns the results in request order.  With
    realtime, each request arrives at its spec's arrival time; without, every
    request is submitted at t = 0 and its latency counts from there."""

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
                results[k].arrival = 0.0  # every request submitted at the start
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


class BatchServer:
    """Continuous batching on the batch path: admitted scenes advance together,
    one batched advection step (one graph launch) at a time; a request that
    arrives while others run is admitted at the next step once a slot and
    enough of the HBM budget are free, a scene leaves the batch when it
    reaches its end, and each scene's time to solution is recorded.  A worker
    thread builds and initialises arriving scenes (on their own contexts)
    while the batch steps, so admission overlaps device work.  Every scene's
    result is the one it has alone (the batch is bit-identical per scene)."""

    def __init__(self, lib, max_active=256, hbm_budget=200 << 30, options=None):
        self.lib, self.max_active, self.hbm_budget, self.options = lib, max_active, hbm_budget, options
        self.tuning = None

    def run(self, specs, keep_fields=False, realtime=True):
        import queue
        results = [SceneResult(s.ident, 0, arrival=s.arrival) for s in specs]
        order = sorted(range(len(specs)), key=lambda k: specs[k].arrival)
        ready = queue.Queue()
        admit = threading.Semaphore(0)
        state = {"bytes": 0, "error": None}
        lock = threading.Lock()
        t0 = time.perf_counter()

        def builder():
            # builds scenes in arrival order, each once admission allows it
            for k in order:
                spec = specs[k]
                need = spec.footprint_bytes()
                if need > self.hbm_budget:
                    state["error"] = ValueError(f"scene {spec.ident} needs {need} B, more than the HBM budget")
                    ready.put(None)
                    return
                if realtime:
                    wait = spec.arrival - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(wait)
                else:
                    results[k].arrival = 0.0  # every request submitted at the start
                admit.acquire()  # a slot
                while True:
                    with lock:
                        if state["bytes"] + need <= self.hbm_budget:
                            state["bytes"] += need
                            break
                    time.sleep(0.001)
                try:
                    sc = cases.DamBreak(self.lib, spec.geometry(), spec.precision)
                    sc.initialize()
                    sc.ctx.synchronize()
                except Exception as e:  # reported by the stepping loop
                    state["error"] = e
                    ready.put(None)
                    return
                ready.put((k, sc, need))

        for _ in range(self.max_active):
            admit.release()
        th = threading.Thread(target=builder, daemon=True)
        th.start()
        from . import ck
        batch = ck.BatchAdvection(self.lib, [], options=self.options)
        active = []  # (k, scene, bytes)
        done = 0
        try:
            while done < len(specs):
                # admit what is ready; wait for the next scene when nothing runs
                while True:
                    try:
                        item = ready.get(block=not active, timeout=None if not active else 0)
                    except queue.Empty:
                        break
                    if item is None:
                        raise state["error"]
                    k, sc, need = item
                    results[k].start = time.perf_counter() - t0
                    results[k].n_fluid = sc.fluid.n
                    batch.add(sc.advection, 0.0)
                    active.append((k, sc, need))
                # ParticleSortCK where each scene's own loop sorts (cases.DamBreak.advection_step)
                st = batch.exec(sort_now=[1 if (sc.iterations + 1) % 100 == 0 else 0 for _, sc, _ in active])
                if (st != 0).any():
                    bad = [active[i][0] for i in np.flatnonzero(st)]
                    raise capi.SphxError(f"scenes {bad} diverged")
                for i, (k, sc, _) in enumerate(active):
                    sc.time = float(batch.time[i])
                    sc.iterations += 1
                    sc.acoustic_steps += int(batch.steps[i])
                finished = [i for i, (k, sc, _) in enumerate(active)
                            if (specs[k].steps and sc.iterations >= specs[k].steps)
                            or (specs[k].t_end and sc.time >= specs[k].t_end)]
                for i in reversed(finished):
                    k, sc, need = active.pop(i)
                    batch.remove(i)
                    res = results[k]
                    res.finish = time.perf_counter() - t0
                    res.acoustic_steps = sc.acoustic_steps
                    res.advection_steps = sc.iterations
                    res.physical_time = sc.time
                    if keep_fields:
                        res.fields = {f: sc.fluid.read(f) for f in ("Position", "Velocity", "Compression")}
                    sc.close()
                    with lock:
                        state["bytes"] -= need
                    admit.release()
                    done += 1
            self.tuning = batch.tuning()
        finally:
            batch.close()
            for _, sc, _ in active:
                sc.close()
        th.join(timeout=60)
        return results, time.perf_counter() - t0


def summarize(results, wall):
    lat = np.array([r.latency for r in results])
    work = sum(r.n_fluid * r.acoustic_steps for r in results)
    return {"scenes": len(results), "wall_s": wall, "particle_steps": work,
            "particle_steps_per_s": work / wall,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_max_s": float(lat.max())}
