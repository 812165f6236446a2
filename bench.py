#!/usr/bin/env python3
"""bench.py -- particle-steps/s of the 3D dam break (weakly-compressible SPH,
CK dual-time loop) on one MI355X, FP32-mixed and plain FP64.

Workload: tank 3.2 x 1.0 x 1.5 (z up), water column 1.2 x 1.0 x 0.6 at x = 0,
dp = 0.00448 (8,008,376 fluid particles), 3 wall layers, rho0 = 1000,
g = 9.81, c0 = 10 * 2 sqrt(g 0.6), AcousticRiemannSolverCK, free-surface
density regularization.  One "step" is one advection step of the CK 3D
dam-break loop: compression summation, regularization, setup, advection dt,
the acoustic sub-steps (dt, 1st half, 2nd half) until they cover the
advection dt, position update, cell linked list and relations (and the
particle sort on every 100th step).  value = fluid particles x acoustic
sub-steps / wall time of the K timed steps, state resident in HBM.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
FP32_VALU_PEAK = 157.3e12      # MI355X vector FP32 (AMD figure, packed issue)
FP64_VALU_PEAK = 78.6e12       # MI355X vector FP64
METRIC = "particle-steps/s, 3D dam break FP32-mixed vs FP64; p50/p99 scene latency"

# Algorithmic HBM bytes per fluid particle (3D), counting each per-particle
# field once per pass and neighbour data as cache-resident:
#  pass 2 (AcousticStep2ndHalf interact + update + fused CFL signal speed):
#    reads (x,y,z,Vol) record, v, C, dC, F_prior, m; writes dC, F, C, rho
#  whole acoustic+advection step: the enumerated figure of the WCSPH step
#    (two fused passes + amortised advection-level work, walls included).
def _bytes(vec, scalar, rec):
    """Per-particle algorithmic bytes of the two acoustic passes (3D)."""
    pass1 = (rec + scalar * 4 + vec * 3) + (vec + scalar + vec) + 2 * 3 * 8
    #       reads: (x,y,z,Vol), p, C, m, rho, F_prior, F, v | writes F, dC, v | dpos r+w (f64)
    pass2 = (rec + vec + scalar * 2 + vec + scalar) + (scalar + vec + scalar + scalar)
    #       reads: (x,y,z,Vol), v, C, dC, F_prior, m      | writes dC, F, C, rho
    return pass1, pass2


PASS1_BYTES = {"mixed": _bytes(12, 4, 16)[0], "fp64": _bytes(24, 8, 32)[0]}
PASS2_BYTES = {"mixed": _bytes(12, 4, 16)[1], "fp64": _bytes(24, 8, 32)[1]}
STEP_BYTES = {"mixed": 340.0, "fp64": 545.0}
# Algorithmic FP32 flops per neighbour pair (FMA = 2, 1/sqrt = 1), counted from
# the pair arithmetic of the reference (acoustic_step_1st_half.hpp:89-111,157-180,
# acoustic_step_2nd_half.hpp:53-73,117-137) with the tabulated kernel as a cubic:
#  pass 1 inner: r (3), |r|^2 (5), 1/|r| (1), |r| and q (2), cubic (6), dW V (1),
#    (p_i + p_j) dW V / |r| (3), force += s r (6), dC += (p_i - p_j) dW V (3)       = 30
#  pass 1 wall: as inner plus e (3), the wall pressure p_w from the face acceleration (12)  = 45
#  pass 2 inner: r, |r|^2, 1/|r|, |r|, q, cubic, dW V (19), e (3), v_i - v_j (3), u (5),
#    div += u dW V (2), limiter PJump (4), s (1), force += s e (6)                  = 43
#  pass 2 wall: as inner plus 2 (v_i - v_w) (3), e.n and its sign (7), n' (3), u on n' (5) = 57
PASS1_FLOP_INNER, PASS1_FLOP_WALL = 30.0, 45.0
PASS2_FLOP_INNER, PASS2_FLOP_WALL = 43.0, 57.0
STEP_FLOP_PER_PAIR = PASS1_FLOP_INNER + PASS2_FLOP_INNER


PROFILE = os.path.join(ROOT, "profiles", "dambreak3d_8M_mixed.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_traffic(stage):
    """HBM bytes per launch of `stage` measured with rocprofv3 FETCH_SIZE and
    WRITE_SIZE on this workload (scaled by the copy calibration), from the
    committed profile summary; None when it holds no figure for the stage."""
    try:
        with open(PROFILE) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic" and parts[1] == stage:
                    return float(parts[2])
    except OSError:
        pass
    return None


class ProfileWindow:
    """Host timestamps around chosen regions of the run, for picking their
    kernel dispatches out of a rocprofv3 trace or counter collection of the
    whole run: CLOCK_BOOTTIME in ns, the clock of rocprofv3's dispatch
    timestamps, read after a synchronisation on both sides, so a dispatch of
    the region starts after the first stamp and ends before the second.  Each
    region is printed to stderr as "profile-window NAME T0 T1" and returned in
    the JSON line (sphinxsys_amd.profile_summary reads either).  Off, both
    calls do nothing."""

    def __init__(self, on, name="headline"):
        self.on, self.name, self.spans = on, name, {}
        self._t0 = None

    @staticmethod
    def now():
        return time.clock_gettime_ns(time.CLOCK_BOOTTIME)

    def resume(self, name=None):
        if self.on:
            self._t0 = (name or self.name, self.now())

    def pause(self):
        if self.on and self._t0 is not None:
            name, t0 = self._t0
            t1 = self.now()
            self.spans[name] = [t0, t1]
            log(f"profile-window {name} {t0} {t1}")
            self._t0 = None


NO_WINDOW = ProfileWindow(False)


def timed_window(sc, steps, collect=NO_WINDOW):
    """K advection steps timed by the host clock (synchronised on both sides),
    with HIP events around the two acoustic passes only (events around every
    small kernel would slow the step by 2-3 %); `collect` brackets the steps
    for a profiler."""
    sc.ctx.synchronize()
    sc.ctx.profile(True, passes_only=True)
    a0, i0, t_phys = sc.acoustic_steps, sc.iterations, sc.time
    collect.resume()
    t0 = time.perf_counter()
    for _ in range(steps):
        sc.advection_step()
    sc.ctx.synchronize()
    wall = time.perf_counter() - t0
    collect.pause()
    passes = {k: sc.ctx.profile_read(k) for k in ("acoustic_pass1", "acoustic_pass2")}
    sc.ctx.profile(False)
    return {"wall_s": wall, "acoustic_steps": sc.acoustic_steps - a0, "advection_steps": sc.iterations - i0,
            "timed_passes": passes, "t_start": t_phys, "t_end": sc.time,
            "sorts": sum(1 for k in range(i0, sc.iterations) if (k + 1) % 100 == 0)}


def stage_breakdown(sc, steps):
    """HIP events around every launch over `steps` advection steps (outside
    any timed window): ms per stage."""
    sc.ctx.profile(True)
    for _ in range(steps):
        sc.advection_step()
    sc.ctx.synchronize()
    prof = {k: sc.ctx.profile_read(k) for k in (
        "acoustic_pass1", "acoustic_pass2", "acoustic_init", "compression_summation",
        "relation_build", "relation_build_contact", "relation_scan", "relation_stage", "cell_linked_list",
        "particle_sort",
        "time_step_reduce", "state_update")}
    sc.ctx.profile(False)
    return prof


def run_gpu(lib, geom, precision, warmup, steps, sustain_t=0.0, collect=NO_WINDOW):
    """The scene on the GPU: a fresh-lattice window (after `warmup` steps) and,
    with sustain_t > 0, a mid-run window after the run has passed physical time
    sustain_t (started 40 advection steps after a ParticleSortCK, so that no
    sort falls inside it); then one ParticleSortCK timed alone for the amortised
    cost (the CK 3D case sorts every 100 advection steps, dambreak.cpp:213-216).
    `collect` brackets the headline window (the mid-run one when there is one)."""
    from sphinxsys_amd import cases
    t0 = time.time()
    sc = cases.DamBreak(lib, geom, precision)
    sc.initialize()
    log(f"  scene ready in {time.time() - t0:.1f}s ({sc.fluid.n} fluid, {sc.wall.n} wall)")
    for _ in range(warmup):
        sc.advection_step()
    out = {"n_fluid": sc.fluid.n, "n_wall": sc.wall.n,
           "fresh": timed_window(sc, steps, collect if sustain_t <= 0.0 else NO_WINDOW)}
    window = out["fresh"]
    if sustain_t > 0.0:
        t0 = time.perf_counter()
        last = t0
        while sc.time < sustain_t or sc.iterations % 100 != 40:
            sc.advection_step()
            if time.perf_counter() - last > 10.0:
                last = time.perf_counter()
                log(f"  t = {sc.time:.4f} s after {sc.iterations} advection steps, {last - t0:.0f} s")
        log(f"  mid-run window from t = {sc.time:.4f} s, advection step {sc.iterations}")
        out["sustained"] = timed_window(sc, steps, collect)
        window = out["sustained"]
    out["profile"] = stage_breakdown(sc, steps)
    inner_size, _, _ = sc.inner.read()
    wall_size, _, _ = sc.contact.read()
    out["mean_neighbors_inner"] = float(inner_size.mean())
    out["mean_neighbors_wall"] = float(wall_size.mean())
    probe, energy = sc.observe()
    out["observables"] = {"physical_time": sc.time, "probe_pressure": float(probe[0]), "energy": energy,
                          "front": sc.front.exec()}
    # one ParticleSortCK alone, with events around its launches
    sc.ctx.synchronize()
    sc.ctx.profile(True)
    sc.particle_sort.exec()
    sc.ctx.synchronize()
    out["sort_ms"] = sc.ctx.profile_read("particle_sort")[0]
    sc.ctx.profile(False)
    out["window"] = window
    sc.close()
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_cpu_baseline(geom, threads, precision):
    """The oracle (the CPU restatement of the CK par_host path, std::thread pool
    in place of TBB) on the same 8M workload: one full advection step after the
    initial configuration, in FP64 (precision 0) or FP32 (precision 2)."""
    from sphinxsys_amd import cases
    from tests._util import oracle_lib
    lib = oracle_lib()
    lib.lib.sphx_oracle_set_threads(threads)
    sc = cases.DamBreak(lib, geom, precision)
    sc.initialize()
    t0 = time.perf_counter()
    n = sc.advection_step()
    wall = time.perf_counter() - t0
    sc.close()
    return {"value": sc.fluid.n * n / wall, "wall_s": wall, "acoustic_steps": n}


def full_run(lib, geom, precision, t_end, observe_dt):
    """The 3D dam break from rest to t_end: probe pressure, total mechanical
    energy and surge front (UpperFrontInAxisDirectionCK along x) every
    observe_dt of physical time; returns the series and the time to solution."""
    from sphinxsys_amd import cases
    sc = cases.DamBreak(lib, geom, precision)
    sc.initialize()
    sc.ctx.synchronize()
    series = {"t": [], "p": [], "e": [], "front": []}

    def obs():
        p, e = sc.observe()
        series["t"].append(sc.time)
        series["p"].append(float(p[0]))
        series["e"].append(e)
        series["front"].append(sc.front.exec())

    obs()
    next_obs = observe_dt
    t0 = time.perf_counter()
    last_log = t0
    while sc.time < t_end:
        sc.advection_step()
        if sc.time >= next_obs:
            obs()
            next_obs += observe_dt
        now = time.perf_counter()
        if now - last_log > 20.0:
            log(f"  t = {sc.time:.4f} s after {sc.iterations} advection steps, {now - t0:.0f} s")
            last_log = now
    sc.ctx.synchronize()
    wall = time.perf_counter() - t0
    out = {"series": series, "wall_s": wall, "advection_steps": sc.iterations,
           "acoustic_steps": sc.acoustic_steps, "n_fluid": sc.fluid.n}
    sc.close()
    return out


def compare_runs(a, b):
    """Error of run a against run b (same observation times): probe pressure
    history, its time average and first-impact peak, and the surge front."""
    import numpy as np
    pa, pb = np.array(a["p"]), np.array(b["p"])
    fa, fb = np.array(a["front"]), np.array(b["front"])
    n = min(len(pa), len(pb))
    pa, pb, fa, fb = pa[:n], pb[:n], fa[:n], fb[:n]
    t = np.array(b["t"][:n])
    scale = max(np.abs(pb).max(), 1e-300)
    ka, kb = int(np.argmax(pa)), int(np.argmax(pb))
    return {
        "probe_max_abs_err_over_peak": float(np.abs(pa - pb).max() / scale),
        "probe_rel_l2": float(np.linalg.norm(pa - pb) / max(np.linalg.norm(pb), 1e-300)),
        "probe_time_avg_rel_err": float(abs(pa.mean() - pb.mean()) / max(abs(pb.mean()), 1e-300)),
        "probe_peak_rel_err": float(abs(pa[ka] - pb[kb]) / scale),
        "probe_peak_time": [float(t[ka]), float(t[kb])],
        "front_max_rel_err": float(np.abs(fa - fb).max() / max(np.abs(fb).max(), 1e-300)),
        # first observation with the front within 0.2% of its largest value (the end wall)
        "front_reaches_end_wall_at": [float(next((tt for tt, f in zip(t, fa) if f >= 0.998 * fb.max()), -1)),
                                      float(next((tt for tt, f in zip(t, fb) if f >= 0.998 * fb.max()), -1))],
    }


def traversal_ab(lib, geom, warmup, calls):
    """The second acoustic half step on the stored lists against the same step
    with the inner neighbours found by a cell traversal, on two copies of the
    scene advanced identically (deterministic kernels: bit-identical states):
    the first call's outputs are compared bit for bit, then `calls` calls of
    each are timed with HIP events on the library's stream, alternating."""
    import numpy as np
    from sphinxsys_amd import _capi as capi
    from sphinxsys_amd import cases, ck
    scenes = []
    for _ in range(2):
        sc = cases.DamBreak(lib, geom, capi.MIXED)
        sc.initialize()
        for _ in range(warmup):
            sc.advection_step()
        sc.ctx.synchronize()
        scenes.append(sc)
    a, b = scenes
    steps = {"lists": ck.AcousticStep2ndHalf(a.inner, a.contact),
             "traversal": ck.AcousticStep2ndHalfTraversal(b.inner, b.contact)}
    owner = {"lists": a, "traversal": b}
    dt = 1.0e-5
    for name in steps:
        steps[name].exec(dt)
        owner[name].ctx.synchronize()
    outputs = ("Force", "CompressionRate", "Compression", "Density", "Displacement")
    identical = all(np.array_equal(a.fluid.read(k), b.fluid.read(k)) for k in outputs)
    ms = {name: [] for name in steps}
    for _ in range(calls):
        for name, step in steps.items():
            ctx = owner[name].ctx
            ctx.profile(True, passes_only=True)
            step.exec(dt)
            ctx.synchronize()
            t, n = ctx.profile_read("acoustic_pass2")
            ctx.profile(False)
            ms[name].append(t / max(n, 1))
    inner_size, _, _ = a.inner.read()
    out = {"metric": "second acoustic half step, stored lists against cell traversal (pass-2 kernel ms)",
           "workload": f"3D dam break, {a.fluid.n} fluid + {a.wall.n} wall particles, FP32-mixed, "
                       f"after {warmup} advection steps",
           "pass2_ms": {k: float(np.median(v)) for k, v in ms.items()},
           "pass2_ms_min": {k: float(np.min(v)) for k, v in ms.items()},
           "calls": calls, "outputs_bit_identical": identical,
           "mean_inner_neighbors": float(inner_size.mean())}
    out["traversal_over_lists"] = out["pass2_ms"]["traversal"] / out["pass2_ms"]["lists"]
    for sc in scenes:
        sc.close()
    return out


def copy_calibration(nbytes=1 << 30, repeats=3):
    """Plain device-to-device copies of a known size through the HIP runtime:
    the reference reading of the HBM counters (FETCH_SIZE / WRITE_SIZE)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(repeats):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {repeats} x {nbytes >> 20} MiB device-to-device copies")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30, help="timed advection steps")
    ap.add_argument("--warmup", type=int, default=3, help="untimed advection steps")
    ap.add_argument("--dp", type=float, default=0.00448, help="particle spacing")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=len(os.sched_getaffinity(0)),
                    help="threads of the CPU baseline (default: every CPU this process may run on)")
    ap.add_argument("--sustain", type=float, default=0.6, metavar="T",
                    help="physical time after which the mid-run (headline) window is timed; 0: fresh only")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device copy (reference kernel for HBM counters)")
    ap.add_argument("--profile-window", action="store_true",
                    help="print host timestamps of the headline window of the FP32-mixed run (of the serial "
                         "and batched phases with --batched / --served, and of the copy calibration) for "
                         "python -m sphinxsys_amd.profile_summary")
    ap.add_argument("--full-run", type=float, default=0.0, metavar="T_END",
                    help="instead of the throughput line: run the scene to T_END in FP32-mixed and FP64 "
                         "and compare probe pressure and surge front (one JSON line)")
    ap.add_argument("--observe", type=float, default=0.01, help="observation interval of --full-run")
    ap.add_argument("--batched", type=int, default=0, metavar="N",
                    help="instead of the throughput line: N independent 2D tanks (1k-64k particles), "
                         "serial plain runs against the batch (one graph launch per advection step)")
    ap.add_argument("--served", type=int, default=0, metavar="N",
                    help="instead of the throughput line: a Poisson stream of N 3D dam-break requests "
                         "(20k-2M particles, t_end 1.0 s); p50/p99 time to solution")
    ap.add_argument("--serial-only", action="store_true",
                    help="with --batched / --served: only the serial plain runs (per-scene times)")
    ap.add_argument("--rate", type=float, default=4.0, help="arrival rate of --served (requests/s)")
    ap.add_argument("--slots", type=int, default=256, help="concurrent scenes of --batched / --served")
    ap.add_argument("--traversal-ab", action="store_true",
                    help="instead of the throughput line: the second acoustic half step on the stored lists "
                         "against the cell traversal, on the same scene (bit-identity and pass-2 ms)")
    ap.add_argument("--microbench", type=int, default=0, metavar="M",
                    help="instead of the throughput line: the neighbour-search + density-summation "
                         "microbenchmark on an M^3 jittered lattice (M=256: 16.8M particles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(independent scenes shard across GPUs without communication)", file=sys.stderr)
        sys.exit(2)

    import sphinxsys_amd
    from sphinxsys_amd import _capi as capi
    from sphinxsys_amd import cases

    lib = sphinxsys_amd.load()
    collect = ProfileWindow(args.profile_window)
    if args.copy_calibration:
        collect.resume("copy_calibration")
        copy_calibration()
        collect.pause()
    t0 = time.time()
    geom = cases.dambreak_3d(dp=args.dp)
    log(f"geometry: {len(geom.fluid_pos)} fluid, {len(geom.wall_pos)} wall in {time.time() - t0:.1f}s")

    if args.batched > 0 or args.served > 0:
        from sphinxsys_amd import serving
        realtime = args.served > 0
        if args.batched > 0:
            specs = serving.batched_2d_specs(args.batched)
            what = f"{args.batched} independent 2D dam-break tanks, 40 advection steps each, FP32-mixed"
        else:
            specs = serving.served_3d_specs(args.served, rate=args.rate)
            what = (f"Poisson stream of {args.served} 3D dam breaks at {args.rate}/s, "
                    "20k-2M fluid particles, t_end 1.0 s, FP32-mixed")
        t0 = time.time()
        for sp in specs:
            sp.geometry()  # the requests carry their particle sets
        log(f"{len(specs)} scenes generated in {time.time() - t0:.1f}s "
            f"({sum(sp.fluid_count() for sp in specs)} fluid particles)")
        arrivals = "arrivals in real time" if realtime else "all requests at t = 0"
        log(f"serial: one plain run after another, {arrivals}")
        collect.resume("serial")
        res1, wall1 = serving.Server(lib, max_active=1).run(specs, realtime=realtime)
        collect.pause()
        serial = serving.summarize(res1, wall1)
        log(f"  {serial['particle_steps_per_s']:.3g} particle-steps/s in {wall1:.1f}s, "
            f"p50 {serial['latency_p50_s']:.2f}s p99 {serial['latency_p99_s']:.2f}s")
        per_scene = [{"n_fluid": r.n_fluid, "acoustic_steps": r.acoustic_steps, "serial_run_s": r.finish - r.start,
                      "serial_rate": r.n_fluid * r.acoustic_steps / max(r.finish - r.start, 1e-9)} for r in res1]
        if args.serial_only:
            print(json.dumps({"metric": "serial plain runs of independent scenes", "workload": what,
                              "serial": serial, "scenes": per_scene}), flush=True)
            return
        log(f"batched: continuous batching, one graph launch per advection step, {arrivals}")
        server = serving.BatchServer(lib, max_active=args.slots)
        collect.resume("batched")
        res2, wall2 = server.run(specs, realtime=realtime)
        collect.pause()
        conc = serving.summarize(res2, wall2)
        log(f"  {conc['particle_steps_per_s']:.3g} particle-steps/s in {wall2:.1f}s, "
            f"p50 {conc['latency_p50_s']:.2f}s p99 {conc['latency_p99_s']:.2f}s")
        same = all(a.acoustic_steps == b.acoustic_steps and a.physical_time == b.physical_time
                   and a.advection_steps == b.advection_steps for a, b in zip(res1, res2))
        for d, r in zip(per_scene, res2):
            d["batched_latency_s"] = r.latency
        last_arrival = max(sp.arrival for sp in specs)
        print(json.dumps({"metric": "aggregate particle-steps/s and time to solution of independent scenes",
                          "workload": what, "slots": args.slots, "serial": serial, "batched": conc,
                          "speedup": conc["particle_steps_per_s"] / serial["particle_steps_per_s"],
                          "last_arrival_s": last_arrival,
                          "arrival_bound": bool(realtime and last_arrival > 0.5 * wall2),
                          "same_steps_and_time_alone_and_in_mix": same,
                          "scenes": per_scene,
                          "batch_tuning": server.tuning}), flush=True)
        return
    if args.traversal_ab:
        print(json.dumps(traversal_ab(lib, geom, args.warmup, args.steps)), flush=True)
        return
    if args.microbench > 0:
        import subprocess
        exe = os.path.join(ROOT, "sphinxsys_amd", "lib", "microbench")
        out = subprocess.run([exe, "--m", str(args.microbench), "--repeat", "5"], capture_output=True, text=True)
        if out.returncode != 0:
            log(out.stderr)
            sys.exit(out.returncode)
        res = [json.loads(l) for l in out.stdout.splitlines() if l.startswith("{")]
        print(json.dumps({"metric": "neighbour search + density summation microbenchmark, stage ms",
                          "lattice": f"{args.microbench}^3 jittered, mt19937_64 seed 20261202",
                          "results": res}), flush=True)
        return
    if args.full_run > 0:
        runs = {}
        for name, prec in (("mixed", capi.MIXED), ("fp64", capi.FP64)):
            log(f"{name} run to t = {args.full_run} s")
            runs[name] = full_run(lib, geom, prec, args.full_run, args.observe)
        cmp = compare_runs(runs["mixed"]["series"], runs["fp64"]["series"])
        line = {"metric": "3D dam break time to solution and FP32-mixed error against FP64",
                "t_end": args.full_run, "n_fluid": runs["mixed"]["n_fluid"],
                "time_to_solution_s": {k: v["wall_s"] for k, v in runs.items()},
                "particle_steps_per_s": {k: v["n_fluid"] * v["acoustic_steps"] / v["wall_s"]
                                         for k, v in runs.items()},
                "advection_steps": {k: v["advection_steps"] for k, v in runs.items()},
                "acoustic_steps": {k: v["acoustic_steps"] for k, v in runs.items()},
                "mixed_vs_fp64": cmp,
                "series": {k: v["series"] for k, v in runs.items()}}
        print(json.dumps(line), flush=True)
        return

    log("FP32-mixed run")
    mixed = run_gpu(lib, geom, capi.MIXED, args.warmup, args.steps, sustain_t=args.sustain, collect=collect)
    fp64 = None
    if not args.no_fp64:
        log("FP64 run")
        fp64 = run_gpu(lib, geom, capi.FP64, args.warmup, args.steps)
    cpu = {}
    if not args.no_cpu:
        for name, prec in (("fp64", capi.FP64), ("fp32", capi.FP32)):
            log(f"CPU baseline (oracle, {name}, {args.cpu_threads} threads)")
            cpu[name] = run_cpu_baseline(geom, args.cpu_threads, prec)

    n_f = mixed["n_fluid"]

    def rate(run, window):
        """particle-steps/s of a timed window with the amortised ParticleSortCK
        (one sort per 100 advection steps, as the CK 3D case) added to its wall
        time; sorts inside the window are already in it."""
        w = run[window]
        extra = (w["advection_steps"] / 100.0 - w["sorts"]) * run["sort_ms"] * 1e-3
        wall = w["wall_s"] + max(extra, 0.0)
        return n_f * w["acoustic_steps"] / wall, wall

    head = "sustained" if "sustained" in mixed else "fresh"
    value, wall = rate(mixed, head)
    ms_per_step = 1000.0 * wall / args.steps
    win = mixed[head]

    # dominant kernel: the acoustic pass with the larger total time in the headline window
    p1_ms, p1_n = win["timed_passes"]["acoustic_pass1"]
    p2_ms, p2_n = win["timed_passes"]["acoustic_pass2"]
    k_in, k_w = mixed["mean_neighbors_inner"], mixed["mean_neighbors_wall"]
    if p2_ms >= p1_ms:
        dom, dom_ms, dom_n, dom_b = "acoustic_pass2", p2_ms, p2_n, PASS2_BYTES["mixed"]
        dom_f = PASS2_FLOP_INNER * k_in + PASS2_FLOP_WALL * k_w
    else:
        dom, dom_ms, dom_n, dom_b = "acoustic_pass1", p1_ms, p1_n, PASS1_BYTES["mixed"]
        dom_f = PASS1_FLOP_INNER * k_in + PASS1_FLOP_WALL * k_w
    avg_ms = dom_ms / max(dom_n, 1)
    achieved = dom_b * n_f / (avg_ms * 1e-3) / 1e9
    flops = dom_f * n_f / (avg_ms * 1e-3) / 1e12
    hbm_frac, valu_frac = achieved / HBM_PEAK_GBS, flops * 1e12 / FP32_VALU_PEAK
    k_bar = k_in + k_w
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-mixed (f32 pair math, f64 position/time)",
        "data": "synthetic (lattice-generated dam-break geometry)",
        "config": {
            "workload": f"3D dam break, tank 3.2x1.0x1.5, column 1.2x1.0x0.6, dp={args.dp}, "
                        f"{n_f} fluid + {mixed['n_wall']} wall particles, FP32-mixed",
            "fluid_particles": n_f,
            "wall_particles": mixed["n_wall"],
            "acoustic_steps_timed": win["acoustic_steps"],
            "window": f"{head}: advection steps {args.warmup if head == 'fresh' else 'from ' + str(round(win['t_start'], 4)) + ' s'}"
                      f" to t = {win['t_end']:.4f} s, ParticleSortCK amortised at 1 per 100 advection steps",
            "mean_neighbors": k_bar,
            "parallelism": "single GPU",
        },
        # the roof that binds the dominant kernel (the larger fraction) gives the
        # top-level figures; the other roof is reported beside it
        "roofline": dict(
            ({"bound": "valu", "achieved": flops, "peak": FP32_VALU_PEAK / 1e12, "unit": "TFLOP/s",
              "frac": valu_frac} if valu_frac > hbm_frac else
             {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm_frac}),
            kernel=dom,
            traffic=profiled_traffic(dom),
            traffic_source="profiles/dambreak3d_8M_mixed.md (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)",
            avg_launch_ms=avg_ms,
            hbm={"achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm_frac,
                 "bytes_per_particle": dom_b},
            valu={"achieved": flops, "peak": FP32_VALU_PEAK / 1e12, "unit": "TFLOP/s", "frac": valu_frac,
                  "flop_per_particle": dom_f,
                  "note": "algorithmic FP32 flops of the pair loops (FMA = 2) at the measured mean "
                          "neighbour counts against the published vector FP32 peak, which counts packed "
                          "FP32; this build issues no packed FP32, so its issue-limited ceiling is half "
                          "of that peak"}),
        "step_roofline": {
            "hbm_frac": STEP_BYTES["mixed"] * value / (HBM_PEAK_GBS * 1e9),
            "valu_frac": STEP_FLOP_PER_PAIR * k_bar * value / FP32_VALU_PEAK,
            "bytes_per_particle_step": STEP_BYTES["mixed"],
            "flop_per_particle_step": STEP_FLOP_PER_PAIR * k_bar,
        },
        "windows": {name: {"value": rate(mixed, name)[0], "ms_per_step": 1000.0 * rate(mixed, name)[1] / args.steps,
                           "t_start": mixed[name]["t_start"], "t_end": mixed[name]["t_end"],
                           "acoustic_steps": mixed[name]["acoustic_steps"]}
                    for name in ("fresh", "sustained") if name in mixed},
        "sort_ms": mixed["sort_ms"],
        "stage_ms": {k: round(v[0], 3) for k, v in mixed["profile"].items()},
        "stage_ms_source": f"HIP events around every launch over the {args.steps} advection steps after the "
                           f"headline window (the timed steps time only the two acoustic passes)",
        "observables": mixed["observables"],
        **({"profile_windows": collect.spans} if collect.on else {}),
    }
    if fp64:
        v64 = rate(fp64, "fresh")[0]
        line["fp64"] = {"value": v64, "window": "fresh",
                        "ms_per_step": 1000.0 * rate(fp64, "fresh")[1] / args.steps,
                        "acoustic_steps_timed": fp64["fresh"]["acoustic_steps"], "sort_ms": fp64["sort_ms"],
                        "stage_ms": {k: round(v[0], 3) for k, v in fp64["profile"].items()},
                        "observables": fp64["observables"]}
        # like for like: both fresh-lattice windows
        line["mixed_over_fp64"] = rate(mixed, "fresh")[0] / v64
    if cpu:
        c64 = cpu["fp64"]
        line["cpu_baseline"] = {
            "value": c64["value"], "unit": "particle-steps/s", "cores": args.cpu_threads,
            "kind": "port",
            "sample": f"one advection step ({c64['acoustic_steps']} acoustic sub-steps + configuration update) "
                      f"of the same 8M workload from the initial lattice, FP64, {c64['wall_s']:.1f}s on "
                      f"{args.cpu_threads} threads of {cpu_model()} (nproc {os.cpu_count()})",
            "fp32": {"value": cpu["fp32"]["value"], "wall_s": cpu["fp32"]["wall_s"],
                     "acoustic_steps": cpu["fp32"]["acoustic_steps"]},
            "cpu_model": cpu_model(), "nproc": os.cpu_count()}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
