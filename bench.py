"""Benchmark of the fermion-matrix hot path on one MI355X.

Workload (default, the largest single-GPU configuration): 4D 32^3 x 64 SU(3)
Wilson D^dag D, kappa = 0.1340, links U = exp(i 0.2 H) (seeded), one
Gaussian source, multishift CG over 24 shifts log-spaced in [1e-5, 20] (a
rational (D^dag D)^{-1/2}-style shift list), every shift to true relative
residual 1e-10 recomputed in FP64; Krylov vectors FP64, then complex64, then
16-bit storage (int16 + per-site norm) with FP32 arithmetic, shift solutions
and recurrences in FP64.  One step = one complete multishift solve.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c5|c2|c3|...]
                    [--precision half|mixed|double]

Prints one JSON line: metric, value = solves/s, ms_per_step, p50/p99 solve
latency, the plain FP64 solver's rate on the same solve, every shift's FP64
true residual, the roofline of the solve's dominant kernel (HIP-event
durations of its launches over a further solve of the same system after the
timed region, algorithmic bytes), and the CPU baseline (torch-CPU complex128
restatement of the reference's op sequence, 4 threads and all cores).
Beside it, the default line carries the 24^3 x 48 measurements: D and the
force sum with their rooflines, the batched 28-source CG (16-bit storage),
and the batched multishift CG over 8 sources x 12 shifts, each batch with
its per-kernel table and the roofline of its dominant and dslash kernels.
Serving workloads (c4, c4e) report capacity, the sustained rate at the
offered rate, and latency at 70% of capacity.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    "c2": dict(geometry=(16, 16, 16, 32), nc=3, kappa=0.13, eps=0.3, nshift=12, lo=1e-4, hi=10.0,
               name="wilson_ddag_d_16c3x32_su3_multishift12"),
    "c5": dict(geometry=(32, 32, 32, 64), nc=3, kappa=0.1340, eps=0.2, nshift=24, lo=1e-5, hi=20.0,
               name="wilson_ddag_d_32c3x64_su3_multishift24"),
    "c3": dict(geometry=(24, 24, 24, 48), nc=3, kappa=0.1335, eps=0.25, nshift=12, lo=1e-4, hi=10.0,
               name="wilson_ddag_d_24c3x48_su3_multishift12"),
}

BATCHED = {
    # 12 point sources (one per spin-colour at the origin) + 16 Z2-noise sources,
    # CG on D^dag D (shift 0) to true relative residual 1e-10
    # (precision: half = FP16 Krylov storage with per-site norms, FP32
    # arithmetic, FP64 reliable updates)
    "c3b": dict(geometry=(24, 24, 24, 48), nc=3, kappa=0.1335, eps=0.25, npoint=12, nz2=16, precision="half",
                name="wilson_ddag_d_24c3x48_su3_batched28"),
    "c2b": dict(geometry=(16, 16, 16, 32), nc=3, kappa=0.13, eps=0.3, npoint=12, nz2=16, precision="half",
                name="wilson_ddag_d_16c3x32_su3_batched28"),
    # batched multishift: 4 point + 4 Z2 sources, 12 shifts log-spaced in
    # [1e-4, 10], every (source, shift) to 1e-10 (16-bit storage phases)
    "c3m": dict(geometry=(24, 24, 24, 48), nc=3, kappa=0.1335, eps=0.25, npoint=4, nz2=4, precision="half",
                nshift=12, lo=1e-4, hi=10.0, name="wilson_ddag_d_24c3x48_su3_batched8_multishift12"),
}

SERVING = {
    "c4": dict(volumes=[(8, 8, 8, 8), (12, 12, 12, 12), (16, 16, 16, 32), (24, 24, 24, 48)], colours=[2, 3],
               rate=200.0, kappa=(0.12, 0.135), max_shifts=16, eps=0.25, requests=96, slots=4, per_field=False,
               name="serving_poisson200_8fields_1to16shifts"),
    # the same fields and stream, but kappa and the shift list drawn once per
    # field (an ensemble: every request on a field is the same operator and
    # rational approximation), so waiting requests on a field merge
    "c4e": dict(volumes=[(8, 8, 8, 8), (12, 12, 12, 12), (16, 16, 16, 32), (24, 24, 24, 48)], colours=[2, 3],
                rate=200.0, kappa=(0.12, 0.135), max_shifts=16, eps=0.25, requests=96, slots=4, per_field=True,
                name="serving_poisson200_8fields_per_field_operator"),
}

HBM_PEAK_GBS = 8000.0


def run_serving(args):
    """Serving stream against the continuous-admission solve server, in two
    phases on one server after a warm-up: (1) the configured Poisson stream
    (offered 200/s): the sustained rate, which is the capacity when the offer
    exceeds it, and latency, which then measures the queue; (2) a Poisson
    stream offered at 70% of that capacity: p50/p99 latency at a load the
    server sustains."""
    import torch
    from tfleonardym_amd import _lib, fields
    from tfleonardym_amd.serving import GaugeStore, SolveRequest, SolveServer, latency_summary
    _lib.require_device()
    W = SERVING[args.workload]
    rng = np.random.default_rng(20270)
    store = GaugeStore()
    keys = []
    for vi, g in enumerate(W["volumes"]):
        for nc in W["colours"]:
            k = f"{'x'.join(map(str, g))}_su{nc}"
            store.add(k, fields.gauge_field(g, nc, seed=20280 + 10 * vi + nc, kind="near_identity", eps=W["eps"]), g)
            ns = int(rng.integers(1, W["max_shifts"] + 1))
            keys.append(dict(key=k, g=g, nc=nc, kappa=float(rng.uniform(*W["kappa"])), ns=ns))
    nreq = args.steps if args.steps > 5 else W["requests"]
    counter = [0]

    def stream(rate, n):
        """n requests with Poisson arrivals at `rate` (rate None: all at t = 0)."""
        out, t = [], 0.0
        for _ in range(n):
            i = counter[0]
            counter[0] += 1
            if rate is not None:
                t += rng.exponential(1.0 / rate)
            f = keys[rng.integers(len(keys))]
            kappa, ns = (f["kappa"], f["ns"]) if W["per_field"] else (
                float(rng.uniform(*W["kappa"])), int(rng.integers(1, W["max_shifts"] + 1)))
            shifts = fields.log_shifts(1e-4, 10.0, ns) if ns > 1 else [1e-4]
            src = torch.from_numpy(fields.gaussian_spinor(f["g"], 4, f["nc"], seed=30000 + i))
            out.append((t, SolveRequest(field=f["key"], kappa=kappa, shifts=shifts, source=src, rid=i)))
        return out

    def play(srv, reqs):
        """Submit at the arrival times; returns the results and the wall time
        from the first arrival to the last completion."""
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        futs = []
        for ta, r in reqs:
            delay = t0 + ta - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            r.arrival = t0 + ta
            futs.append(srv.submit(r))
        res = [f.result() for f in futs]
        wall = max(r.done for r in res) - t0
        bad = [r for r in res if r.status != "converged" or max(r.stats["true_rel_resid"]) > 1e-10]
        if bad:
            print(f"[bench] {len(bad)} requests missed 1e-10", file=sys.stderr)
            sys.exit(1)
        return res, wall

    print(f"[bench] serving: {len(keys)} resident fields ({store.resident_bytes() / 2**30:.2f} GiB), "
          f"{nreq} requests per phase", file=sys.stderr, flush=True)
    with SolveServer(store, slots=W["slots"], precision_mode=args.precision) as srv:
        for f in keys:  # warm the one-shift and multishift paths once per field
            for shifts in ([1.0], fields.log_shifts(1e-4, 10.0, W["max_shifts"])):
                srv.submit(SolveRequest(field=f["key"], kappa=0.12, shifts=shifts, source=torch.from_numpy(
                    fields.gaussian_spinor(f["g"], 4, f["nc"], seed=1)))).result()
        srv.batches.clear()
        off_res, off_wall = play(srv, stream(W["rate"], nreq))
        sustained = nreq / off_wall
        lat = latency_summary(off_res)
        print(f"[bench] offered {W['rate']:.0f}/s: sustained {sustained:.1f} solves/s, p50 {lat['p50_ms']:.0f} ms, "
              f"p99 {lat['p99_ms']:.0f} ms", file=sys.stderr, flush=True)
        # above capacity the sustained rate is the capacity; the second stream
        # is offered at 70% of it
        capacity = sustained
        low_rate = 0.7 * capacity
        low_res, low_wall = play(srv, stream(low_rate, nreq))
        low_lat = latency_summary(low_res)
        print(f"[bench] offered {low_rate:.1f}/s: sustained {nreq / low_wall:.1f} solves/s, "
              f"p50 {low_lat['p50_ms']:.0f} ms, p99 {low_lat['p99_ms']:.0f} ms", file=sys.stderr, flush=True)
        inflight = srv.max_inflight_seen
        batches = {int(k): int(v) for k, v in sorted(srv.batches.items())}
    service_ms = np.array([r.done - r.start for r in low_res]) * 1e3
    out = {
        "metric": "serving solves/s at 1e-10 true resid, p50/p99 latency",
        "value": sustained, "unit": "solves/s", "n_gpus": 1, "steps": nreq, "warmup": len(keys),
        "ms_per_step": off_wall / nreq * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"mixed": "f32+f64", "double": "f64", "half": "f16 storage (int16 + per-site norm) + f32 + f64"}[
            args.precision],
        "data": "synthetic (seeded Poisson arrivals, near-identity SU(2)/SU(3) fields, Gaussian sources)",
        "config": {"workload": W["name"], "rate_per_s": W["rate"], "slots": W["slots"], "fields": len(keys),
                   "kappa": "per field" if W["per_field"] else "per request", "precision": args.precision},
        "offered_rate_per_s": W["rate"],
        "sustained_rate_per_s": sustained,
        "latency_ms": {"p50": lat["p50_ms"], "p99": lat["p99_ms"], "mean": lat["mean_ms"]},
        "capacity_solves_per_s": capacity,
        "at_sustainable_rate": {"offered_rate_per_s": low_rate, "sustained_rate_per_s": nreq / low_wall,
                                "latency_ms": {"p50": low_lat["p50_ms"], "p99": low_lat["p99_ms"],
                                               "mean": low_lat["mean_ms"]},
                                "service_ms": {"p50": float(np.percentile(service_ms, 50)),
                                               "p99": float(np.percentile(service_ms, 99))}},
        "merged_groups": batches,
        "max_inflight": inflight,
    }
    print(json.dumps(out))


def batched_result(workload, precision, steps, warmup, baseline_steps, autotune=True, timing=True):
    """Batched multi-RHS CG (all sources in the same launches; 16-bit Krylov
    storage with FP32 arithmetic and FP64 reliable updates unless `precision`
    says otherwise) against the plain FP64 solver run source by source, on
    sources spread over both kinds (point and Z2).  Returns the result dict
    (None values where skipped); exits if a solve misses 1e-10."""
    import torch
    from tfleonardym_amd import _lib, dirac, dirac_solver, fields, lattice
    dev = _lib.require_device()
    torch.cuda.empty_cache()  # earlier workloads' freed blocks go back to the device
    W = BATCHED[workload]
    geometry, nc = W["geometry"], W["nc"]
    U = fields.gauge_field(geometry, nc, seed=20250, kind="near_identity", eps=W["eps"])
    srcs = [fields.point_source(geometry, 4, nc, site=0, color=c, spin=s) for s in range(4) for c in range(nc)]
    srcs = srcs[::max(1, len(srcs) // W["npoint"])][:W["npoint"]]
    srcs = srcs + [fields.z2_noise(geometry, 4, nc, seed=20260 + i) for i in range(W["nz2"])]
    shifts = fields.log_shifts(W["lo"], W["hi"], W["nshift"]) if "nshift" in W else [0.0]
    srcs = [torch.from_numpy(s).to(dev) for s in srcs]
    op = dirac.DiracWilsonOperator(lattice.Configuration(U, geometry, nc, 4), W["kappa"], True,
                                   representation="fundamental")
    tiles = op.autotune() if autotune else {"f64": op.tile_info("f64"), "f32": op.tile_info("f32")}
    A = dirac.SquareOperator(op)
    mode = precision or W["precision"]

    # The plain FP64 solver, one source at a time, on sources spread over both
    # kinds.  Measured first, in the device-memory state of a standalone run,
    # and again after the batched solves: its shift updates stream 2 vectors
    # per shift and iteration, and their rate depends on where the allocator
    # places its buffers (at 24^3x48, 12 shifts: 2.00 solves/s standalone,
    # 2.16 after the batched solves' workspace has been cached).
    def plain_rate():
        ps = dirac_solver.multishift_solver(20000, 0.0, algorithm="cg", tolerance=1e-10, precision_mode="double",
                                            verbose=False)
        nb = min(len(srcs), max(2, baseline_steps))
        pick = sorted(set(np.linspace(0, len(srcs) - 1, nb).round().astype(int).tolist()))
        ps.solve(A, srcs[0], shifts)
        torch.cuda.synchronize()
        tb = time.perf_counter()
        for s in [srcs[i] for i in pick]:
            ps.solve(A, s, shifts)
        torch.cuda.synchronize()
        worst = max(ps.last_stats["true_rel_resid"])
        del ps
        return len(pick) / (time.perf_counter() - tb), worst

    plain = plain_after = None
    if baseline_steps > 0:
        plain, worst = plain_rate()
        torch.cuda.empty_cache()
        print(f"[bench] plain FP64 CG, one source at a time (standalone): {plain:.2f} solves/s, "
              f"true residual {worst:.2e}", file=sys.stderr, flush=True)
    solver = dirac_solver.multishift_solver(20000, 0.0, algorithm="cg", tolerance=1e-10, precision_mode=mode,
                                            verbose=False)
    for _ in range(warmup):
        solver.solve_batch(A, srcs, shifts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        solver.solve_batch(A, srcs, shifts)
        st = solver.last_stats
        worst = max(max(s["true_rel_resid"]) for s in st)
        print(f"[bench] batched {workload} step {i}: {st[0]['seconds'] * 1e3:.0f} ms, {st[0]['iterations']} it, "
              f"max true residual {worst:.2e}", file=sys.stderr, flush=True)
        if solver.last_status != "converged" or worst > 1e-10:
            print(f"[bench] batched solve did not reach 1e-10", file=sys.stderr)
            sys.exit(1)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    value = steps * len(srcs) / total
    # per-kernel durations: one further batch with HIP events around every
    # launch (direct launches, not part of the timed region)
    kernels = roofline = dslash_roofline = None
    if timing:
        _lib.kernel_timing(True)
        tk = time.perf_counter()
        solver.solve_batch(A, srcs, shifts)
        torch.cuda.synchronize()
        kt_total = time.perf_counter() - tk
        _lib.kernel_timing(False)
        kt = _lib.kernel_timing_read()
        if kt:
            kernels, roofline, dslash_roofline = kernel_table(
                kt, kt_total, W["name"],
                "HIP events on the solve stream around every launch of the class, over one further batch solve "
                "after the timed region (direct launches; the timed solves replay HIP graphs)")
            for tag, r in (("dominant kernel", roofline), ("dslash", dslash_roofline)):
                if r:
                    print(f"[bench] batched {workload} {tag} {r['kernel']}: {r['launches']} launches, "
                          f"{r['avg_us']:.1f} us avg, {r['achieved']:.0f} GB/s ({100 * r['frac']:.1f}%)",
                          file=sys.stderr, flush=True)
    if baseline_steps > 0:
        plain_after, _ = plain_rate()
        print(f"[bench] plain FP64 CG, one source at a time (after the batched solves): {plain_after:.2f} solves/s",
              file=sys.stderr, flush=True)
    return {
        "metric": ("batched multi-RHS CG solves/s at 1e-10 true resid" if len(shifts) == 1 else
                   "batched multishift CG solves/s (one solve = one source, all shifts) at 1e-10 true resid"),
        "value": value, "unit": "solves/s", "steps": steps, "warmup": warmup,
        "ms_per_step": total / steps * 1e3,
        "dtype": {"half": "f16 storage + f32 arithmetic + f64 reliable updates", "mixed": "f32+f64",
                  "double": "f64"}[mode],
        "config": {"workload": W["name"], "geometry": list(geometry), "kappa": W["kappa"], "rhs": len(srcs),
                   "shifts": len(shifts), "precision": mode, "tolerance": 1e-10},
        "iterations": solver.last_stats[0]["iterations"],
        "plain_fp64_single_rhs_solves_per_s": plain,
        "speedup_vs_plain_fp64": (value / plain) if plain else None,
        "plain_fp64_single_rhs_after_batch_solves_per_s": plain_after,
        "speedup_vs_plain_fp64_after_batch": (value / plain_after) if plain_after else None,
        "wilson_tiles": tiles,
        "roofline": roofline,
        "dslash_roofline": dslash_roofline,
        "kernels": kernels,
    }


def run_batched(args):
    """The batched workloads as the bench's own JSON line."""
    r = batched_result(args.workload, args.precision, args.steps, args.warmup, args.baseline_steps,
                       autotune=not args.no_autotune, timing=not args.no_kernel_timing)
    if args.calib_copy:
        calib_copy()
    out = {
        "metric": r["metric"], "value": r["value"], "unit": r["unit"], "n_gpus": 1, "steps": r["steps"],
        "warmup": r["warmup"], "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": r["dtype"],
        "data": "synthetic (seeded near-identity SU(3) links; point + Z2 sources)",
        "config": r["config"], "iterations": r["iterations"],
        "plain_fp64_single_rhs_solves_per_s": r["plain_fp64_single_rhs_solves_per_s"],
        "speedup_vs_plain_fp64": r["speedup_vs_plain_fp64"],
        "plain_fp64_single_rhs_after_batch_solves_per_s": r["plain_fp64_single_rhs_after_batch_solves_per_s"],
        "speedup_vs_plain_fp64_after_batch": r["speedup_vs_plain_fp64_after_batch"],
        "wilson_tiles": r["wilson_tiles"],
        "roofline": r["roofline"], "dslash_roofline": r["dslash_roofline"], "kernels": r["kernels"],
    }
    print(json.dumps(out))


def bytes_per_site_dslash(nc, ndim, wl, ws, link_rows=None):
    """Compulsory bytes of one D application per site: links read once
    (link_rows of nc complex each: nc, or nc - 1 when the kernels rebuild the
    last row of SU(nc) links), input and output spinor once (w = bytes per
    real)."""
    ns = 4 if ndim == 4 else 2
    rows = nc if link_rows is None else link_rows
    return ndim * rows * nc * 2 * wl + 2 * ns * nc * 2 * ws


# D at 24^3 x 48 SU(3) (the batched-solve lattice): roofline fractions are
# quoted here, where the working set does not fit the 256 MB MALL
DSLASH_C3 = dict(geometry=(24, 24, 24, 48), nc=3, kappa=0.1335, eps=0.25, seed=20250)


KERNEL_TIME_REPLAYS = 5


def kernel_time_us(launch, reps):
    """Duration in us of one `launch()` (a single kernel): the median over
    KERNEL_TIME_REPLAYS timed runs of the average over `reps` back-to-back
    launches.  The launches are captured into a HIP graph and replayed, so
    host-side launch overhead (ctypes, descriptor validation, a profiler's API
    interception) does not enter the figure; HIP events are recorded on the
    stream the graph replays on.  Falls back to timing the direct launches if
    the capture fails."""
    import torch
    for _ in range(3):
        launch()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    graph = None
    try:
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            for _ in range(reps):
                launch()
        graph.replay()
        torch.cuda.synchronize()
    except Exception as exc:  # capture not possible: direct launches
        print(f"[bench] graph capture failed ({exc}); timing direct launches", file=sys.stderr, flush=True)
        graph = None
        torch.cuda.synchronize()
    runs = []
    for _ in range(KERNEL_TIME_REPLAYS):
        e0.record(stream)
        if graph is not None:
            graph.replay()
        else:
            for _ in range(reps):
                launch()
        e1.record(stream)
        e1.synchronize()
        runs.append(e0.elapsed_time(e1) * 1e3 / reps)
    return float(np.median(runs))


def time_dslash(op, src, reps):
    """Average launch time of D (lym_wilson_apply, projected form) in
    complex128 and complex64 (kernel_time_us: HIP events around a graph of
    back-to-back launches); GB/s from the compulsory bytes per site."""
    import ctypes
    import torch
    from tfleonardym_amd import _lib
    desc = op.ensure_f32()
    nc, V = op.number_of_colors, op.volume
    rows = nc - 1 if desc.link_format == _lib.LINKS_SU else nc
    adjoint = op.representation == "adjoint"
    res = {}
    for prec, name, w in ((_lib.PREC_F64, "f64", 8), (_lib.PREC_F32, "f32", 4)):
        dt = torch.complex128 if prec == _lib.PREC_F64 else torch.complex64
        x = src.to(dt)
        y = torch.empty_like(x)

        def launch():
            _lib.check(_lib.lib().lym_wilson_apply(ctypes.byref(desc), prec, 0, _lib.ptr(x), _lib.ptr(y),
                                                   _lib.stream_handle()))
        us = kernel_time_us(launch, reps)
        # adjoint spinors are Ns x Nc x Nc matrices; their links are read in full
        bps = (4 * nc * nc * 2 * w + 2 * 4 * nc * nc * 2 * w) if adjoint else bytes_per_site_dslash(nc, 4, w, w, rows)
        gbs = bps * V / (us * 1e-6) / 1e9
        res[name] = dict(us=us, bytes_per_site=bps, gbs=gbs, frac=gbs / HBM_PEAK_GBS)
    return res


def bytes_per_site_force(nc, ndim, npairs, link_rows=None):
    """Compulsory bytes of one force launch per site (complex128): links read
    once (link_rows of nc complex per direction), the (D, nc, nc) output
    written once, X_p and Y_p of every pair read once."""
    ns = 4 if ndim == 4 else 2
    rows = nc if link_rows is None else link_rows
    return ndim * rows * nc * 16 + ndim * nc * nc * 16 + npairs * 2 * ns * nc * 16


def time_force(op, reps, npairs=4, seed=7):
    """Average launch time of the symmetrised force sum over `npairs` pairs
    (x_p, Q x_p) with the force_level1 rational weights (one
    lym_lie_derivative launch), timed by kernel_time_us."""
    import ctypes
    import torch
    from tfleonardym_amd import _lib, approximations
    desc = op.desc
    nc, V, D = op.number_of_colors, op.volume, op.number_of_dimensions
    g = torch.Generator(device="cpu").manual_seed(seed)
    xs = [torch.randn(op.spinor_shape, dtype=torch.complex128, generator=g).cuda() for _ in range(npairs)]
    ys = [op.apply(x) for x in xs]
    al, _ = approximations.split(approximations.force_level1_rational_approximation_coefficients)
    w = [al[p % len(al)] for p in range(npairs)]
    out = torch.empty((D, nc, nc, V), dtype=torch.complex128, device="cuda")
    xp = (ctypes.c_void_p * npairs)(*[t.data_ptr() for t in xs])
    yp = (ctypes.c_void_p * npairs)(*[t.data_ptr() for t in ys])
    wp = (ctypes.c_double * npairs)(*w)

    def launch():
        _lib.check(_lib.lib().lym_lie_derivative(ctypes.byref(desc), npairs, xp, yp, wp, 1, 0, _lib.ptr(out),
                                                 _lib.stream_handle()))
    us = kernel_time_us(launch, reps)
    rows = nc - 1 if desc.link_format == _lib.LINKS_SU else nc
    bps = bytes_per_site_force(nc, D, npairs, rows)
    gbs = bps * V / (us * 1e-6) / 1e9
    return dict(us=us, npairs=npairs, bytes_per_site=bps, gbs=gbs, frac=gbs / HBM_PEAK_GBS,
                kernel="k_force (complex128, symmetrised, force_level1 weights)")


def dslash_c3(reps, geometry=None, autotune=True):
    """D at 24^3 x 48 (seeded near-identity links, Gaussian input); another
    4D geometry with the same generator if given."""
    import torch
    from tfleonardym_amd import dirac, fields, lattice
    W = DSLASH_C3
    g = tuple(geometry) if geometry else W["geometry"]
    t0 = time.time()
    U = fields.gauge_field(g, W["nc"], seed=W["seed"], kind="near_identity", eps=W["eps"])
    src = torch.from_numpy(fields.gaussian_spinor(g, 4, W["nc"], seed=W["seed"] + 1)).cuda()
    op = dirac.DiracWilsonOperator(lattice.Configuration(U, g, W["nc"], 4), W["kappa"], True,
                                   representation="fundamental")
    tiles = op.autotune() if autotune else {n: op.tile_info(n) for n in ("f64", "f32")}
    res = time_dslash(op, src, reps)
    res["tiles"] = tiles
    res["force"] = time_force(op, max(5, reps // 5))
    res["force"]["traffic"] = measured_force_traffic(g)
    print(f"[bench] force {g} {res['force']['npairs']} pairs: {res['force']['us']:.1f} us/launch, "
          f"{res['force']['gbs']:.0f} GB/s ({100 * res['force']['frac']:.1f}% of {HBM_PEAK_GBS:.0f})",
          file=sys.stderr, flush=True)
    for name in ("f64", "f32"):
        res[name]["traffic"] = measured_traffic(g, name)
    res["geometry"] = list(g)
    for name in ("f64", "f32"):
        print(f"[bench] wilson {g} {name}: {res[name]['us']:.2f} us/launch, {res[name]['gbs']:.0f} GB/s "
              f"({100 * res[name]['frac']:.1f}% of {HBM_PEAK_GBS:.0f})", file=sys.stderr, flush=True)
    # the adjoint representation (the reference driver's default, matrix
    # spinors U X U^dag) on the same links
    src_a = torch.from_numpy(fields.gaussian_spinor(g, 4, W["nc"], seed=W["seed"] + 2, representation="adjoint")).cuda()
    op_a = dirac.DiracWilsonOperator(lattice.Configuration(U, g, W["nc"], 4), W["kappa"], True,
                                     representation="adjoint")
    tiles_a = op_a.autotune() if autotune else {n: op_a.tile_info(n) for n in ("f64", "f32")}
    res["adjoint"] = time_dslash(op_a, src_a, reps)
    res["adjoint"]["tiles"] = tiles_a
    for name in ("f64", "f32"):
        r = res["adjoint"][name]
        print(f"[bench] wilson adjoint {g} {name}: {r['us']:.2f} us/launch, {r['gbs']:.0f} GB/s "
              f"({100 * r['frac']:.1f}% of {HBM_PEAK_GBS:.0f})", file=sys.stderr, flush=True)
    print(f"[bench] dslash set-up + timing {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    return res


def measured_traffic(geometry, prec):
    """HBM bytes per launch of D at `geometry` / `prec` from the committed
    rocprofv3 summaries (FETCH_SIZE and WRITE_SIZE passes, calibrated on a
    copy kernel of known size), or None."""
    return measured_class_traffic("x".join(str(L) for L in geometry), prec)


def measured_force_traffic(geometry):
    """HBM bytes per force launch at `geometry` from the committed rocprofv3
    summaries (calibrated FETCH_SIZE + WRITE_SIZE), or None."""
    return measured_class_traffic("x".join(str(L) for L in geometry), "force")


def calib_copy():
    """Three copies of a 256 MiB float32 tensor on the device: the kernel of
    known size that calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    c = torch.empty_like(a)
    for _ in range(3):
        c.copy_(a)
    torch.cuda.synchronize()


def run_dslash(args):
    """Only D (for kernel profiles) at --dslash-geometry, then a 256 MiB
    device copy (the calibration kernel of the FETCH_SIZE / WRITE_SIZE
    counters): one JSON line."""
    import torch
    geometry = tuple(int(v) for v in args.dslash_geometry.split("x"))
    res = dslash_c3(args.dslash_reps, geometry, autotune=not args.no_autotune)
    calib_copy()
    f32 = res["f32"]
    print(json.dumps({"metric": f"Wilson D GB/s ({args.dslash_geometry} SU(3), compulsory bytes)", "value": f32["gbs"],
                      "unit": "GB/s", "n_gpus": 1, "steps": args.dslash_reps, "warmup": 5,
                      "ms_per_step": f32["us"] * 1e-3, "higher_is_better": True, "scaling": "weak",
                      "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded near-identity SU(3) links)",
                      "config": {"workload": f"wilson_d_{args.dslash_geometry}_su3"}, "dslash": res}))


def cpu_info():
    """Host CPU of this box: logical CPUs, the CPUs this process may use, model."""
    import platform
    model = platform.processor() or ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"nproc": os.cpu_count(), "available": avail, "model": model}


def cpu_baseline_torch(U, b, geometry, kappa, shifts, gpu_iterations, iters):
    """The reference's op sequence restated in torch-CPU complex128
    (oracle/torch_cpu.py: project -> gather -> colour einsum -> rebuild ->
    add -> g5, dirac.py:458-501) driving multishift CG iterations with every
    shift active; timed with 4 intra-op threads (the reference's
    TfLeonardYMNDim.py:16) and with every CPU this process may use, on
    `iters` iterations after one warm-up iteration, extrapolated to the GPU
    solve's iteration count.  TensorFlow is not installed, so the reference
    itself cannot be timed."""
    import torch
    from oracle import torch_cpu as otc
    info = cpu_info()
    threads_all = max(1, min(int(os.environ.get("OMP_NUM_THREADS", info["available"])), info["available"]))
    t0 = time.time()
    op = otc.TorchWilson(U, geometry, kappa, True)
    bt = torch.from_numpy(b)
    res = {}
    for nt in sorted({4, threads_all}):
        torch.set_num_threads(nt)
        otc.cgm_iterations(op, bt, shifts, 1)
        tc = time.perf_counter()
        otc.cgm_iterations(op, bt, shifts, iters)
        per_it = (time.perf_counter() - tc) / iters
        res[nt] = per_it
        print(f"[bench] cpu baseline ({nt} threads): {per_it:.2f} s per CG-M iteration", file=sys.stderr, flush=True)
    print(f"[bench] cpu baseline total {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    best_nt = min(res, key=res.get)
    sample = (f"{iters} multishift CG iterations ({len(shifts)} shifts, all active) at {'x'.join(map(str, geometry))} "
              f"by a torch-CPU complex128 restatement of the reference's op sequence (oracle/torch_cpu.py; "
              f"TensorFlow unavailable offline), after one warm-up iteration, extrapolated to the "
              f"{gpu_iterations} iterations of the GPU solve")
    return {"value": 1.0 / (res[best_nt] * gpu_iterations), "unit": "solves/s", "cores": best_nt, "kind": "port",
            "sample": sample,
            "per_threads": {str(nt): {"s_per_iteration": res[nt], "solves_per_s": 1.0 / (res[nt] * gpu_iterations)}
                            for nt in res},
            "cpu": info}


# Kernel classes of lym_kernel_timing and the kernels they are in a rocprof trace
KERNEL_OF_CLASS = {
    "shift_update_f64": "k_fast_k3", "shift_update_f32": "k_fast_k3s<1>", "shift_update_f16": "k_fast_k3s<2>",
    "wilson_cgm_k1_f64": "k_wilson_tile<..., 0, EpiCgmK1<0, 12>, ...>",
    "wilson_cgm_k2f_f64": "k_wilson_tile<..., 0, EpiCgmK2F<0, 12>, ...>",
    "wilson_cgm_k1_f32": "k_wilson_tile<..., 1, EpiCgmK1<1, 12>, ...>",
    "wilson_cgm_k2f_f32": "k_wilson_tile<..., 1, EpiCgmK2F<1, 12>, ...>",
    "wilson_cgm_k1_f16": "k_wilson_tile<..., 2, EpiCgmK1<2, 12>, ...>",
    "wilson_cgm_k2f_f16": "k_wilson_tile<..., 2, EpiCgmK2F<2, 12>, ...>",
    "direction_update_f64": "k_fast_kp<0>", "direction_update_f32": "k_fast_kp<1>",
    "direction_update_f16": "k_fast_kp16<12>", "finalize": "k_finalize",
    "wilson_cg_k1_f16": "k_wilson_tile<..., 2, EpiCgmK1<2, 12>, ...>",
    "wilson_cg_k2x_f16": "k_wilson_tile<..., 2, EpiCgK2X<2, 12>, ...>",
    "wilson_cg_k1_f32": "k_wilson_tile<..., 1, EpiCgmK1<1, 12>, ...>",
    "wilson_cg_k2x_f32": "k_wilson_tile<..., 1, EpiCgK2X<1, 12>, ...>",
    "cg_direction_update_f16": "k_rel_kp<2, 12>", "cg_direction_update_f32": "k_rel_kp<1, 12>",
    "reliable_update": "k_ru_add + k_wilson_tile<..., 0, EpiRu1/EpiRu2H> + k_finalize (gated per RHS)",
}


def kernel_table(kt, kt_total, workload, timing_note):
    """Per-class durations of a kernel-timed solve (lym_kernel_timing) as the
    bench reports them, the roofline of the dominant class with algorithmic
    bytes, and the roofline of the dominant Wilson (dslash) class."""
    kernels = {k: dict(v, avg_us=v["us"] / max(1, v["launches"]),
                       gbs=(v["bytes"] / (v["us"] * 1e-6) / 1e9) if v["us"] > 0 and v["bytes"] > 0 else None,
                       share=v["us"] / (kt_total * 1e6))
               for k, v in kt.items()}

    def roof(cls):
        d = kt[cls]
        gbs = d["bytes"] / (d["us"] * 1e-6) / 1e9
        return {"kernel": f"{cls} ({KERNEL_OF_CLASS.get(cls, cls)})", "bound": "hbm", "achieved": gbs,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                "traffic": measured_class_traffic(workload, cls),
                "bytes_per_launch": d["bytes"] / d["launches"], "avg_us": d["us"] / d["launches"],
                "launches": d["launches"], "share_of_solve": d["us"] / (kt_total * 1e6), "timing": timing_note}
    with_bytes = [k for k in kt if kt[k]["bytes"] > 0]
    dom = max(with_bytes, key=lambda k: kt[k]["us"]) if with_bytes else None
    wil = [k for k in with_bytes if k.startswith("wilson")]
    dwil = max(wil, key=lambda k: kt[k]["us"]) if wil else None
    return kernels, (roof(dom) if dom else None), (roof(dwil) if dwil else None)


def measured_class_traffic(workload, cls):
    """HBM bytes per launch of kernel class `cls` in `workload` from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summaries (profiles/*traffic*.md),
    or None."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*traffic*.md"))):
        for line in open(path):
            m = re.search(r"traffic_bytes_per_launch (\S+) (\S+) = (\d+)", line)
            if m and m.group(1) == workload and m.group(2) == cls:
                return int(m.group(3))
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c5", choices=sorted(WORKLOADS) + sorted(BATCHED) + sorted(SERVING) + ["dslash"])
    ap.add_argument("--precision", default=None, choices=["mixed", "double", "half"],
                    help="Krylov storage (default: half for c5, mixed for the other multishift workloads; "
                         "batched workloads: their own)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-iters", type=int, default=2)
    ap.add_argument("--dslash-reps", type=int, default=50)
    ap.add_argument("--no-dslash-c3", action="store_true", help="skip the 24^3x48 dslash measurement")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched 28-source CG and the batched 8-source 12-shift multishift CG at 24^3x48 "
                         "reported beside the default workload")
    ap.add_argument("--dslash-geometry", default="24x24x24x48", help="lattice of --workload dslash")
    ap.add_argument("--no-autotune", action="store_true", help="keep the built-in Wilson tile rule")
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="skip the per-kernel HIP-event timing (and so the roofline) after the timed solves")
    ap.add_argument("--timing-steps", type=int, default=1,
                    help="solves timed per kernel with HIP events after the timed region (roofline, kernel table)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="end with 256 MiB device copies (the calibration kernel of rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--baseline-steps", type=int, default=2,
                    help="timed solves of the plain FP64 CG-M for the speed-up figure (0: skip)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); multi-GPU domain decomposition "
              "is not built in this version", file=sys.stderr)
        sys.exit(2)
    if args.workload in BATCHED:
        return run_batched(args)
    if args.workload == "dslash":
        return run_dslash(args)
    if args.precision is None:
        args.precision = "half" if args.workload == "c5" else "mixed"
    if args.workload in SERVING:
        return run_serving(args)

    import torch
    from tfleonardym_amd import _lib, dirac, dirac_solver, fields, lattice

    dev = _lib.require_device()
    W = WORKLOADS[args.workload]
    geometry, nc = W["geometry"], W["nc"]
    V = int(np.prod(geometry))
    t0 = time.time()
    U = fields.gauge_field(geometry, nc, seed=20240, kind="near_identity", eps=W["eps"])
    b = fields.gaussian_spinor(geometry, 4, nc, seed=20241)
    shifts = fields.log_shifts(W["lo"], W["hi"], W["nshift"])
    print(f"[bench] inputs generated in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    cfg = lattice.Configuration(U, geometry, nc, 4)
    op = dirac.DiracWilsonOperator(cfg, W["kappa"], True, representation="fundamental")
    tiles = None
    if not args.no_autotune:
        tiles = op.autotune()
        print(f"[bench] autotuned tiles: {tiles}", file=sys.stderr, flush=True)
    A = dirac.SquareOperator(op)
    bd = torch.from_numpy(b).to(dev)
    solver = dirac_solver.multishift_solver(20000, 0.0, algorithm="cg", tolerance=1e-10,
                                            precision_mode=args.precision, verbose=False)

    # ---- warmup + timed solves (inputs resident in HBM) ---------------------
    for i in range(args.warmup):
        solver.solve(A, bd, shifts)
        print(f"[bench] warmup {i} done: {solver.last_stats['iterations']} it, "
              f"{solver.last_stats['seconds']:.3f}s", file=sys.stderr, flush=True)
    timing = not args.no_kernel_timing and args.precision != "double"
    lat = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    stats = None
    for i in range(args.steps):
        ts = time.perf_counter()
        xs = solver.solve(A, bd, shifts)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
        stats = solver.last_stats
        if stats["status"] != "converged" or max(stats["true_rel_resid"]) > 1e-10:
            print(f"[bench] solve {i} did not reach 1e-10: {stats}", file=sys.stderr)
            sys.exit(1)
        print(f"[bench] step {i}: {lat[-1] * 1e3:.1f} ms, {stats['iterations']} it, refinement "
              f"{sum(stats['refine_cycles'])} solves / {stats['refine_iterations']} it", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    total = time.perf_counter() - t_start
    solves_per_s = args.steps / total
    # per-kernel durations: further solves of the same system with HIP events
    # around every launch on the solve stream.  Events need direct launches
    # (no graph replay), which adds launch gaps, so these solves are not part
    # of the timed region; the kernels' own durations are unaffected.
    kt = None
    kt_total = None
    if timing and args.timing_steps > 0:
        _lib.kernel_timing(True)
        tk = time.perf_counter()
        for _ in range(args.timing_steps):
            solver.solve(A, bd, shifts)
        torch.cuda.synchronize()
        kt_total = time.perf_counter() - tk
        _lib.kernel_timing(False)
        kt = _lib.kernel_timing_read()

    # ---- the plain FP64 multishift CG (textbook CG-M, one RHS) on the same solve
    plain = None
    if args.precision != "double" and args.baseline_steps > 0:
        ps = dirac_solver.multishift_solver(20000, 0.0, algorithm="cg", tolerance=1e-10, precision_mode="double",
                                            verbose=False)
        ps.solve(A, bd, shifts)
        torch.cuda.synchronize()
        tb = time.perf_counter()
        for _ in range(args.baseline_steps):
            ps.solve(A, bd, shifts)
        torch.cuda.synchronize()
        plain = args.baseline_steps / (time.perf_counter() - tb)
        print(f"[bench] plain FP64 CG-M: {plain:.3f} solves/s, {ps.last_stats['iterations']} it", file=sys.stderr,
              flush=True)
        del ps
    del xs

    # ---- roofline: the solve's dominant kernel class ------------------------
    roofline = None
    kernels = None
    if kt:
        kernels, roofline, _ = kernel_table(
            kt, kt_total, W["name"],
            f"HIP events on the solve stream around every launch of the class, over {args.timing_steps} further "
            f"solve(s) of the benchmarked system after the timed region (direct launches; the timed solves replay "
            f"HIP graphs)")
        print(f"[bench] dominant kernel {roofline['kernel']}: {roofline['launches']} launches, "
              f"{roofline['avg_us']:.1f} us avg, {roofline['achieved']:.0f} GB/s "
              f"({100 * roofline['frac']:.1f}% of {HBM_PEAK_GBS:.0f})", file=sys.stderr, flush=True)

    if args.calib_copy:
        calib_copy()

    # ---- Wilson kernel bandwidth (HIP events on the launch stream) ----------
    res = time_dslash(op, bd, args.dslash_reps)
    for name in ("f64", "f32"):
        print(f"[bench] wilson {name}: {res[name]['us']:.2f} us/launch, {res[name]['gbs']:.0f} GB/s", file=sys.stderr,
              flush=True)
    del op, A, solver
    torch.cuda.empty_cache()
    res_c3 = None
    if not args.no_dslash_c3:
        res_c3 = dslash_c3(args.dslash_reps, autotune=not args.no_autotune)

    # ---- the batched 28-source CG at 24^3x48 (16-bit storage) beside it ---
    res_b = None
    if not args.no_batched:
        res_b = batched_result("c3b", None, 1, 1, 2, autotune=args.no_dslash_c3 and not args.no_autotune,
                               timing=not args.no_kernel_timing)
    # ---- the batched multishift CG at 24^3x48 (8 sources x 12 shifts, the
    # mixed-precision batched multishift solve of the dslash target) ---------
    res_m = None
    if not args.no_batched:
        res_m = batched_result("c3m", None, 1, 1, 2, autotune=False, timing=not args.no_kernel_timing)

    # ---- CPU baseline -----------------------------------------------------
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_torch(U, b, geometry, W["kappa"], shifts, stats["iterations"], args.cpu_iters)

    out = {
        "metric": "multishift-CG solves/s at 1e-10 true resid",
        "value": solves_per_s,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"half": "f16 storage (int16 + per-site norm) + f32 + f64", "mixed": "f32+f64",
                  "double": "f64"}[args.precision],
        "data": "synthetic (seeded near-identity SU(3) links, Gaussian source)",
        "config": {"workload": W["name"], "geometry": list(geometry), "kappa": W["kappa"], "shifts": len(shifts),
                   "shift_range": [W["lo"], W["hi"]], "precision": args.precision, "tolerance": 1e-10},
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3), "p99": float(np.percentile(lat, 99) * 1e3)},
        "iterations": stats["iterations"],
        "precision_phase_reached": stats["precision_phase"],
        "refinement": {"solves": sum(stats["refine_cycles"]), "iterations": stats["refine_iterations"]},
        "plain_fp64_solves_per_s": plain,
        "speedup_vs_plain_fp64": (solves_per_s / plain) if plain else None,
        "max_true_rel_resid": max(stats["true_rel_resid"]),
        "true_rel_resid_per_shift": [float(v) for v in stats["true_rel_resid"]],
        "roofline": roofline,
        "kernels": kernels,
        "dslash_gbs": {"f64": res["f64"]["gbs"], "f32": res["f32"]["gbs"]},
        "dslash": res,
        "dslash_24c3x48": res_c3,
        "batched_24c3x48": res_b,
        "batched_multishift_24c3x48": res_m,
        "wilson_tiles": tiles,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
