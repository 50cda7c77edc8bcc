"""Benchmark: Newton timesteps/s of the device pipeline on the 2000-body mixed
pile at ~30k contacts per step.

A step is one full timestep on the device-resident scene: grid-SDF contact
generation, the primal-dual Newton solve (FP64 outer loop, FP32 block-Jacobi
PCG inner solve, FP64 merit line search) and integration.  The timed region
starts from the settled state of the seeded pile (the state after the
workload's 1000 steps), with every input already in HBM.

Prints one JSON line.  The CPU baseline is the FP64 restatement of the same
algorithm (oracle/), run on the host cores on a bounded sample of steps from
the same state.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from primal_dual_contacts_miror_amd import pdc, scenes, serve  # noqa: E402

METRIC = "Newton timesteps/sec at 30k contacts; batched scene-steps/sec + p50/p99 latency"
PEAK_HBM = 8.0e12
# kernel classes timed with HIP events (include/pdc.h PDC_KT_*)
KT_NAMES = {"pcg": pdc.KT_PCG, "pcg_per_launch": pdc.KT_SPMV, "contact_terms": pdc.KT_CONTACT_TERMS,
            "contacts": pdc.KT_CONTACTS}
KT_KERNEL = {"pcg": "k_pcg_iter (one PCG iteration: block-CSR SpMV with A P, block-Jacobi apply, vector "
                    "updates, partial dot products), timed by one event pair around each solve's launches, "
                    "queued %d at a time (the launches that find the solve converged included)",
             "pcg_per_launch": "k_pcg_iter, launched without graph replay, timed per launch",
             "contact_terms": "k_constraints + k_contact_res", "contacts": "k_pair_hits"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def settled_state(sc, name, settle_steps=None, path=None):
    if path is None:
        path = os.path.join(ROOT, "tests", "golden", "%s_settled.npz" % name)
    if settle_steps is None and os.path.exists(path):
        z = np.load(path)
        return (z["pos"], z["quat"], z["vel"], z["angvel"]), "fixture"
    n = sc.steps if settle_steps is None else settle_steps
    log("settling %s on the device (%d steps)" % (name, n))
    ds = pdc.DynamicSystem(sc)
    for k in range(n):
        ds.step(1)
        if k % 100 == 0:
            log("settle step", k)
    st = ds.state()
    ds.ctx.close()
    return st, "device-settled"


def cpu_baseline(sc, state, budget_s, threads, first_gpu=None):
    """CPU leg; with first_gpu (the device state one step after `state`) also
    the precision statement of that step against the FP64 restatement."""
    from oracle import orc
    # same parameters as the device run (tolerances, PCG branch and its tolerance), so both do the same work
    o = orc.OracleScene(sc, sc.params)
    o.set_threads(threads)
    o.set_state(*state)
    t0 = time.perf_counter()
    n = 0
    contacts = 0
    precision = None
    step_ms = []
    while True:
        ts = time.perf_counter()
        st = o.step(1)
        step_ms.append(1e3 * (time.perf_counter() - ts))
        n += 1
        contacts += st.n_contacts
        el = time.perf_counter() - t0
        if n == 1 and first_gpu is not None:
            t_pause = time.perf_counter()
            precision = step_precision(sc, first_gpu, o.get_state(), state, threads)
            t0 += time.perf_counter() - t_pause
        log("cpu baseline step %d: %d contacts, %.2fs" % (n, st.n_contacts, el))
        if el > budget_s or n >= 50:
            break
    return precision, {
        "value": n / el,
        "unit": "timesteps/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d timestep(s) of the FP64 CPU restatement (OpenMP) from the same settled state, same "
                  "parameters (%d contacts/step avg; block-Jacobi PCG branch, pcg_tol %g)" % (
                      n, contacts // n, sc.params.pcg_tol),
        "median_step_ms": float(np.median(step_ms)),
        "p99_step_ms": float(np.percentile(step_ms, 99)),
        "host": {"nproc": os.cpu_count(), "cpu_model": _cpu_model()},
    }


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def measured_copy_rate():
    """Device-to-device copy rate of 1 GiB on this GPU (read + write bytes per
    second), the achievable counterpart of the 8 TB/s HBM figure, through the
    HIP runtime the library itself is linked against."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    V = ctypes.c_void_p
    hip.hipMalloc.argtypes = [ctypes.POINTER(V), ctypes.c_size_t]
    hip.hipFree.argtypes = [V]
    hip.hipMemset.argtypes = [V, ctypes.c_int, ctypes.c_size_t]
    hip.hipMemcpyAsync.argtypes = [V, V, ctypes.c_size_t, ctypes.c_int, V]
    hip.hipEventCreate.argtypes = [ctypes.POINTER(V)]
    hip.hipEventRecord.argtypes = [V, V]
    hip.hipEventSynchronize.argtypes = [V]
    hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), V, V]
    hip.hipEventDestroy.argtypes = [V]
    n = 1 << 30
    a, b, e0, e1 = V(), V(), V(), V()
    if hip.hipMalloc(ctypes.byref(a), n) or hip.hipMalloc(ctypes.byref(b), n):
        raise RuntimeError("hipMalloc failed")
    try:
        hip.hipMemset(a, 0, n)
        hip.hipEventCreate(ctypes.byref(e0))
        hip.hipEventCreate(ctypes.byref(e1))
        best = 0.0
        for _ in range(6):
            hip.hipEventRecord(e0, None)
            hip.hipMemcpyAsync(b, a, n, 3, None)  # device to device
            hip.hipEventRecord(e1, None)
            hip.hipEventSynchronize(e1)
            ms = ctypes.c_float(0)
            hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
            if ms.value > 0:
                best = max(best, 2.0 * n / (ms.value * 1e-3))
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    return best


def batched_leg(n_scenes, steps, plain_scenes, plain_steps, seed_base=1000):
    """Many independent seeded random piles (10-500 bodies, log-uniform) in
    one context, all stepped together from their seeded drop; then, at the
    same (contact-rich) state, the batched rate over a few steps against the
    plain per-scene path (one scene per step call, one after another)."""
    scs = [scenes.random_pile(seed_base + k) for k in range(n_scenes)]
    bs = pdc.BatchedSystem(scs)
    log("batched: %d scenes, %d bodies, %d steps" % (n_scenes, sum(s.n for s in scs), steps))
    t0 = time.perf_counter()
    for k in range(steps):
        st = bs.step(1)
        if k % 50 == 0:
            log("batched step %d: %d contacts" % (k, sum(x.n_contacts for x in st)))
    el = time.perf_counter() - t0
    contacts = [x.n_contacts for x in st]
    tail = 10
    # the same tail of steps from the same state, with the default work
    # shapes and with the shapes tuned per size class at this state
    state = bs.ctx.get_state()
    t0 = time.perf_counter()
    for _ in range(tail):
        bs.step(1)
    el_default = time.perf_counter() - t0
    bs.ctx.set_state(*state)
    t0 = time.perf_counter()
    shapes = bs.ctx.autotune(2)
    el_tune = time.perf_counter() - t0
    log("autotune (%.1fs): %s" % (el_tune, shapes.as_dict()))
    t0 = time.perf_counter()
    for _ in range(tail):
        bs.step(1)
    el_tail = time.perf_counter() - t0
    t0 = time.perf_counter()
    for k in range(plain_scenes):
        for _ in range(plain_steps):
            bs.step(1, [k])
    el_plain = time.perf_counter() - t0
    batched_tail = n_scenes * tail / el_tail
    plain = plain_scenes * plain_steps / el_plain
    out = {
        "workload": "%d seeded random piles, 10-500 bodies (log-uniform), spheres + boxes, dt 0.005, 32^3 grids, "
                    "FP32 PCG / FP64 Newton, stepped together from the seeded drop" % n_scenes,
        "steps": steps,
        "scene_steps_per_s": n_scenes * steps / el,
        "contacts_last_step": {"total": int(sum(contacts)), "max_per_scene": int(max(contacts))},
        "autotune": {
            "shapes": shapes.as_dict(),
            "seconds": el_tune,
            "default_shapes_scene_steps_per_s": n_scenes * tail / el_default,
            "tuned_shapes_scene_steps_per_s": batched_tail,
        },
        "contact_phase": {
            "batched_scene_steps_per_s": batched_tail,
            "plain_scene_steps_per_s": plain,
            "speedup_vs_plain": batched_tail / plain,
            "sample": "%d batched steps of all scenes vs %d steps of each of %d scenes alone, same state" % (
                tail, plain_steps, plain_scenes),
        },
    }
    return bs, out


def served_leg(n_scenes, settle_steps, rates, duration, plain_rate, seed_base=5000, seed=7):
    """Continuous-batching server over resident scenes of its own: seeded
    random piles of 10-1000 bodies (log-uniform), admitted by HBM footprint,
    stepped together from their seeded drop for `settle_steps` steps so that
    the requests meet contact-rich states; then Poisson request streams of
    one-step requests at each rate, and the plain per-scene server (one
    request per step call) at `plain_rate`."""
    t0 = time.perf_counter()
    scs = [scenes.random_pile(seed_base + k, n_min=10, n_max=1000) for k in range(n_scenes)]
    srv = serve.SceneServer(scs, hbm_budget=200 * 2 ** 30)
    n = len(srv.scenes)
    bodies = sum(sc.n for sc in srv.scenes)
    log("served: %d scenes built and resident (%d rejected), %d bodies, %.1f GiB, %.1fs" % (
        n, len(srv.rejected), bodies, srv.footprint / 2 ** 30, time.perf_counter() - t0))
    t0 = time.perf_counter()
    st = []
    for k in range(settle_steps):
        st = srv.system.step(1)
        if k % 50 == 0:
            log("served settle step %d: %d contacts" % (k, sum(x.n_contacts for x in st)))
    settle_s = time.perf_counter() - t0
    out = {"workload": "%d seeded random piles of 10-1000 bodies (log-uniform), spheres + boxes, 32^3 grids, "
                       "FP32 PCG / FP64 Newton, admitted by HBM footprint, %d steps from the seeded drop before "
                       "the requests" % (n_scenes, settle_steps),
           "resident_scenes": n, "rejected_scenes": len(srv.rejected), "bodies": bodies,
           "footprint_gib": srv.footprint / 2 ** 30,
           "contacts_resident": int(sum(x.n_contacts for x in st)) if st else 0,
           "settle_s": settle_s, "duration_s": duration, "rates": []}
    for rate in rates:
        reqs = serve.poisson_requests(rate, duration, n, seed=seed)
        served = srv.run(reqs, batched=True, time_limit=duration * 4)
        summ = serve.latency_summary(served)
        summ["rate_req_per_s"] = rate
        summ["dropped"] = len(reqs) - summ["requests"]
        log("served %g req/s: %s" % (rate, summ))
        out["rates"].append(summ)
    reqs = serve.poisson_requests(plain_rate, duration, n, seed=seed)
    served = srv.run(reqs, batched=False, time_limit=duration * 4)
    summ = serve.latency_summary(served)
    summ["rate_req_per_s"] = plain_rate
    summ["dropped"] = len(reqs) - summ["requests"]
    log("plain %g req/s: %s" % (plain_rate, summ))
    out["plain_per_scene"] = summ
    srv.close()
    return out


def _state_errors(sc, a, b):
    dyn = sc.is_static == 0
    size = np.max(sc.dims, axis=1)[dyn]
    ep = np.linalg.norm(a[0][dyn] - b[0][dyn], axis=1) / size
    ea = 2 * np.arccos(np.clip(np.abs(np.sum(a[1][dyn] * b[1][dyn], axis=1)), 0, 1))
    gdt = float(np.linalg.norm(sc.params.gravity[:])) * sc.params.timestep
    ev = np.linalg.norm(a[2][dyn] - b[2][dyn], axis=1) / gdt
    return {"max_position_error_rel_body_size": float(ep.max()), "max_angle_error_rad": float(ea.max()),
            "max_velocity_error_rel_g_dt": float(ev.max())}


def step_precision(sc, g, o, state, threads):
    """Errors of one device timestep (FP32 PCG inside the FP64 Newton loop)
    against the FP64 CPU restatement from the same state: positions relative
    to the body size, orientations in radians, velocities relative to |g| dt.
    Next to them, the same errors between two runs of the restatement itself
    that differ only in the inner solve's tolerance (the workload's and
    1e-12): at the workload's Newton tolerance the step is defined only to
    that level.  Per-contact forces are not compared: bodies resting on
    several coplanar contacts have a nonunique force split (only the body
    impulses are fixed)."""
    from oracle import orc
    q = type(sc.params).from_buffer_copy(sc.params)
    q.pcg_tol = 1e-12
    o2 = orc.OracleScene(sc, q)
    o2.set_threads(threads)
    o2.set_state(*state)
    o2.step(1)
    out = {"steps_from_identical_state": 1}
    out.update(_state_errors(sc, g, o))
    out["restatement_self_difference"] = _state_errors(sc, o, o2.get_state())
    out["reference"] = ("FP64 CPU restatement, same parameters (Newton tolerance %g, inner PCG to %g), "
                        "same settled state; self difference: the restatement with its inner PCG to 1e-12"
                        % (sc.params.tolerance, sc.params.pcg_tol))
    return out


def measured_traffic(kernel, workload):
    """HBM bytes per launch of the timed kernel from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (profiles/<kernel>_<workload>_traffic.json), or None when there is none."""
    if kernel != "pcg":
        return None
    path = os.path.join(ROOT, "profiles", "pcg_iter_%s_traffic.json" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return float(json.load(f)["traffic_bytes_per_launch"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="pile_2000")
    ap.add_argument("--kernel", default="pcg", choices=sorted(KT_NAMES))
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--state", default=None, help="npz with pos, quat, vel, angvel to start from")
    ap.add_argument("--settle-steps", type=int, default=None,
                    help="settle on the device for this many steps instead of loading the settled state")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched and served legs")
    ap.add_argument("--batched-scenes", type=int, default=256)
    ap.add_argument("--batched-steps", type=int, default=200)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--served-rates", default="100,200,500,1000,2000,5000")
    ap.add_argument("--served-duration", type=float, default=2.0)
    ap.add_argument("--served-scenes", type=int, default=1024)
    ap.add_argument("--served-settle", type=int, default=150)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(scene-level sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    sc = scenes.WORKLOADS[args.workload]()
    state, origin = settled_state(sc, args.workload, args.settle_steps, args.state)
    ds = pdc.DynamicSystem(sc)
    ds.set_state(*state)
    log("warmup %d steps" % args.warmup)
    first_gpu = None
    for k in range(args.warmup):
        st = ds.step(1)
        if k == 0:
            first_gpu = tuple(x.copy() for x in ds.state())
        log("warmup: %d contacts, %d Newton it, %d PCG it" % (st.n_contacts, st.newton_iterations,
                                                              st.pcg_iterations))
    ds.ctx.enable_kernel_timing(KT_NAMES[args.kernel])
    ds.ctx.reset_timings()
    contacts = []
    newton = []
    pcg_it = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        st = ds.step(1)
        contacts.append(st.n_contacts)
        newton.append(st.newton_iterations)
        pcg_it.append(st.pcg_iterations)
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms, launches, nbytes = ds.ctx.kernel_timing()
    log("timed %d steps in %.3fs: contacts %d..%d, Newton it %.1f, PCG it %.1f" % (
        args.steps, elapsed, min(contacts), max(contacts), np.mean(newton), np.mean(pcg_it)))
    achieved = nbytes / (ms * 1e-3) if ms > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (Newton, residuals, line search) + f32 (PCG inner solve)",
        "data": "synthetic: seeded 2000-body mixed pile, %s state after %d steps" % (origin, sc.steps),
        "config": {
            "workload": "pile_2000: 1000 spheres + 1000 boxes in a 20x20x40 container, "
                        "96^3 SDF grids, mu 0.6, dt 0.005, settled",
            "contacts_per_step": float(np.mean(contacts)),
            "newton_iterations_per_step": float(np.mean(newton)),
            "pcg_iterations_per_step": float(np.mean(pcg_it)),
            "bodies": int(sc.n),
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": KT_KERNEL[args.kernel] % (ds.ctx.get_shapes().pcg_chunk or 8) if args.kernel == "pcg"
            else KT_KERNEL[args.kernel],
            "achieved": achieved / 1e9,
            "peak": PEAK_HBM / 1e9,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM,
            "traffic": measured_traffic(args.kernel, args.workload),
            "launches": launches,
            "avg_launch_us": 1e3 * ms / max(launches, 1),
            "algorithmic_bytes_per_launch": nbytes / max(launches, 1),
        },
    }
    try:
        out["roofline"]["measured_copy_GBps"] = measured_copy_rate() / 1e9
    except Exception as exc:  # the copy is a reference figure only
        log("copy rate not measured: %s" % exc)
    ds.ctx.close()
    if not args.no_batched:
        bs, out["batched"] = batched_leg(args.batched_scenes, args.batched_steps, 16, 10)
        bs.ctx.close()
        if not args.no_served:
            rates = [float(r) for r in args.served_rates.split(",") if r]
            out["served"] = served_leg(args.served_scenes, args.served_settle, rates, args.served_duration,
                                       plain_rate=40.0)
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        precision, out["cpu_baseline"] = cpu_baseline(sc, state, args.cpu_budget, threads, first_gpu)
        if precision is not None:
            out["precision"] = precision
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
