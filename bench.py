#!/usr/bin/env python3
"""Benchmark of the isolated narrow-phase workload on one MI355X.

Workload: 2^24 unique triangle pairs, vertices uniform in a 1 mm cube,
separation log-uniform in [1e-7, 1e-2] m, 5% near-parallel edge pairs
(angle < 1e-4 rad), 1% intersecting; generated in HBM with a fixed seed.
A step is one pass of the narrow phase over all pairs.  The timed region
brackets exactly --steps passes with device synchronisation and HIP events on
the stream the kernels run on.  Inputs are resident in HBM before timing.

Prints ONE JSON line.  The CPU baseline (the C restatement of the reference,
OpenMP over pairs, on this host's cores) is timed on a bounded sample of the
same pairs, and the GPU results of that sample are checked against it.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6


def host_cpu(threads):
    """CPU model, logical CPUs, physical cores (distinct physical/core id
    pairs of /proc/cpuinfo) and the threads a baseline used."""
    model, cores = "unknown", set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name":
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(), "physical_cores": len(cores) or None,
            "threads_used": threads}


def omp_threads():
    """Threads of the OpenMP baselines: OMP_NUM_THREADS (the GPU host's CPU
    share), else every logical CPU."""
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--pairs-log2", type=int, default=24)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="target wall time of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dem", action="store_true",
                    help="leave out the DEM sections of the default run (C1, funnel, drum)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default launch parameters (every workload)")
    ap.add_argument("--tuning", default="",
                    help="fixed launch parameters instead of autotuning (every workload), "
                         "e.g. learned_wg_per_cu=16,exact_block=256 (profiling runs)")
    ap.add_argument("--methods", default="bf64,bf64_exact,penalty64,hybrid64,mixed,"
                                         "penalty_cert,learned_penalty,learned")
    ap.add_argument("--workload", default="pairs",
                    choices=["pairs", "c1-dem", "hopper", "funnel", "drum"],
                    help="pairs: 2^24 unique triangle pairs (the headline); c1-dem: the "
                         "64-icosphere DEM step (particle-steps/s); hopper: the reference's "
                         "hopper scene (input/hopper.py, 15,625 spheres as 80-triangle "
                         "icospheres in 26 wall prisms); funnel: 10^4 uneven particles "
                         "through a 2,048-triangle cone; drum: 2*10^5 particles in a "
                         "rotating box (both as captured steps)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 256 MiB device-to-device copy kernel three times "
                         "(reference for scaling HBM byte counters) and exit")
    ap.add_argument("--train-pairs-log2", type=int, default=24)
    ap.add_argument("--train-epochs", type=int, default=20)
    return ap.parse_args()


def time_method(batch, method, steps, warmup):
    from delta_amd import _lib
    L = _lib.lib()
    for _ in range(warmup):
        batch.run(method)
    _lib.check(L.dh_synchronize(), "sync")
    e0, e1 = C.c_void_p(), C.c_void_p()
    _lib.check(L.dh_event_create(C.byref(e0)), "ev")
    _lib.check(L.dh_event_create(C.byref(e1)), "ev")
    t0 = time.perf_counter()
    _lib.check(L.dh_event_record(e0, None), "rec")
    for _ in range(steps):
        batch.run(method)
    _lib.check(L.dh_event_record(e1, None), "rec")
    _lib.check(L.dh_synchronize(), "sync")
    wall = time.perf_counter() - t0
    ms = C.c_float()
    _lib.check(L.dh_event_elapsed_ms(C.byref(ms), e0, e1), "elapsed")
    L.dh_event_destroy(e0)
    L.dh_event_destroy(e1)
    return ms.value / steps, wall


def train_network(a):
    """Train the feature network on pairs of the same distribution drawn with
    another seed (untimed setup)."""
    from delta_amd import learn, narrow
    t0 = time.perf_counter()
    tb = narrow.PairBatch(1 << a.train_pairs_log2, seed=101)
    ds = learn.Dataset(tb.soa, tb.n)
    tr = learn.Trainer(seed=0)
    losses = tr.fit(ds, epochs=a.train_epochs, batch=8192, lr=3e-3)
    top1, top2 = tr.accuracy(ds, min(ds.n, 1 << 18))
    tr.install()
    info = {"train_pairs": tb.n, "epochs": a.train_epochs, "final_loss": losses[-1],
            "top1": float(top1), "top2": float(top2), "seconds": time.perf_counter() - t0}
    print(f"[bench] trained network: {info}", file=sys.stderr, flush=True)
    return tr, ds, info


def autotune(batch, methods):
    """Launch parameters for this device and workload, chosen before the
    timed region by the library's autotuner on the batch itself (one knob per
    path: the FP64 block size from bf64, the persistent grids from mixed and
    learned); every candidate gives identical results."""
    from delta_amd import tuning
    trials = {}
    for m in ("bf64", "mixed", "learned"):
        if m not in methods:
            continue
        try:
            trials[m] = [(t["value"], round(t["ms"], 4)) for t in tuning.autotune_pairs(batch, m)]
        except Exception as ex:
            print(f"[bench] autotune {m}: {ex}", file=sys.stderr, flush=True)
    table = tuning.get()
    print(f"[bench] tuning {table}", file=sys.stderr, flush=True)
    return {"table": table, "trials_value_ms": trials}


def contact_errors(P0, Q0, d0, P, Q, d, margin=2e-4):
    """Contact-record errors (contactpoint.cpp:116-155) over the pairs that
    are contacts in both: point (P+Q)/2, normal (P-Q)/|P-Q|, depth."""
    contact = (d0 <= margin) & (d <= margin)
    if not contact.any():
        return {"contacts": 0, "max_contact_point_error_m": 0.0, "max_normal_error": 0.0,
                "max_depth_error_m": 0.0}
    x0, x1 = (P0 + Q0) / 2, (P + Q) / 2
    n0 = (P0 - Q0) / np.where(d0 > 0, d0, 1.0)
    n1 = (P - Q) / np.where(d > 0, d, 1.0)
    return {"contacts": int(contact.sum()),
            "max_contact_point_error_m": float(np.abs(x1 - x0)[:, contact].max()),
            "max_normal_error": float(np.abs(n1 - n0)[:, contact].max()),
            "max_depth_error_m": float(np.abs(d - d0)[contact].max())}


def learned_accuracy(batch, method="learned"):
    """Learned results against the FP64 brute force over every pair."""
    import ctypes as C
    from delta_amd import _lib
    batch.run("bf64")
    P0, Q0, d0, _ = batch.results()
    batch.run(method)
    P, Q, d, fl = batch.results()
    cnt = _lib.DhPairCounters()
    _lib.check(_lib.lib().dh_tri_pairs_counters(C.byref(cnt)), "counters")
    rel = np.abs(d - d0) / np.where(d0 > 0, d0, 1.0)
    margin = 2e-4
    band = np.abs(d0 - margin) <= 1e-9
    mism = int(np.count_nonzero((d <= margin)[~band] != (d0 <= margin)[~band]))
    cert = (fl & 8) != 0
    out = {"pairs": int(batch.n), "certified_fraction": float(cert.mean()),
           "brute_force_fraction": float(1 - cert.mean()),
           "features_evaluated_per_pair": cnt.features_fp64 / batch.n,
           "max_rel_distance_error": float(rel.max()),
           "contact_set_mismatches_outside_1e-9_band": mism,
           "pairs_in_band": int(band.sum())}
    out.update(contact_errors(P0, Q0, d0, P, Q, d, margin))
    out["bitwise_equal_fraction"] = float(np.mean(d.view(np.uint64) == d0.view(np.uint64)))
    out["pq_bitwise_equal_fraction"] = float(np.mean(
        np.all(P.view(np.uint64) == P0.view(np.uint64), 0) &
        np.all(Q.view(np.uint64) == Q0.view(np.uint64), 0)))
    return out


def warm_start_report(batch):
    """Penalty Newton from the network's barycentric outputs against the same
    from the reference's 0.33 start: Newton steps per pair and certified
    fraction (the rest goes to the brute force)."""
    from delta_amd import _lib
    out = {}
    for m in ("learned_penalty", "penalty_cert"):
        batch.run(m)
        cnt = _lib.DhPairCounters()
        _lib.check(_lib.lib().dh_tri_pairs_counters(C.byref(cnt)), "counters")
        out[m] = {"newton_steps_per_pair": cnt.features_fp64 / batch.n,
                  "certified_fraction": cnt.certified / batch.n}
    return out


def mixed_accuracy(batch):
    """FP32 screen + FP64 recheck against the FP64 brute force."""
    from delta_amd import _lib
    batch.run("bf64")
    P0, Q0, d0, _ = batch.results()
    batch.run("mixed")
    P, Q, d, fl = batch.results()
    re = (fl & 2) != 0
    rel = np.abs(d - d0) / np.where(d0 > 0, d0, 1.0)
    margin = 2e-4
    band = np.abs(d0 - margin) <= 1e-9
    out = {"rechecked_fraction": float(re.mean()),
           "max_rel_distance_error": float(rel.max()),
           "max_rel_distance_error_fp32_pairs": float(rel[~re].max()) if (~re).any() else 0.0,
           "pairs_above_1e-6_rel_error": int(np.count_nonzero(rel > 1e-6)),
           "contact_set_mismatches_outside_1e-9_band":
               int(np.count_nonzero((d <= margin)[~band] != (d0 <= margin)[~band]))}
    out.update(contact_errors(P0, Q0, d0, P, Q, d, margin))
    return out


def cpu_baseline(batch, target_s):
    """Time the oracle (OpenMP over pairs) on a bounded prefix of the pairs and
    check the GPU's bf64 results on that prefix bit for bit."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    path = os.path.join(ROOT, "oracle", "liboracle_omp.so")
    if not os.path.exists(path):
        return None, None
    orc = C.CDLL(path)
    orc.or_tri_pairs_omp.argtypes = [C.c_int, C.c_void_p, C.c_int64, C.c_double, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    threads = omp_threads()
    soa_all = batch.soa.download()
    batch.run("bf64")
    P_g, Q_g, d_g, _ = batch.results()

    def run(m):
        soa = np.ascontiguousarray(soa_all[:, :m])
        P = np.empty((3, m))
        Q = np.empty((3, m))
        d = np.empty(m)
        t0 = time.perf_counter()
        orc.or_tri_pairs_omp(0, soa.ctypes.data, m, 2e-4, P.ctypes.data, Q.ctypes.data,
                             d.ctypes.data, None, threads)
        return time.perf_counter() - t0, P, Q, d

    m = min(batch.n, 1 << 16)
    run(m)  # warm-up
    dt, _, _, _ = run(m)
    rate = m / dt
    m = int(min(batch.n, max(m, rate * target_s)))
    dt, P, Q, d = run(m)
    mism = int(np.count_nonzero(d.view(np.uint64) != d_g[:m].view(np.uint64)) +
               np.count_nonzero(P.view(np.uint64) != P_g[:, :m].view(np.uint64)) +
               np.count_nonzero(Q.view(np.uint64) != Q_g[:, :m].view(np.uint64)))
    return ({"value": m / dt, "unit": "triangle-pair checks/s", "cores": threads,
             "kind": "port", "cpu": host_cpu(threads),
             "sample": f"first {m} of the {batch.n} pairs, FP64 brute force (oracle bfSolver, "
                       f"OpenMP over pairs, {threads} threads), {dt:.1f} s; the unique pairs "
                       "have no particle pairs, so the reference's row-parallel driver does not "
                       "apply here"},
            {"pairs": m, "bitwise_mismatches": mism})


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (scaled by a copy kernel of known size),
    or None."""
    path = os.path.join(ROOT, "profiles", "k_learned_tri_pairs_2p24_hbm_bytes.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            return int(f[2])
    return None


def copy_calibration():
    """A plain copy of known size (256 MiB read + 256 MiB written per
    launch) for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    x = torch.ones(1 << 28, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": 1 << 28, "launches": 3}), flush=True)


def dem_c1(a, steps=2000, learned=False):
    """64 jittered 80-triangle icospheres on a 4x4x4 lattice (1.05 mm) above a
    2-triangle floor, FP64 brute force, spring-dashpot k=1e5 N/m, damping 0.1,
    margin 1e-4 m, dt = 1e-6 s, `steps` (2000) steps from the initial state:
    eager (dh_step) and captured (one graph replay per step), each on its own
    copy of the scene after `--warmup` steps on a third copy; the two final
    states must be equal bit for bit.  Unless --no-autotune the detection
    tiling is tuned on the initial geometry first.  The lattice's initial
    overlaps fling the particles apart within about eight steps; most of the
    2000 steps then carry no contact."""
    from delta_amd import _lib, engine
    hs = engine.c1_scene()
    prm = engine.step_params(method=_lib.DH_BF, dt=1e-6, spring=1e5, damper=0.1, friction=0.0,
                             cell=hs.cell_size())
    L = _lib.lib()
    warm = engine.DeviceScene(hs)
    tuned = None if a.no_autotune or a.tuning else tune_detection(warm, hs, prm, L, 1_000_000)
    max_pairs = max_contacts = 0
    for _ in range(max(a.warmup, 1)):
        info = engine.step(warm, prm)
    _lib.check(L.dh_synchronize(), "sync")
    del warm
    ds = engine.DeviceScene(hs)
    _lib.check(L.dh_synchronize(), "sync")
    checks = contacts = pairs = culled = 0
    with_contacts = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        info = engine.step(ds, prm)
        checks += info.triangle_checks
        culled += info.pairs_culled
        contacts += info.contacts
        pairs += info.candidate_pairs
        with_contacts += 1 if info.contacts else 0
        max_pairs = max(max_pairs, info.candidate_pairs)
        max_contacts = max(max_contacts, info.contacts)
    _lib.check(L.dh_synchronize(), "sync")
    edt = time.perf_counter() - t0
    n = hs.n
    # the same steps as one captured graph per step (no host synchronisation
    # between steps)
    dg = engine.DeviceScene(hs)
    g = engine.StepGraph(dg, prm, pair_cap=2 * max_pairs + 64, contact_cap=2 * max_contacts + 1024)
    try:
        _lib.check(L.dh_synchronize(), "sync")
        t0 = time.perf_counter()
        for _ in range(steps):
            g.launch()
        ginfo, overflow = g.status()
        gdt = time.perf_counter() - t0
    finally:
        g.close()
    same = all(np.array_equal(ds.download(f).view(np.uint64), dg.download(f).view(np.uint64))
               for f in ("x", "y", "z", "pos", "linvel", "angvel", "rot"))
    out = {"metric": "particle-steps/s (64-icosphere DEM step, FP64 brute force, captured step)",
           "value": n * steps / gdt, "unit": "particle-steps/s", "n_gpus": 1, "steps": steps,
           "warmup": a.warmup, "ms_per_step": gdt * 1e3 / steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "64 jittered 80-triangle icospheres, 4x4x4 lattice at 1.05 mm, "
                                  "2-triangle floor, dt 1e-6 s, 2000 steps, FP64 brute force",
                      "particles": n},
           "overflow": overflow, "state_equals_eager": bool(same),
           "timing": "host clock from the first launch to the stream's end",
           "eager": {"ms_per_step": edt * 1e3 / steps, "value": n * steps / edt},
           "triangle_checks_per_s": checks / gdt, "culled_fraction": culled / max(checks, 1),
           "mean_candidate_pairs": pairs / steps, "mean_contacts": contacts / steps,
           "steps_with_contacts": with_contacts, "max_contacts": max_contacts,
           "update_and_binning": hopper_hbm(hs, dg, prm, L)}
    if tuned:
        out["tuning"] = {"table": tuned[0], "trials_value_ms": tuned[1]}
    if learned:
        # the same 2000 steps with the survivor queue through the learned
        # narrow phase (network installed by the C4 section): the final
        # state must equal the brute-force run's bit for bit
        pl = engine.step_params(method=_lib.DH_BF_LEARNED, dt=1e-6, spring=1e5, damper=0.1,
                                friction=0.0, cell=hs.cell_size())
        dl = engine.DeviceScene(hs)
        _lib.check(L.dh_synchronize(), "sync")
        t0 = time.perf_counter()
        for _ in range(steps):
            engine.step(dl, pl)
        _lib.check(L.dh_synchronize(), "sync")
        ldt = time.perf_counter() - t0
        dlg = engine.DeviceScene(hs)
        g = engine.StepGraph(dlg, pl, pair_cap=2 * max_pairs + 64,
                             contact_cap=2 * max_contacts + 1024)
        try:
            _lib.check(L.dh_synchronize(), "sync")
            t0 = time.perf_counter()
            for _ in range(steps):
                g.launch()
            _, lov = g.status()
            lgdt = time.perf_counter() - t0
        finally:
            g.close()
        lsame = all(np.array_equal(ds.download(f).view(np.uint64), x.download(f).view(np.uint64))
                    for f in ("x", "y", "z", "pos", "linvel", "angvel", "rot") for x in (dl, dlg))
        out["learned_survivor_queue"] = {
            "ms_per_step": lgdt * 1e3 / steps, "value": n * steps / lgdt,
            "eager_ms_per_step": ldt * 1e3 / steps, "overflow": lov,
            "state_equals_brute_force": bool(lsame)}
    if not a.no_cpu:
        # the reference-shaped OpenMP step (rows of each particle pair in
        # parallel, bf.cpp:92-95,137-141) from the same initial state, over
        # the first steps that fit the sample time
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as ol
        threads = omp_threads()
        hc = engine.c1_scene()
        k = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < a.cpu_seconds and k < steps:
            ol.or_step(hc, _lib.DH_BF, 1e5, 0.1, 0.0, 1e-6, (0.0, -9.81, 0.0),
                       omp_threads=threads)
            k += 1
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n * k / dt, "unit": "particle-steps/s", "cores": threads,
                               "kind": "port", "cpu": host_cpu(threads),
                               "sample": f"the first {k} of the {steps} steps from the same "
                                         f"initial state (all the contact-carrying ones), "
                                         f"oracle step with OpenMP over the triangles of A in "
                                         f"each particle pair and a critical append (then sorted into row order), "
                                         f"{threads} threads, {dt:.1f} s"}
    return out


def cpu_rows_baseline(hs, pairs, margin_each, seconds):
    """The reference-shaped OpenMP brute force (or_detect_rows_omp: rows of A
    in parallel, contact count under a critical section, bf.cpp:92-95,
    137-141) over the given particle pairs until `seconds` have passed:
    (triangle checks/s, pairs done, checks, seconds, threads)."""
    path = os.path.join(ROOT, "oracle", "liboracle_omp.so")
    orc = C.CDLL(path)
    f = orc.or_detect_rows_omp
    f.restype = C.c_int64
    f.argtypes = [C.c_void_p] * 3 + [C.c_int] + [C.c_void_p] * 3 + [C.c_int, C.c_double, C.c_int]
    threads = omp_threads()
    off = hs.tri_offset
    x, y, z = (np.ascontiguousarray(v, dtype=np.float64) for v in (hs.x, hs.y, hs.z))
    ptr = lambda v, i: v.ctypes.data + 8 * 3 * int(off[i])
    t0 = time.perf_counter()
    k = nchk = 0
    while time.perf_counter() - t0 < seconds and k < len(pairs):
        i, j = pairs[k]
        ni, nj = int(off[i + 1] - off[i]), int(off[j + 1] - off[j])
        f(ptr(x, i), ptr(y, i), ptr(z, i), ni, ptr(x, j), ptr(y, j), ptr(z, j), nj,
          2 * margin_each, threads)
        nchk += ni * nj
        k += 1
    dt = time.perf_counter() - t0
    return nchk / dt, k, nchk, dt, threads


def hopper_hbm(hs, ds, prm, L, reps=20):
    """HBM rates of binning (dh_broadphase: particle boxes from every vertex
    replica, then the grid) and of the fused update (dh_integrate on every
    particle, zero forces), each timed with HIP events over `reps` calls.
    Algorithmic bytes: binning 24 B per vertex replica read plus 64 B per
    particle (offsets, margin, box); update 48 B per vertex replica
    (referential read, spatial write) plus 560 B of particle state."""
    from delta_amd import _lib
    from delta_amd.device import DeviceArray
    nv = 3 * int(hs.tri_offset[-1])
    n = hs.n
    e0, e1 = C.c_void_p(), C.c_void_p()
    L.dh_event_create(C.byref(e0))
    L.dh_event_create(C.byref(e1))
    ms = C.c_float()
    cap = 16 * (n + 16)
    pairs = DeviceArray((2 * cap,), np.int32)
    npairs = C.c_int(0)
    L.dh_broadphase.argtypes = [C.POINTER(_lib.DhGeometry), C.c_void_p, C.c_double, C.c_void_p,
                                C.c_int, C.POINTER(C.c_int), C.c_void_p]
    geom = C.byref(ds.scene.geom)
    L.dh_broadphase(geom, ds.buf["obstacle"].ptr, prm.cell, pairs.ptr, cap, C.byref(npairs), None)
    L.dh_event_record(e0, None)
    for _ in range(reps):
        L.dh_broadphase(geom, ds.buf["obstacle"].ptr, prm.cell, pairs.ptr, cap, C.byref(npairs),
                        None)
    L.dh_event_record(e1, None)
    _lib.check(L.dh_synchronize(), "sync")
    L.dh_event_elapsed_ms(C.byref(ms), e0, e1)
    t_bin = ms.value * 1e-3 / reps
    zero = DeviceArray.from_host(np.zeros(3 * n))
    g = (C.c_double * 3)(0.0, 0.0, -10.0)
    L.dh_integrate.argtypes = [C.POINTER(_lib.DhBodies), C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_double,
                               C.POINTER(C.c_double), C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    sc = ds.scene
    args = (C.byref(sc.bodies), zero.ptr, zero.ptr, None, 1e-9, g, sc.x, sc.y, sc.z, sc.refx,
            sc.refy, sc.refz, sc.geom.tri_offset, None)
    L.dh_integrate(*args)
    L.dh_event_record(e0, None)
    for _ in range(reps):
        L.dh_integrate(*args)
    L.dh_event_record(e1, None)
    _lib.check(L.dh_synchronize(), "sync")
    L.dh_event_elapsed_ms(C.byref(ms), e0, e1)
    t_int = ms.value * 1e-3 / reps
    b_bin = 24.0 * nv + 64.0 * n
    b_int = 48.0 * nv + 560.0 * n
    return {"binning_GBps": b_bin / t_bin / 1e9, "binning_ms": t_bin * 1e3,
            "binning_bytes": b_bin, "update_GBps": b_int / t_int / 1e9, "update_ms": t_int * 1e3,
            "update_bytes": b_int, "peak_GBps": 8000.0,
            "note": "binning time includes the grid passes and the pair-count read-back"}


def tune_detection(ds, hs, prm, L, contact_cap):
    """Detection launch parameters autotuned on the scene's current
    candidate pairs (the broad phase run once)."""
    from delta_amd import _lib, tuning
    from delta_amd.device import DeviceArray
    cap = 16 * (hs.n + 16)
    d_pairs = DeviceArray((2 * cap,), np.int32)
    npairs = C.c_int(0)
    L.dh_broadphase.argtypes = [C.POINTER(_lib.DhGeometry), C.c_void_p, C.c_double, C.c_void_p,
                                C.c_int, C.POINTER(C.c_int), C.c_void_p]
    _lib.check(L.dh_broadphase(C.byref(ds.scene.geom), ds.buf["obstacle"].ptr, prm.cell,
                               d_pairs.ptr, cap, C.byref(npairs), None), "dh_broadphase")
    if npairs.value <= 0:
        # a scene whose particles have all separated has nothing to tune on
        raise RuntimeError("tune_detection: the scene has no candidate pairs "
                           "(tune on a geometry with contacts)")
    out_c = DeviceArray((contact_cap * 144,), np.uint8)
    # the median of 7 timings per value: single timings of these short
    # launches vary by more than the settings differ
    trials = tuning.autotune_detect(ds.scene.geom, d_pairs.ptr, npairs.value, out_c.ptr,
                                    contact_cap, reps=7)
    table = tuning.get()
    print(f"[bench] tuning {table}", file=sys.stderr, flush=True)
    return table, [(t["knob"], t["value"], round(t["ms"], 4)) for t in trials]


def detection_defaults():
    """The current table with the detection knobs at the library's defaults
    (an earlier section of the run may have tuned them)."""
    from delta_amd import tuning
    t = tuning.get()
    t.update({k: v for k, v in tuning.default().items() if k.startswith("detect_")})
    return t


def keep_faster_table(ds, prm, tuned, default, pair_cap, contact_cap, replays=40, rounds=3):
    """The detection trials time the detection alone, where the settings
    differ by less than the trials' spread; on the whole captured step their
    pick can lose to the library's default table.  Both tables are captured
    into step graphs and replayed `replays` times, alternately, in `rounds`
    rounds (the scene advances over these replays); the table with the lower
    median round time is set and returned with the two medians (ms per step)."""
    from delta_amd import engine, tuning
    tables = {"tuned": tuned, "default": default}
    times = {k: [] for k in tables}
    for _ in range(rounds):
        for k, t in tables.items():
            tuning.set(**t)
            g = engine.StepGraph(ds, prm, pair_cap=pair_cap, contact_cap=contact_cap)
            try:
                g.launch()
                g.status()
                t0 = time.perf_counter()
                for _ in range(replays):
                    g.launch()
                g.status()
                times[k].append((time.perf_counter() - t0) * 1e3 / replays)
            finally:
                g.close()
    med = {k: sorted(v)[len(v) // 2] for k, v in times.items()}
    kept = min(med, key=med.get)
    tuning.set(**tables[kept])
    print(f"[bench] whole-step check: {med}, kept {kept}", file=sys.stderr, flush=True)
    return tables[kept], {"kept": kept, "median_ms_per_step": {k: round(v, 4) for k, v in med.items()}}


def dem_hopper(a):
    """The hopper of input/hopper.py (also input/hopper.mbfcp): 25^3 touching
    spheres of radius 0.02 as 80-triangle icospheres, 26 wall prisms,
    density 1000, gravity -10 z, spring 1e6 (the scene's SPRING), damping
    0.1, friction off, margin 1e-3, dt 1e-6 s, FP64 brute force.  The lattice
    starts in contact (every neighbour inside the margin); at dt 1e-6 s it
    stays packed over the timed steps (at 1e-5 s it flies apart within ten
    steps and the later steps measure an almost empty scene)."""
    from delta_amd import _lib, engine, scene_io
    print("[bench] building the hopper scene", file=sys.stderr, flush=True)
    hs = scene_io.hopper_scene(eps=1e-3, sphere_level=1)
    ds = engine.DeviceScene(hs)
    prm = engine.step_params(method=_lib.DH_BF, dt=1e-6, gravity=(0.0, 0.0, -10.0), spring=1e6,
                             damper=0.1, friction=0.0, cell=hs.cell_size())
    L = _lib.lib()
    for _ in range(a.warmup):
        engine.step(ds, prm)
    _lib.check(L.dh_synchronize(), "sync")
    print("[bench] hopper warm", file=sys.stderr, flush=True)
    from delta_amd import tuning
    default = detection_defaults()
    if a.no_autotune or a.tuning:
        table, trials = tuning.get(), []
    else:
        table, trials = tune_detection(ds, hs, prm, L, 4_000_000)
    checks = contacts = pairs = culled = 0
    max_pairs = max_contacts = 0
    t0 = time.perf_counter()
    for _ in range(a.steps):
        info = engine.step(ds, prm)
        checks += info.triangle_checks
        culled += info.pairs_culled
        contacts += info.contacts
        pairs += info.candidate_pairs
        max_pairs = max(max_pairs, info.candidate_pairs)
        max_contacts = max(max_contacts, info.contacts)
    _lib.check(L.dh_synchronize(), "sync")
    dt = time.perf_counter() - t0
    check = None
    if trials:
        table, check = keep_faster_table(ds, prm, table, default, 2 * max_pairs + 1024,
                                         2 * max_contacts + 4096)
    # the same step captured once and replayed `steps` times from the
    # state the eager steps reached (no host synchronisation between steps)
    g = engine.StepGraph(ds, prm, pair_cap=2 * max_pairs + 1024, contact_cap=2 * max_contacts + 4096)
    try:
        g.launch()
        _, overflow = g.status()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            g.launch()
        _, ov = g.status()
        gdt = time.perf_counter() - t0
        overflow |= ov
    finally:
        g.close()
    hbm = hopper_hbm(hs, ds, prm, L)
    out = {"metric": "particle-steps/s (hopper DEM step, FP64 brute force, captured step)",
           "value": hs.n * a.steps / gdt, "unit": "particle-steps/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": gdt * 1e3 / a.steps,
           "overflow": overflow,
           "eager": {"ms_per_step": dt * 1e3 / a.steps, "value": hs.n * a.steps / dt},
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (input/hopper.py geometry)",
           "config": {"workload": "hopper: 15,625 spheres r 0.02 as 80-triangle icospheres, "
                                  "26 wall prisms, dt 1e-6 s, FP64 brute force",
                      "particles": hs.n, "triangles": int(hs.tri_offset[-1])},
           "triangle_checks_per_s": checks / dt, "culled_fraction": culled / max(checks, 1),
           "mean_candidate_pairs": pairs / a.steps, "mean_contacts": contacts / a.steps,
           "tuning": {"table": table, "trials_value_ms": trials, "whole_step_check": check},
           "hbm": hbm}
    if not a.no_cpu:
        # the reference-shaped OpenMP brute force over a sample of lattice
        # neighbours (particle pairs of the step's kind): triangle checks/s
        rng = np.random.default_rng(0)
        sample = []
        for i in rng.choice(np.arange(26, hs.n), 64, replace=False):
            sample.append((int(i), int(i + 1 if i + 1 < hs.n else i - 1)))
        rate, k, nchk, cdt, threads = cpu_rows_baseline(hs, sample, 1e-3, a.cpu_seconds)
        out["cpu_baseline"] = {"value": rate, "unit": "triangle checks/s", "cores": threads,
                               "kind": "port", "cpu": host_cpu(threads),
                               "sample": f"oracle bf over {k} lattice-neighbour particle pairs "
                                         f"({nchk} triangle pairs), OpenMP over the triangles "
                                         f"of A with a critical append (then sorted into row order), {threads} threads, "
                                         f"{cdt:.1f} s; compare with triangle_checks_per_s "
                                         "(the eager steps' checks over their own time)"}
    return out


LARGE = {
    "funnel": dict(metric="particle-steps/s (uneven funnel fill, FP64 brute force, captured step)",
                   workload="10,000 particles of 20/80/320-triangle jittered icospheres and "
                            "60-triangle hulls (radius 0.3-0.8 mm) in a 2,048-triangle cone "
                            "(half-angle 30 deg, 4 mm outlet), dt 1e-6 s, FP64 brute force",
                   n=10000, dt=1e-6),
    "drum": dict(metric="particle-steps/s (rotating box, FP64 brute force, captured step)",
                 workload="200,000 particles: 80-triangle jittered icospheres and 128-triangle "
                          "3:1 ellipsoids (0.5-1.5 mm) in a 12-triangle box turning at 2 rad/s, "
                          "dt 5e-7 s, FP64 brute force",
                 n=200000, dt=5e-7),
}


def dem_large(a, kind, steps=None, warmup=None, learned=False):
    """The uneven funnel fill (10^4 particles) or the rotating box (2*10^5
    particles): eager warm-up steps size the capacities, the detection
    tiling is autotuned on the live candidate pairs, then the whole step is
    captured once and replayed `--steps` times (host clock around the
    replays, synchronised at both ends).  Spring 1e5 N/m, damping 0.1,
    friction off, margin 1e-4 m, gravity -9.81 y.  The particles start on a
    lattice whose pitch exceeds the largest particle: a lattice below the
    particle sizes puts tens of triangle-pair contacts on each overlap, whose
    summed springs fling the particles apart within a few steps at these
    time steps (and the timed steps then measure an emptying scene)."""
    from delta_amd import _lib, engine, tuning, workloads
    from delta_amd.device import DeviceArray
    cfg = LARGE[kind]
    steps = steps or a.steps
    warmup = a.warmup if warmup is None else warmup
    L = _lib.lib()
    _lib.init()
    free0, total = C.c_size_t(), C.c_size_t()
    _lib.check(L.dh_mem_info(C.byref(free0), C.byref(total)), "dh_mem_info")
    print(f"[bench] building the {kind} scene", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    hs = getattr(workloads, kind + "_scene")(n=cfg["n"])
    ds = engine.DeviceScene(hs)
    build_s = time.perf_counter() - t0
    prm = engine.step_params(method=_lib.DH_BF, dt=cfg["dt"], spring=1e5, damper=0.1,
                             friction=0.0, cell=hs.cell_size())
    max_pairs = max_contacts = 0
    for _ in range(max(warmup, 1)):
        info = engine.step(ds, prm)
        max_pairs = max(max_pairs, info.candidate_pairs)
        max_contacts = max(max_contacts, info.contacts)
    _lib.check(L.dh_synchronize(), "sync")
    print(f"[bench] {kind} warm: {max_pairs} pairs, {max_contacts} contacts", file=sys.stderr,
          flush=True)
    # eager steps, timed (each reads its counts back to size the next launches)
    checks = contacts = pairs = culled = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        info = engine.step(ds, prm)
        checks += info.triangle_checks
        culled += info.pairs_culled
        contacts += info.contacts
        pairs += info.candidate_pairs
        max_pairs = max(max_pairs, info.candidate_pairs)
        max_contacts = max(max_contacts, info.contacts)
    _lib.check(L.dh_synchronize(), "sync")
    eager_s = time.perf_counter() - t0
    # detection tiling for these pairs
    cap = max(4 * max_pairs, 1024)
    d_pairs = DeviceArray((2 * cap,), np.int32)
    npairs = C.c_int(0)
    L.dh_broadphase.argtypes = [C.POINTER(_lib.DhGeometry), C.c_void_p, C.c_double, C.c_void_p,
                                C.c_int, C.POINTER(C.c_int), C.c_void_p]
    _lib.check(L.dh_broadphase(C.byref(ds.scene.geom), ds.buf["obstacle"].ptr, prm.cell,
                               d_pairs.ptr, cap, C.byref(npairs), None), "dh_broadphase")
    ccap = 4 * max_contacts + 4096
    out_c = DeviceArray((ccap * 144,), np.uint8)
    default = detection_defaults()
    trials = [] if a.no_autotune or a.tuning else tuning.autotune_detect(
        ds.scene.geom, d_pairs.ptr, npairs.value, out_c.ptr, ccap, reps=7)
    del out_c, d_pairs
    table = tuning.get()
    print(f"[bench] tuning {table}", file=sys.stderr, flush=True)
    check = None
    if trials:
        table, check = keep_faster_table(ds, prm, table, default, 2 * max_pairs + 1024,
                                         2 * max_contacts + 4096)
    g = engine.StepGraph(ds, prm, pair_cap=2 * max_pairs + 1024,
                         contact_cap=2 * max_contacts + 4096)
    try:
        g.launch()
        _, overflow = g.status()
        t0 = time.perf_counter()
        for _ in range(steps):
            g.launch()
        ginfo, ov = g.status()
        gdt = time.perf_counter() - t0
        overflow |= ov
    finally:
        g.close()
    learned_out = None
    if learned:
        # the same window with the survivor queue through the learned narrow
        # phase, captured, from the state the brute-force window reached
        pl = engine.step_params(method=_lib.DH_BF_LEARNED, dt=cfg["dt"], spring=1e5,
                                damper=0.1, friction=0.0, cell=hs.cell_size())
        info = engine.step(ds, pl)
        _lib.check(L.dh_synchronize(), "sync")
        gl = engine.StepGraph(ds, pl, pair_cap=2 * max_pairs + 1024,
                              contact_cap=2 * max_contacts + 4096)
        try:
            gl.launch()
            _, lov = gl.status()
            t0 = time.perf_counter()
            for _ in range(steps):
                gl.launch()
            _, ov2 = gl.status()
            lgdt = time.perf_counter() - t0
            lov |= ov2
        finally:
            gl.close()
        learned_out = {"ms_per_step": lgdt * 1e3 / steps, "value": hs.n * steps / lgdt,
                       "overflow": lov}
    free1 = C.c_size_t()
    _lib.check(L.dh_mem_info(C.byref(free1), C.byref(total)), "dh_mem_info")
    hbm = hopper_hbm(hs, ds, prm, L)
    n = hs.n
    out = {"metric": cfg["metric"], "value": n * steps / gdt, "unit": "particle-steps/s",
           "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": gdt * 1e3 / steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": cfg["workload"], "particles": n,
                      "triangles": int(hs.tri_offset[-1])},
           "overflow": overflow,
           "eager": {"ms_per_step": eager_s * 1e3 / steps,
                     "triangle_checks_per_s": checks / eager_s,
                     "culled_fraction": culled / max(checks, 1),
                     "mean_candidate_pairs": pairs / steps,
                     "mean_contacts": contacts / steps},
           "tuning": {"table": table,
                      "trials_value_ms": [(t["knob"], t["value"], round(t["ms"], 4))
                                          for t in trials],
                      "whole_step_check": check},
           "learned_survivor_queue": learned_out,
           "hbm_footprint_bytes": int(free0.value - free1.value),
           "scene_build_s": build_s,
           "hbm": hbm}
    if not a.no_cpu:
        # the reference-shaped OpenMP brute force over neighbouring particle
        # pairs of the scene's own meshes: triangle checks/s
        rng = np.random.default_rng(0)
        sample = [(int(i), int(i) + 1) for i in rng.integers(1, n - 1, 4096)]
        rate, k, nchk, cdt, threads = cpu_rows_baseline(hs, sample, 1e-4, a.cpu_seconds)
        out["cpu_baseline"] = {"value": rate, "unit": "triangle checks/s", "cores": threads,
                               "kind": "port", "cpu": host_cpu(threads),
                               "sample": f"oracle bf over {k} neighbouring particle pairs "
                                         f"({nchk} triangle pairs), OpenMP over the triangles "
                                         f"of A with a critical append (then sorted into row order), {threads} threads, "
                                         f"{cdt:.1f} s; compare with eager.triangle_checks_per_s"}
    return out


def main():
    a = parse()
    if a.calibrate_copy:
        copy_calibration()
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is a "
                                   "design note only (bench.py --gpus 1)"}))
        sys.exit(2)
    fixed = None
    if a.tuning:
        from delta_amd import tuning
        fixed = tuning.set(**{k: int(v) for k, v in
                              (kv.split("=") for kv in a.tuning.split(","))})
    if a.workload == "c1-dem":
        print(json.dumps(dem_c1(a)), flush=True)
        return
    if a.workload == "hopper":
        print(json.dumps(dem_hopper(a)), flush=True)
        return
    if a.workload in LARGE:
        print(json.dumps(dem_large(a, a.workload)), flush=True)
        return
    t_start = time.perf_counter()
    from delta_amd import narrow
    n = 1 << a.pairs_log2
    print(f"[bench] generating {n} pairs in HBM", file=sys.stderr, flush=True)
    batch = narrow.PairBatch(n, seed=3)
    results = {}
    train_info = None
    if "learned" in a.methods.split(","):
        trainer, _, train_info = train_network(a)
    if "mixed" in a.methods.split(","):
        batch.make_f32()  # the binary32 bulk copy is an input format, made before timing
    if fixed:
        tune_info = {"table": fixed, "fixed": True}
    else:
        tune_info = None if a.no_autotune else autotune(batch, a.methods.split(","))
    for method in a.methods.split(","):
        try:
            ms, wall = time_method(batch, method, a.steps, a.warmup)
        except Exception as ex:  # method not built yet
            print(f"[bench] {method}: {ex}", file=sys.stderr, flush=True)
            continue
        results[method] = {"ms_per_step": ms, "checks_per_s": n / (ms * 1e-3), "wall_s": wall}
        print(f"[bench] {method}: {ms:.3f} ms/pass, {n / (ms * 1e-3):.3e} pairs/s",
              file=sys.stderr, flush=True)
    head = "learned" if "learned" in results else "bf64"
    r = results[head]
    bf = results.get("bf64")
    # roofline of the pass: algorithmic bytes per pair (18 FP64 inputs,
    # P, Q, d, flag) over the measured pass time
    bytes_per = narrow.BYTES_F64
    achieved = bytes_per * n / (r["ms_per_step"] * 1e-3) / 1e9
    kernel = "k_learned" if head == "learned" else "k_tri_pairs_exact<0>"
    traffic = measured_traffic(kernel) if n == 1 << 24 else None
    roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
            "kernel": kernel, "bytes_per_pair": bytes_per,
            "traffic_source": "profiles/k_learned_tri_pairs_2p24_hbm_bytes.md" if traffic else None}
    extra = {}
    if bf:
        fl = narrow.FLOP_BF * n / (bf["ms_per_step"] * 1e-3) / 1e12
        extra["bf64_fp64_tflops"] = fl
        extra["bf64_fp64_frac"] = fl / PEAK_FP64_TFLOPS
        extra["bf64_checks_per_s"] = bf["checks_per_s"]
        extra["speedup_vs_bf64"] = r["checks_per_s"] / bf["checks_per_s"]
    out = {
        "metric": "triangle-pair checks/s, learned-hybrid vs plain FP64 BF, at 1e-6 dist err",
        "value": r["checks_per_s"],
        "unit": "triangle-pair checks/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if head == "bf64" else "bf16+f64",
        "data": "synthetic",
        "config": {"workload": "isolated narrow-phase microbench: 2^24 random triangle pairs, "
                               "1 mm cube, separation log-uniform 1e-7..1e-2 m, 5% near-parallel, "
                               "1% intersecting",
                   "pairs": n, "method": head},
        "roofline": roof,
        "methods": {k: {"ms_per_step": v["ms_per_step"], "checks_per_s": v["checks_per_s"]}
                    for k, v in results.items()},
    }
    out.update(extra)
    if tune_info:
        out["tuning"] = tune_info
    if train_info:
        out["network"] = train_info
        out["learned_vs_bf64"] = learned_accuracy(batch)
    if "mixed" in results:
        out["mixed_vs_bf64"] = mixed_accuracy(batch)
    if "learned_penalty" in results and "penalty_cert" in results:
        out["network_warm_start"] = warm_start_report(batch)
    if "bf64_exact" in results:
        # the generator's intersecting pairs under both semantics: bfSolver
        # (feature distance > 0) and the exact mode (d = 0)
        kind = batch.kind.download()
        batch.run("bf64_exact")
        _, _, d_x, f_x = batch.results()
        batch.run("bf64")
        _, _, d_b, _ = batch.results()
        hit = (f_x & 16) != 0
        out["intersecting_pairs"] = {
            "generated": int((kind == 2).sum()), "exact_mode_d0": int(hit.sum()),
            "bfSolver_d_min_on_them": float(d_b[hit].min()) if hit.any() else None,
            "bfSolver_d_median_on_them": float(np.median(d_b[hit])) if hit.any() else None}
    if not a.no_cpu:
        cb, chk = cpu_baseline(batch, a.cpu_seconds)
        out["cpu_baseline"] = cb
        out["bf64_parity_sample"] = chk
        if cb:
            out["speedup_vs_cpu"] = r["checks_per_s"] / cb["value"]
    if not a.no_dem:
        out["dem"] = dem_sections(a, t_start)
    print(json.dumps(out), flush=True)


def c4_section(a, n_train=4 << 20, n_test=1 << 20, epochs=10):
    """The learned warm start on the DEM step's distribution: the network
    trained on n_train triangle pairs of the 64-icosphere lattice's
    candidate particle pairs (lattices of seed 6), evaluated on n_test
    held-out pairs of other lattices (seed 7): top-1/top-2 feature accuracy,
    the share the learned decision finishes without the brute force, and both
    paths' pass times on the held-out pairs (their results are bfSolver's bit
    for bit).  The trained network stays installed for the learned DEM
    steps."""
    from delta_amd import _lib, learn, narrow
    from delta_amd.device import DeviceArray
    t0 = time.perf_counter()
    train = DeviceArray.from_host(learn.c1_pairs(n_train, seed=6))
    test_soa = learn.c1_pairs(n_test, seed=7)
    build_s = time.perf_counter() - t0
    ds = learn.Dataset(train, n_train)
    t0 = time.perf_counter()
    tr = learn.Trainer(seed=0)
    losses = tr.fit(ds, epochs=epochs, batch=8192, lr=3e-3)
    train_s = time.perf_counter() - t0
    held = learn.Dataset(DeviceArray.from_host(test_soa), n_test)
    top1, top2 = tr.accuracy(held)
    tr.install()
    a._c4_trainer = tr  # keeps the installed parameters alive
    batch = narrow.PairBatch.from_host(test_soa)
    res = {}
    for m in ("bf64", "learned"):
        ms, _ = time_method(batch, m, 5, 2)
        res[m] = ms
    batch.run("learned")
    cnt = _lib.DhPairCounters()
    _lib.check(_lib.lib().dh_tri_pairs_counters(C.byref(cnt)), "counters")
    _, _, d, _ = batch.results()
    batch.run("bf64")
    _, _, d0, _ = batch.results()
    return {"train_pairs": n_train, "held_out_pairs": n_test, "epochs": epochs,
            "final_loss": losses[-1], "held_out_top1": float(top1),
            "held_out_top2": float(top2),
            "held_out_finished_by_network": cnt.certified / n_test,
            "held_out_distance_bitwise_equal": bool(np.array_equal(d.view(np.uint64),
                                                                   d0.view(np.uint64))),
            "held_out_ms_per_pass": res,
            "dataset_build_s": build_s, "train_s": train_s,
            "data": "triangle pairs of the candidate particle pairs of 64-icosphere lattices "
                    "(train: seed 6, held out: seed 7), labelled by the FP64 brute force"}


def dem_sections(a, t_start, budget_s=420.0):
    """The DEM workloads beside the unique-pair headline: the network trained
    on the DEM step's triangle pairs (C4), C1 (all 2000 steps, brute force
    and learned survivor queue), a window of the uneven funnel fill and of
    the rotating box (20 captured steps each after eager warm-up steps, the
    box also with the learned survivor queue); each part is skipped,
    with the reason, when the run has used `budget_s` seconds, and an error
    in one part is recorded without losing the line."""
    parts = {}
    for name, fn in (("c4_learned", lambda: c4_section(a)),
                     ("c1", lambda: dem_c1(a, learned=hasattr(a, "_c4_trainer"))),
                     ("funnel", lambda: dem_large(a, "funnel", steps=20, warmup=3)),
                     ("drum", lambda: dem_large(a, "drum", steps=20, warmup=2,
                                                learned=hasattr(a, "_c4_trainer")))):
        used = time.perf_counter() - t_start
        if used > budget_s:
            parts[name] = {"skipped": f"bench time budget ({used:.0f} s of {budget_s:.0f} s used)"}
            continue
        print(f"[bench] dem: {name}", file=sys.stderr, flush=True)
        try:
            parts[name] = fn()
        except Exception as ex:  # recorded, the headline line is still printed
            parts[name] = {"error": f"{type(ex).__name__}: {ex}"}
        print(f"[bench] dem: {name} done", file=sys.stderr, flush=True)
    return parts


if __name__ == "__main__":
    main()
