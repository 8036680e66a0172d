#!/usr/bin/env python3
"""Throughput of Jess template matching on one MI355X, next to the CPU oracle.

One step = one pass of the hot path over a workload: every template against
every molecule (candidate sets, kd leaf order, Join enumeration,
superposition, RMSD threshold), molecules and templates already resident in
HBM. The metric is template x molecule queries per second.

The headline ``value`` is the library-scale batched search: 64 synthetic
assemblies of 100,000 atoms in 20 chains x 7,000 templates of 2-8 residues
(2.5 A cutoff), streamed through the C ABI in chunks of 16 assemblies, the
events left in HBM; every assembly's hit set is checked against the oracle
(per-pair counts and digests computed on the device). Beside it: the event
copy to host memory and the drop-in ``Jess.query_batch`` (Python hit
objects) on one chunk; ``plain`` is the drop-in ``list(Jess(templates).query(mol,
...))`` on one 8,000-atom protein x 500 templates; ``batched`` is
``Jess.query_batch`` on 4,096 molecules x 200 templates; ``differentiation``
is the differentiable superposition RMSD (forward + backward) over 1M
point-set pairs; ``ingestion`` is the device PDB parse (records, exact
decimals, residue-name index) of 8 of the library's assemblies written as
PDB text. Each carries a CPU baseline (the C restatement of the
reference algorithm built for the host, on the host's cores, on a bounded
sample; torch autograd through an SVD for the RMSD; the host PDB parser for
ingestion) and the search legs the
roofline of the dominant kernel, the join's k_join launches (one per
level), timed with HIP events on the library's stream in one more pass
after the timed ones.

Prints one JSON line on stdout; progress goes to stderr.
"""

import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "template×molecule queries/sec, bit-exact hits, 1 GPU vs CPU; batched vs plain"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# flops of one superposition beyond the 27 per atom pair: centring, X^T X, the
# determinant, a cyclic Jacobi of a 3x3 matrix (about 4 sweeps of 3 rotations
# of ~64 flops) and the rotation matrix
C_EIG = 1100

_M64 = (1 << 64) - 1


def _mix64(x):
    x = (x + 0x9E3779B97F4A7C15) & _M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _M64
    return x ^ (x >> 31)


def hit_digest(t, mi, atoms):
    """Order-independent digest term of a hit, as the CPU baseline computes it."""
    h = _mix64(((t & 0xFFFFFFFF) << 32) ^ (mi & 0xFFFFFFFF))
    for a in atoms:
        h = _mix64(h ^ (int(a) & 0xFFFFFFFF))
    return h


def _mix64_np(x):
    x = x + np.uint64(0x9E3779B97F4A7C15)
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def pair_digests(pair, atoms, sizes, n_templates, n_pairs):
    """Per-pair hit counts and sums (mod 2^64) of hit_digest over the hits
    (pair index, atom rows), vectorised over hits."""
    pair = np.asarray(pair, dtype=np.int64)
    t = pair % n_templates
    mi = pair // n_templates
    n = np.asarray(sizes, dtype=np.int64)[t]
    with np.errstate(over="ignore"):
        h = _mix64_np(((t.astype(np.uint64) & np.uint64(0xFFFFFFFF)) << np.uint64(32))
                      ^ (mi.astype(np.uint64) & np.uint64(0xFFFFFFFF)))
        for j in range(int(n.max()) if len(n) else 0):
            sel = n > j
            a = atoms[sel, j].astype(np.int64).astype(np.uint64) & np.uint64(0xFFFFFFFF)
            h[sel] = _mix64_np(h[sel] ^ a)
    counts = np.bincount(pair, minlength=n_pairs).astype(np.int64)
    dg = np.zeros(n_pairs, dtype=np.uint64)
    np.add.at(dg, pair, h)
    return counts, dg


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_share() -> int:
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, n)


def measure(name, scale, seed, steps, warmup, cpu_seconds, cpu, work_sample):
    from pyjess_amd import Jess, synthetic, _native

    t0 = time.time()
    mols, tpls, kw = synthetic.workload(name, seed=seed, scale=scale)
    n_atoms = sum(len(m) for m in mols)
    n_pairs = len(mols) * len(tpls)
    log(f"workload {name}: {len(mols)} molecules ({n_atoms} atoms) x {len(tpls)} templates, "
        f"generated in {time.time() - t0:.1f}s")
    qargs = (kw["rmsd_threshold"], kw["distance_cutoff"], kw["max_dynamic_distance"])
    qkw = dict(max_candidates=kw["max_candidates"])
    jess = Jess(tpls)

    def step():
        if len(mols) == 1:
            return [list(jess.query(mols[0], *qargs, **qkw))]
        return jess.query_batch(mols, *qargs, **qkw)

    for i in range(warmup):
        step()
        log(f"{name}: warmup step {i + 1}/{warmup}")
    _native.synchronize()
    stage_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        hits = step()
        stage_ms.append(jess._last_result.timings_ms.copy())
    _native.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / steps
    value = n_pairs / (elapsed / steps)
    st = np.array(stage_ms).mean(axis=0)
    # the dominant kernel's launches, timed with HIP events on the library's
    # stream in as many steps again (the events cost stream time, so the
    # steps above run without them)
    ctx = _native.context()
    ctx.set_profiling(True)
    join_ms, join_launches = [], []
    try:
        for _ in range(steps):
            step()
            join_ms.append(jess._last_result.join_kernel_ms)
            join_launches.append(jess._last_result.join_launches)
    finally:
        ctx.set_profiling(False)
    join_ms = float(np.mean(join_ms))
    launches = float(np.mean(join_launches))
    log(f"{name}: {steps} steps in {elapsed:.3f}s: {ms_per_step:.3f} ms/step, {value:,.0f} queries/s; "
        f"device ms: candidates {st[0]:.3f} kd {st[1]:.3f} join {st[2]:.3f} (k_join {join_ms:.3f} "
        f"in {launches:.0f} launches) super {st[3]:.3f} total {st[4]:.3f}")

    # the native call alone (device pipeline + event copy, no Python hit objects)
    t0 = time.perf_counter()
    for _ in range(steps):
        res = jess._device_query(mols, *qargs, 0, True)
    native_ms = 1000.0 * (time.perf_counter() - t0) / steps
    n_hits = sum(len(h) for h in hits)
    log(f"{name}: native query {native_ms:.3f} ms/step; {n_hits} hits, "
        f"{int(res.pair_matches.sum())} complete matches")

    # algorithmic bytes of the whole search (reference algorithm)
    sizes = [len(t) for t in tpls]
    bytes_alg = 40 * n_atoms + sum(32 * n + 8 * n * n for n in sizes) * len(mols)
    hit_sizes = np.array([sizes[int(p) % len(tpls)] for p in res.pair[res.flag == 1]], dtype=np.int64)
    bytes_alg += int((4 * hit_sizes + 128).sum())

    out = {
        "value": value,
        "unit": "queries/s",
        "ms_per_step": ms_per_step,
        "steps": steps,
        "warmup": warmup,
        "config": {
            "workload": name,
            "molecules": len(mols),
            "atoms": n_atoms,
            "templates": len(tpls),
            "pairs": n_pairs,
            "seed": seed,
            "rmsd_threshold": kw["rmsd_threshold"],
            "distance_cutoff": kw["distance_cutoff"],
            "max_dynamic_distance": kw["max_dynamic_distance"],
            "max_candidates": kw["max_candidates"],
        },
        "native_ms_per_step": native_ms,
        "device_ms": {"candidates": st[0], "kd_order": st[1], "join": st[2], "superposition": st[3],
                      "total": st[4], "k_join": join_ms, "k_join_launches": launches},
        "hits": n_hits,
        "complete_matches": int(res.pair_matches.sum()),
        "join_level_partials": [int(x) for x in res.level_partials[1:] if x > 0],
        "bytes_alg": bytes_alg,
    }
    hbm = {"bound": "hbm", "kernel": "whole device pipeline",
           "achieved": bytes_alg / (st[4] / 1000.0) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    hbm["frac"] = hbm["achieved"] / hbm["peak"]
    out["roofline_search_hbm"] = hbm
    out["roofline"] = None
    out["cpu_baseline"] = None
    out["hit_set_parity"] = None

    if cpu:
        from oracle import oracle

        threads = cpu_share()
        mv = [oracle.MolView.of(m) for m in mols]
        tv = oracle.TplView(tpls)
        # instrumented run of the reference algorithm: annulus tests on the
        # candidates inside every join box (sum over levels L of L x in-box)
        wmols = mols if work_sample is None else mols[:work_sample]
        work = {"inbox_tests": 0, "complete": 0, "complete_atoms": 0}
        for m in wmols:
            w = oracle.work(m, tv, kw["distance_cutoff"], kw["max_dynamic_distance"])
            for k in work:
                work[k] += w[k]
        f = len(mols) / len(wmols)
        flops_join = 8.0 * work["inbox_tests"] * f
        reps = 0
        t0 = time.perf_counter()
        while True:
            cm, hc, dg = oracle.batch(mv, tv, *qargs, n_threads=threads, native=True)
            reps += 1
            if time.perf_counter() - t0 >= cpu_seconds:
                break
        cpu_elapsed = time.perf_counter() - t0
        cpu_rate = reps * n_pairs / cpu_elapsed
        out["cpu_baseline"] = {
            "value": cpu_rate, "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"the whole {name} workload ({n_pairs} pairs) run {reps} times in {cpu_elapsed:.1f}s "
                      f"on {threads} threads (C restatement of the reference algorithm, gcc -O3 -march=native)"}
        out["gpu_vs_cpu"] = value / cpu_rate
        log(f"{name}: cpu baseline {cpu_rate:,.0f} queries/s on {threads} threads ({reps} reps)")
        # parity of the hit set (before the best_match / max_candidates policies)
        sel = res.flag == 1
        gpu_hc, gpu_dg = pair_digests(res.pair[sel], res.atoms[sel], sizes, len(tpls), n_pairs)
        parity = bool(np.array_equal(gpu_hc, hc) and np.array_equal(gpu_dg, dg.astype(np.uint64)) and
                      np.array_equal(res.pair_matches, cm))
        out["hit_set_parity"] = parity
        log(f"{name}: hit-set parity with the CPU oracle: {parity}")
        # dominant kernel: the join's k_join launches; FP64 VALU work (no
        # MFMA: nothing here is a dense contraction)
        per_launch_flops = flops_join / max(launches, 1.0)
        avg_launch_ms = join_ms / max(launches, 1.0)
        achieved = per_launch_flops / (avg_launch_ms / 1000.0) / 1e12
        out["roofline"] = {
            "bound": "fp64-valu", "kernel": "k_join", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
            "per_launch_flops": per_launch_flops, "avg_launch_ms": avg_launch_ms,
            "launches_per_step": launches,
            "note": "FP64 vector ALU work of the exact annulus tests (8 flops x level x in-box "
                    "candidate, counted by an instrumented oracle run"
                    + ("" if work_sample is None else f" on the first {len(wmols)} molecules, scaled")
                    + ") against the 78.6 TFLOP/s FP64 peak"}
        prof = os.path.join(ROOT, "profiles", f"k_join.{name}.pmc.md")
        if os.path.exists(prof):
            # HBM bytes per k_join launch from the committed counter passes
            # (FETCH_SIZE and WRITE_SIZE, calibrated against a copy kernel)
            for line in open(prof):
                if line.startswith("traffic_bytes_per_launch:"):
                    out["roofline"]["traffic"] = float(line.split(":")[1])
                    out["roofline"]["traffic_source"] = os.path.relpath(prof, ROOT)
        flops_all = flops_join + (27.0 * work["complete_atoms"] + C_EIG * work["complete"]) * f
        out["flops_alg"] = flops_all
    return out


def measure_library(scale, seed, chunk, steps, warmup, cpu_seconds, cpu, parity_all=True, search_only=False,
                    ingest_steps=0):
    """Library-scale batched search: 64 assemblies of 100,000 atoms x 7,000
    templates, streamed through jam_query in chunks of ``chunk`` assemblies
    (one batched call each, every (template, molecule) pair of the chunk).

    ``value`` counts the search with inputs and the resulting events resident
    in HBM; ``with_event_copy`` adds the copy of every event to host memory
    (pageable numpy arrays, over PCIe), with two contexts taking alternate
    chunks from two host threads so that copies overlap searches;
    ``drop_in`` is ``Jess.query_batch``
    (Python hit objects) on the first chunk. The CPU baseline is the C
    restatement on the first assembly x all templates; the hit set of every
    assembly (of the first chunk with ``parity_all`` false) is checked against
    the restatement run over all of them."""
    import ctypes

    from pyjess_amd import Jess, synthetic, _native

    t0 = time.time()
    mols, tpls, kw = synthetic.workload("library-64", seed=seed, scale=scale)
    n_atoms = sum(len(m) for m in mols)
    T = len(tpls)
    n_pairs = len(mols) * T
    log(f"library: {len(mols)} assemblies ({n_atoms} atoms) x {T} templates, generated in {time.time() - t0:.1f}s")
    qargs = (kw["rmsd_threshold"], kw["distance_cutoff"], kw["max_dynamic_distance"])
    jess = Jess(tpls)
    ctx = _native.context()
    L = _native.lib()
    tset = jess._device_templates(ctx)
    chunks = [mols[i:i + chunk] for i in range(0, len(mols), chunk)]
    msets = [_native.MoleculeSet(ctx, c) for c in chunks]
    params = _native.Params(*qargs, 0, 1)

    part = np.zeros(32, dtype=np.int64)   # partial matches per join level, of the last pass

    def device_pass():
        n_ev = 0
        stage = np.zeros(5)
        jms, jl = 0.0, 0
        part[:] = 0
        for ms in msets:
            h = ctypes.c_void_p()
            with ctx.lock:
                _native._check(L.jam_query(ctx.handle, tset.handle, ms.handle, ctypes.byref(params),
                                           ctypes.byref(h)))
                try:
                    n_ev += L.jam_result_n_events(h)
                    t = np.zeros(5)
                    _native._check(L.jam_result_timings(h, _native._ptr(t)))
                    stage += t
                    x, nl = ctypes.c_double(0.0), ctypes.c_int32(0)
                    _native._check(L.jam_result_join_kernels(h, ctypes.byref(x), ctypes.byref(nl)))
                    jms += x.value
                    jl += nl.value
                    lp = np.zeros(32, dtype=np.int64)
                    _native._check(L.jam_result_level_partials(h, _native._ptr(lp)))
                    np.add(part, lp, out=part)
                finally:
                    L.jam_result_free(h)
        return n_ev, stage, jms, jl

    for _ in range(warmup):
        device_pass()
    _native.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        n_events, stage, _, _ = device_pass()
    _native.synchronize()
    dev_s = (time.perf_counter() - t0) / steps
    # the k_join launches timed with HIP events in one more pass
    ctx.set_profiling(True)
    try:
        _, _, join_ms, join_launches = device_pass()
    finally:
        ctx.set_profiling(False)
    value = n_pairs / dev_s
    log(f"library: device-resident {dev_s * 1000:.1f} ms/step, {value:,.0f} queries/s, {n_events} events; "
        f"device ms: candidates {stage[0]:.1f} kd {stage[1]:.1f} join {stage[2]:.1f} super {stage[3]:.1f} "
        f"total {stage[4]:.1f}")

    copy_s = drop_s = None
    n_hits0 = 0
    partials = part.copy()
    if not search_only:
        # the same with every event copied to host memory: two contexts (each
        # with its own stream and workspace) take alternate chunks from two host
        # threads, so that one chunk's event copy overlaps the next one's search
        ctx2 = _native.Context()
        tset2 = _native.TemplateSet(ctx2, jess._templates)
        lanes = [[(tset, msets[i]) for i in range(0, len(msets), 2)],
                 [(tset2, _native.MoleculeSet(ctx2, chunks[i])) for i in range(1, len(msets), 2)]]
        if lanes[1]:
            _native.query(lanes[1][0][0], lanes[1][0][1], *qargs, 0, True)   # sizes the second workspace
        errors = []

        def copy_lane(k):
            try:
                for ts, ms in lanes[k]:
                    r = _native.query(ts, ms, *qargs, 0, True)
                    del r
            except Exception as e:   # re-raised on the main thread
                errors.append(e)

        threads = [threading.Thread(target=copy_lane, args=(k,)) for k in range(2)]
        t0 = time.perf_counter()
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        copy_s = time.perf_counter() - t0
        if errors:
            raise errors[0]
        log(f"library: with the event copy {copy_s * 1000:.1f} ms, {n_pairs / copy_s:,.0f} queries/s "
            f"(two contexts, alternate chunks)")

        # the drop-in batched API on the first chunk (Python hit objects)
        t0 = time.perf_counter()
        hits = jess.query_batch(chunks[0], *qargs, max_candidates=kw["max_candidates"])
        drop_s = time.perf_counter() - t0
        n_hits0 = sum(len(h) for h in hits)
        del hits
        log(f"library: query_batch on {len(chunks[0])} assemblies {drop_s * 1000:.1f} ms "
            f"({len(chunks[0]) * T / drop_s:,.0f} queries/s, {n_hits0} hits)")

    out = {
        "value": value,
        "unit": "queries/s",
        "ms_per_step": dev_s * 1000.0,
        "steps": steps,
        "config": {"workload": "library-64", "molecules": len(mols), "atoms": n_atoms, "templates": T,
                   "pairs": n_pairs, "chunk": chunk, "seed": seed, "rmsd_threshold": kw["rmsd_threshold"],
                   "distance_cutoff": kw["distance_cutoff"], "max_dynamic_distance": kw["max_dynamic_distance"],
                   "max_candidates": kw["max_candidates"]},
        "events": int(n_events),
        "join_level_partials": [int(x) for x in partials[1:] if x > 0],
        "device_ms": {"candidates": stage[0], "kd_order": stage[1], "join": stage[2], "superposition": stage[3],
                      "total": stage[4]},
        "with_event_copy": None if copy_s is None else {"value": n_pairs / copy_s, "unit": "queries/s",
                                                        "seconds": copy_s, "contexts": 2},
        "drop_in": None if drop_s is None else {"value": len(chunks[0]) * T / drop_s, "unit": "queries/s",
                                                "molecules": len(chunks[0]), "hits": n_hits0, "seconds": drop_s},
        "k_join_ms": join_ms,
        "k_join_launches": join_launches,
        "roofline": None,
        "cpu_baseline": None,
        "hit_set_parity": None,
    }
    if cpu:
        from oracle import oracle

        threads = cpu_share()
        mv = [oracle.MolView.of(mols[0])]
        tv = oracle.TplView(tpls)
        reps = 0
        t0 = time.perf_counter()
        while True:
            cm, hc, dg = oracle.batch(mv, tv, *qargs, n_threads=threads, native=True)
            reps += 1
            if time.perf_counter() - t0 >= cpu_seconds:
                break
        cpu_elapsed = time.perf_counter() - t0
        cpu_rate = reps * T / cpu_elapsed
        out["cpu_baseline"] = {
            "value": cpu_rate, "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"the first assembly ({len(mols[0])} atoms) x all {T} templates, run {reps} times in "
                      f"{cpu_elapsed:.1f}s on {threads} threads (C restatement of the reference algorithm, "
                      f"gcc -O3 -march=native)"}
        out["gpu_vs_cpu"] = value / cpu_rate
        # algorithmic FP64 work of the join (8 flops per annulus test of an
        # in-box candidate) from an instrumented oracle run on the first
        # assembly x every tenth template, scaled to the workload
        wt = tpls[::10]
        w = oracle.work(mols[0], oracle.TplView(wt), kw["distance_cutoff"], kw["max_dynamic_distance"])
        flops_join = 8.0 * w["inbox_tests"] * (T / len(wt)) * len(mols)
        per_launch = flops_join / max(join_launches, 1)
        avg_ms = join_ms / max(join_launches, 1)
        achieved = per_launch / (avg_ms / 1000.0) / 1e12
        out["roofline"] = {
            "bound": "fp64-valu", "kernel": "k_join", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
            "per_launch_flops": per_launch, "avg_launch_ms": avg_ms, "launches_per_step": join_launches,
            "note": "FP64 vector ALU work of the exact annulus tests (8 flops x level x in-box candidate, "
                    f"oracle run on the first assembly x {len(wt)} templates, scaled) against the 78.6 "
                    "TFLOP/s FP64 peak"}
        prof = os.path.join(ROOT, "profiles", "k_join.library-64.pmc.md")
        if os.path.exists(prof):
            for line in open(prof):
                if line.startswith("traffic_bytes_per_launch:"):
                    out["roofline"]["traffic"] = float(line.split(":")[1])
                    out["roofline"]["traffic_source"] = os.path.relpath(prof, ROOT)
        # hit-set parity on every assembly: per pair, complete matches, hits
        # and an order-independent digest of the hits -- on the device from
        # the events in HBM (jam_result_pair_digests), chunk by chunk, and by
        # the oracle over all assemblies on the host's cores
        n_par = len(mols) if parity_all else min(len(mols), chunk)
        gpu_cm = np.zeros(n_par * T, dtype=np.int64)
        gpu_hc = np.zeros(n_par * T, dtype=np.int64)
        gpu_dg = np.zeros(n_par * T, dtype=np.uint64)
        for ci, ms in enumerate(msets):
            m0 = ci * chunk
            if m0 >= n_par:
                break
            h = ctypes.c_void_p()
            with ctx.lock:
                _native._check(L.jam_query(ctx.handle, tset.handle, ms.handle, ctypes.byref(params),
                                           ctypes.byref(h)))
                try:
                    npc = L.jam_result_n_pairs(h)
                    cm_c = np.zeros(npc, dtype=np.int64)
                    hc_c = np.zeros(npc, dtype=np.int64)
                    dg_c = np.zeros(npc, dtype=np.uint64)
                    _native._check(L.jam_result_pair_matches(h, _native._ptr(cm_c)))
                    _native._check(L.jam_result_pair_digests(h, m0, T, _native._ptr(hc_c), _native._ptr(dg_c)))
                finally:
                    L.jam_result_free(h)
            k = min(npc, (n_par - m0) * T)
            gpu_cm[m0 * T:m0 * T + k] = cm_c[:k]
            gpu_hc[m0 * T:m0 * T + k] = hc_c[:k]
            gpu_dg[m0 * T:m0 * T + k] = dg_c[:k]
        t0 = time.perf_counter()
        cm_all, hc_all, dg_all = oracle.batch([oracle.MolView.of(m) for m in mols[:n_par]], tv, *qargs,
                                              n_threads=threads)
        full_s = time.perf_counter() - t0
        parity = bool(np.array_equal(gpu_cm, cm_all) and np.array_equal(gpu_hc, hc_all) and
                      np.array_equal(gpu_dg, dg_all.astype(np.uint64)))
        out["hit_set_parity"] = parity
        out["hit_set_parity_scope"] = {"assemblies": n_par, "of": len(mols), "pairs": int(n_par * T),
                                       "hits": int(hc_all.sum()), "complete_matches": int(cm_all.sum()),
                                       "oracle_seconds": full_s, "oracle_threads": threads}
        log(f"library: cpu baseline {cpu_rate:,.0f} queries/s on {threads} threads; hit-set parity on "
            f"{n_par}/{len(mols)} assemblies ({int(hc_all.sum())} hits, oracle {full_s:.1f}s): {parity}")
    if ingest_steps > 0:
        # 8 assemblies (65.5 MB of PDB text), the ingest-8 workload
        out["ingestion"] = measure_ingest(mols[:8], ingest_steps, cpu_seconds, cpu)
    return out


def measure_ingest(mols, steps, cpu_seconds, cpu):
    """PDB ingestion of library assemblies (8 in the default run): their PDB texts
    (``Molecule.dumps`` of the assemblies translated to non-negative
    coordinates, serial numbers wrapped at 99,999 as the five serial
    columns force) parsed on the device into a batch with its residue-name
    index (jam_molecules_parse_pdb: text upload from pageable host memory,
    line split, records, exact decimal conversion, index; ``value``), and
    the drop-in ``Molecule.load_many`` (host ``Molecule`` objects as well).
    The CPU baseline is the host parser, ``Molecule.load(..., device=False)``
    (a pure-Python restatement of PDBMoleculeParser.load + Atom.loads), on
    one core over the first text. Parity: every field of the first molecule
    equals the host parse, bit for bit, and every molecule's coordinates,
    names, residues, chains and serials equal the ones written."""
    import io

    from pyjess_amd import Molecule, _native

    t0 = time.time()
    written, texts = [], []
    for m in mols:
        # translated to non-negative coordinates (eight columns hold -999.999
        # to 9999.999), rounded to the 0.001 A the text keeps
        xyz = np.round(m.xyz - m.xyz.min(axis=0), 3)
        wrapped = np.arange(len(m), dtype=np.int32) % 99999 + 1
        w = Molecule.from_arrays(xyz, m.names, m.resnames, m.chains, m.resseq, wrapped, id=m.id)
        written.append(w)
        texts.append(w.dumps().encode())
    n_bytes = sum(len(t) for t in texts)
    n_atoms = sum(len(m) for m in mols)
    log(f"ingestion: {len(texts)} PDB texts, {n_bytes / 1e6:.1f} MB, {n_atoms} atoms, written in "
        f"{time.time() - t0:.1f}s")
    ctx = _native.context()
    _native.ParsedMolecules(ctx, texts)          # warm-up
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        batch = _native.ParsedMolecules(ctx, texts)
        times.append(time.perf_counter() - t0)
        del batch
    dev_s = float(np.median(times))
    t0 = time.perf_counter()
    loaded = Molecule.load_many([io.BytesIO(t) for t in texts])
    drop_s = time.perf_counter() - t0
    fields = ("xyz", "names", "resnames", "chains", "resseq", "serial", "altloc", "icode", "occupancy",
              "tempfactor", "segid", "element", "charge")

    def bits(a):
        return a.view(np.dtype(f"u{a.itemsize}")) if a.dtype.kind == "f" else a

    parity = all(np.array_equal(bits(getattr(d, f)), bits(getattr(w, f)))
                 for d, w in zip(loaded, written) for f in fields[:6])
    out = {
        "value": n_atoms / dev_s, "unit": "atoms/s", "ms_per_chunk": dev_s * 1e3,
        "text_GB_per_s": n_bytes / dev_s / 1e9, "steps": steps,
        "config": {"texts": len(texts), "bytes": n_bytes, "atoms": n_atoms},
        "drop_in": {"value": n_atoms / drop_s, "unit": "atoms/s", "ms": drop_s * 1e3,
                    "note": "Molecule.load_many: device parse, header fields, Molecule objects on the host"},
        "cpu_baseline": None,
    }
    if cpu:
        text0 = texts[0].decode()
        reps, t0 = 0, time.perf_counter()
        while True:
            host = Molecule.load(io.StringIO(text0), device=False)
            reps += 1
            if time.perf_counter() - t0 >= cpu_seconds:
                break
        cpu_s = time.perf_counter() - t0
        cpu_rate = reps * len(mols[0]) / cpu_s
        parity = parity and all(np.array_equal(bits(getattr(loaded[0], f)), bits(getattr(host, f))) for f in fields)
        out["cpu_baseline"] = {
            "value": cpu_rate, "unit": "atoms/s", "cores": 1, "kind": "port",
            "sample": f"the first text ({len(texts[0]) / 1e6:.1f} MB, {len(mols[0])} atoms) parsed {reps} times in "
                      f"{cpu_s:.1f}s by Molecule.load(device=False) (pure-Python PDBMoleculeParser.load + Atom.loads)"}
        out["gpu_vs_cpu"] = out["value"] / cpu_rate
    out["parity"] = bool(parity)
    log(f"ingestion: device {dev_s * 1e3:.1f} ms per chunk ({out['value'] / 1e6:.1f} M atoms/s, "
        f"{out['text_GB_per_s']:.2f} GB/s of text), load_many {drop_s * 1e3:.0f} ms, parity {parity}")
    return out


def measure_diff(n_sets, seed, steps, warmup, cpu_seconds, cpu):
    """Differentiable superposition RMSD: forward + backward over n_sets point-set
    pairs (k uniform in 3..12, 2% near-collinear k = 3), FP64 on the GPU."""
    import torch

    from pyjess_amd import synthetic
    from pyjess_amd.autograd import superposition_rmsd

    t0 = time.time()
    a, b, off = synthetic.superposition_pairs(np.random.default_rng(seed), n_sets)
    n_pts = int(off[-1])
    log(f"differentiation: {n_sets} point-set pairs ({n_pts} points) generated in {time.time() - t0:.1f}s")
    dev = torch.device("cuda:0")
    ta = torch.tensor(a, device=dev, requires_grad=True)
    tb = torch.tensor(b, device=dev, requires_grad=True)
    toff = torch.tensor(off, device=dev)

    def step():
        ta.grad = None
        tb.grad = None
        r = superposition_rmsd(ta, tb, toff)
        r.sum().backward()
        return r

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = n_sets * steps / elapsed
    # the two kernels alone: jam_rmsd_forward / jam_rmsd_backward called
    # through the C ABI (no offset check, no torch reduction) between HIP
    # events on the stream they are launched on (torch's current stream)
    import ctypes

    from pyjess_amd import _native

    L = _native.lib()
    rm = torch.empty(n_sets, dtype=torch.float64, device=dev)
    rot = torch.empty((n_sets, 9), dtype=torch.float64, device=dev)
    gr = torch.ones(n_sets, dtype=torch.float64, device=dev)
    ga_k = torch.empty_like(ta)
    gb_k = torch.empty_like(tb)
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    fwd_ms, bwd_ms = [], []
    for _ in range(steps + 1):
        ev[0].record()
        _native._check(L.jam_rmsd_forward(ptr(ta), ptr(tb), ptr(toff), n_sets, n_pts, ptr(rm), ptr(rot), stream))
        ev[1].record()
        _native._check(L.jam_rmsd_backward(ptr(ta), ptr(tb), ptr(toff), n_sets, n_pts, ptr(rm), ptr(rot), ptr(gr),
                                           ptr(ga_k), ptr(gb_k), stream))
        ev[2].record()
        torch.cuda.synchronize()
        fwd_ms.append(ev[0].elapsed_time(ev[1]))
        bwd_ms.append(ev[1].elapsed_time(ev[2]))
    fwd_ms, bwd_ms = float(np.median(fwd_ms[1:])), float(np.median(bwd_ms[1:]))
    del rm, rot, gr, ga_k, gb_k
    # algorithmic HBM bytes: forward reads a, b and the offsets, writes the
    # RMSD and the rotation (kept for the backward); backward reads a, b,
    # offsets, RMSD, rotation and upstream gradient, writes both gradients
    fwd_bytes = 2 * 24 * n_pts + 8 * (n_sets + 1) + 8 * n_sets + 72 * n_sets
    bwd_bytes = 2 * 24 * n_pts + 8 * (n_sets + 1) + 16 * n_sets + 72 * n_sets + 2 * 24 * n_pts
    bwd_gbs = bwd_bytes / (bwd_ms / 1000.0) / 1e9
    log(f"differentiation: {value:,.0f} matches/s forward+backward ({1000 * elapsed / steps:.3f} ms/step); "
        f"forward {fwd_ms:.3f} ms ({fwd_bytes / fwd_ms / 1e6:.0f} GB/s), backward {bwd_ms:.3f} ms "
        f"({bwd_gbs:.0f} GB/s), the kernels alone")
    out = {
        "value": value,
        "unit": "matches/s (forward+backward)",
        "ms_per_step": 1000.0 * elapsed / steps,
        "config": {"workload": "rmsd-grad-1m", "sets": n_sets, "points": n_pts, "k": "3..12", "seed": seed,
                   "near_collinear": 0.02},
        "kernel_ms": {"forward": fwd_ms, "backward": bwd_ms},
        "roofline": {"bound": "hbm", "kernel": "k_rmsd_fwd", "achieved": fwd_bytes / (fwd_ms / 1000.0) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": None, "bytes_per_launch": fwd_bytes},
        "roofline_backward": {"bound": "hbm", "kernel": "k_rmsd_bwd", "achieved": bwd_gbs, "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "traffic": None, "bytes_per_launch": bwd_bytes},
        "cpu_baseline": None,
        "gradient_parity": None,
    }
    for k in ("roofline", "roofline_backward"):
        out[k]["frac"] = out[k]["achieved"] / out[k]["peak"]
    prof = os.path.join(ROOT, "profiles", "k_rmsd.rmsd-grad-1m.pmc.md")
    if os.path.exists(prof) and n_sets == 1_000_000:
        # HBM bytes per launch from the committed counter passes (FETCH_SIZE
        # and WRITE_SIZE over this workload, calibrated against a copy)
        for line in open(prof):
            for key, k in (("traffic_bytes_per_launch_fwd:", "roofline"),
                           ("traffic_bytes_per_launch_bwd:", "roofline_backward")):
                if line.startswith(key):
                    out[k]["traffic"] = float(line.split(":")[1])
                    out[k]["traffic_source"] = os.path.relpath(prof, ROOT)
    if cpu:
        from oracle.kabsch_torch import kabsch_rmsd_grad

        # parity on a sample: gradients against autograd through an SVD (CPU)
        ga = ta.grad.detach().cpu().numpy()
        gb = tb.grad.detach().cpu().numpy()
        rr = r.detach().cpu().numpy()
        worst = 0.0
        for s in range(0, n_sets, max(1, n_sets // 200)):
            A, B = a[off[s]:off[s + 1]], b[off[s]:off[s + 1]]
            r0, xa, xb = kabsch_rmsd_grad(A, B)
            x = np.concatenate([ga[off[s]:off[s + 1]], gb[off[s]:off[s + 1]]]).ravel()
            y = np.concatenate([xa, xb]).ravel()
            worst = max(worst, float(np.linalg.norm(x - y) / np.linalg.norm(y)), abs(rr[s] - r0) / r0)
        out["gradient_parity"] = {"max_rel_error_vs_cpu_autograd": worst, "sample": 200}
        # CPU baseline: the same autograd-through-SVD, batched by set size in
        # float64 torch on the host cores, on a bounded sample
        threads = cpu_share()
        torch.set_num_threads(threads)
        m = min(n_sets, 200_000)
        sizes = np.diff(off[:m + 1])
        groups = {}
        for k in np.unique(sizes):
            idx = np.nonzero(sizes == k)[0]
            rows = (off[idx][:, None] + np.arange(k)[None, :]).ravel()
            groups[int(k)] = (torch.tensor(a[rows].reshape(len(idx), k, 3)), torch.tensor(b[rows].reshape(len(idx), k, 3)))

        def cpu_step():
            for A, B in groups.values():
                A = A.clone().requires_grad_(True)
                B = B.clone().requires_grad_(True)
                Ac = A - A.mean(dim=1, keepdim=True)
                Bc = B - B.mean(dim=1, keepdim=True)
                X = Bc.transpose(1, 2) @ Ac
                S = torch.linalg.svdvals(X)
                sg = torch.where(torch.linalg.det(X) < 0, -1.0, 1.0).to(A.dtype)
                e = (Ac * Ac).sum((1, 2)) + (Bc * Bc).sum((1, 2)) - 2.0 * (S[:, 0] + S[:, 1] + sg * S[:, 2])
                rr = torch.sqrt(torch.clamp(e, min=0.0) / A.shape[1])
                rr.sum().backward()

        reps = 0
        t0 = time.perf_counter()
        while True:
            cpu_step()
            reps += 1
            if time.perf_counter() - t0 >= cpu_seconds:
                break
        cpu_elapsed = time.perf_counter() - t0
        cpu_rate = reps * m / cpu_elapsed
        out["cpu_baseline"] = {"value": cpu_rate, "unit": "matches/s (forward+backward)", "cores": threads,
                               "kind": "port",
                               "sample": f"the first {m} sets, batched by size, autograd through torch's SVD in "
                                         f"float64 on {threads} threads, {reps} passes in {cpu_elapsed:.1f}s"}
        out["gpu_vs_cpu"] = value / cpu_rate
        log(f"differentiation: cpu baseline {cpu_rate:,.0f} matches/s on {threads} threads; "
            f"gradient parity (max rel vs CPU autograd) {worst:.2e}")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed passes of the headline workload")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="library-64",
                    help="library-64 (the headline, with the other workloads beside it) | plain-2k | plain-8k | "
                         "batch-4k | rmsd-grad-1m | ingest-8 (one workload alone) | copy-1g (the counter "
                         "calibration copy)")
    ap.add_argument("--plain-steps", type=int, default=200,
                    help="timed steps of the plain-8k measurement beside the headline (0: skip it)")
    ap.add_argument("--parity-first-chunk", action="store_true",
                    help="library-64: check the hit set of the first chunk of assemblies only")
    ap.add_argument("--library-chunk", type=int, default=16, help="assemblies per batched call of library-64")
    ap.add_argument("--scale", type=float, default=1.0,
                    help="fraction of the molecules of the workload measured (library-64: of its 64 assemblies)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="budget of each CPU baseline")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--search-only", action="store_true",
                    help="library-64: only the device-resident search (no event copy, no query_batch leg)")
    ap.add_argument("--ingest-steps", type=int, default=3,
                    help="timed device parses of one chunk of library PDB texts (0: skip the ingestion leg)")
    ap.add_argument("--batch-scale", type=float, default=1.0,
                    help="fraction of the 4,096 molecules of the batched measurement (0: skip it)")
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--diff-sets", type=int, default=1_000_000,
                    help="point-set pairs of the differentiable-RMSD measurement (0: skip it)")
    args = ap.parse_args()
    if args.gpus != 1:
        log(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (queries are independent: "
            "they would be dealt to GPUs by molecule, with no exchange during the search)")
        sys.exit(2)

    if args.workload == "rmsd-grad-1m":
        # the differentiable RMSD alone
        d = measure_diff(args.diff_sets or 1_000_000, args.seed, args.steps, args.warmup, args.cpu_seconds,
                         not args.no_cpu)
        out = {"metric": "superposition RMSD forward+backward matches/sec, 1 GPU vs CPU", "value": d["value"],
               "unit": d["unit"], "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
               "ms_per_step": d["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "f64", "data": "synthetic", "config": dict(d["config"], parallelism="1 GPU"),
               "roofline": d["roofline"], "cpu_baseline": d["cpu_baseline"], "differentiation": d}
        print(json.dumps(out), flush=True)
        return
    if args.workload == "copy-1g":
        # a plain device copy of 1 GiB: the reference for scaling the HBM
        # counters (FETCH_SIZE, WRITE_SIZE) of the profiles to bytes
        import torch

        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        print(json.dumps({"metric": "device copy GB/s (read + write)", "value": 2 * src.numel() * 8 / dt / 1e9,
                          "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": 0, "ms_per_step": dt * 1e3,
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                          "data": "synthetic", "config": {"workload": "copy-1g", "bytes": src.numel() * 8}}),
              flush=True)
        return
    if args.workload == "ingest-8":
        # the PDB ingestion of one chunk of library assemblies alone
        from pyjess_amd import synthetic

        mols, _, _ = synthetic.workload("library-64", seed=args.seed, scale=8 / 64)
        g = measure_ingest(mols, args.steps, args.cpu_seconds, not args.no_cpu)
        out = {"metric": "PDB atoms ingested/sec (device parse + residue-name index), 1 GPU vs CPU",
               "value": g["value"], "unit": g["unit"], "n_gpus": 1, "steps": args.steps, "warmup": 1,
               "ms_per_step": g["ms_per_chunk"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "u8", "data": "synthetic", "config": dict(g["config"], workload="ingest-8"),
               "roofline": None, "cpu_baseline": g["cpu_baseline"], "ingestion": g}
        print(json.dumps(out), flush=True)
        return
    if args.workload in ("plain-2k", "plain-8k", "batch-4k"):
        # one search workload alone
        if args.workload == "batch-4k":
            args.steps = min(args.steps, 10)
        run = measure(args.workload, args.scale, args.seed, args.steps, args.warmup, args.cpu_seconds,
                      not args.no_cpu, None if args.workload != "batch-4k" else 32)
        out = {"metric": METRIC, "value": run["value"], "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "ms_per_step": run["ms_per_step"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
               "config": dict(run["config"], parallelism="1 GPU"), "roofline": run["roofline"],
               "cpu_baseline": run["cpu_baseline"], "run": run}
        print(json.dumps(out), flush=True)
        return

    # the headline: the library-scale batched search (64 assemblies of
    # 100,000 atoms x 7,000 templates), exactly --steps timed passes
    library = measure_library(args.scale, args.seed, args.library_chunk, args.steps, args.warmup,
                              args.cpu_seconds, not args.no_cpu, parity_all=not args.parity_first_chunk,
                              search_only=args.search_only, ingest_steps=args.ingest_steps)
    plain = None
    if args.plain_steps > 0:
        plain = measure("plain-8k", 1.0, args.seed, args.plain_steps, 3, args.cpu_seconds, not args.no_cpu, None)
    batched = None
    if args.batch_scale > 0:
        batched = measure("batch-4k", args.batch_scale, args.seed, args.batch_steps, 1, args.cpu_seconds,
                          not args.no_cpu, 32)
        if plain is not None:
            batched["vs_plain"] = batched["value"] / plain["value"]
    if plain is not None:
        library["vs_plain"] = library["value"] / plain["value"]
    diff = None
    if args.diff_sets > 0:
        diff = measure_diff(args.diff_sets, args.seed, max(3, args.batch_steps), 1, args.cpu_seconds,
                            not args.no_cpu)
    rest = {k: v for k, v in library.items() if k not in ("value", "unit", "ms_per_step", "steps", "config",
                                                           "roofline", "cpu_baseline")}
    out = {
        "metric": METRIC,
        "value": library["value"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": library["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": dict(library["config"], parallelism="1 GPU"),
        "roofline": library["roofline"],
        "cpu_baseline": library["cpu_baseline"],
        "library": rest,
        "plain": plain,
        "batched": batched,
        "differentiation": diff,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
