#!/usr/bin/env python3
"""METCF SpMM benchmark on one MI355X.

Workload (default): the 232,965-node symmetric Chung-Lu power-law graph
(exponent 2.1, mean degree 492, self loops, seeded relabel, ~114.8M nnz,
values 1/deg of the row) times a dense 232,965 x 128 FP32 B, U[-1,1].
A step is one SpMM C = A*B with A resident on the GPU in METCF form and B
resident in HBM: the B conversion pass (BF16 forms) plus the SpMM kernel.

Prints ONE JSON line on stdout: GFLOP/s (2*nnz*K/t) of the headline form
(split-BF16), the roofline of the SpMM kernel against HBM, the CPU baseline
(the oracle's OpenMP CSR SpMM on the host cores) and per-form numbers with
their max-rel / Frobenius errors against the FP64 oracle.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

HBM_PEAK_GBS = 8000.0

from accspmm_amd.workloads import (SERVING_KS, WORKLOADS, batch_specs, make_graph,  # noqa: E402
                                   poisson_gaps, serving_pool, serving_requests)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def alg_bytes(M, N, nnz, K, form):
    s = 2 if form == 2 else 4
    return 4 * (M + 1) + nnz * (4 + s) + N * K * s + M * K * 4


# sources of the SpMM kernels: a traffic profile is valid for the build it
# was measured on, identified by this digest
KERNEL_SOURCES = ("accspmm_amd/csrc/spmm_device.hpp", "accspmm_amd/csrc/spmm.hip",
                  "accspmm_amd/csrc/internal.hpp", "accspmm_amd/csrc/builder.hip")


def kernel_source_stamp() -> str:
    import hashlib
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        h.update((REPO / rel).read_bytes())
    return h.hexdigest()[:16]


def sig(x: float, digits: int = 4) -> float:
    """x rounded to `digits` significant digits (errors like 3e-7 stay visible)."""
    return float(f"{x:.{digits}g}") if x else 0.0


def traffic_file(workload: str, K: int, form: str) -> Path:
    """Traffic profile of the workload at K in the form, in bench.py's
    default layout and library tile."""
    return REPO / "profiles" / f"traffic_{workload}_k{K}_{form}.json"


def kernel_short_name(full: str) -> str:
    """`spmm_kernel<1, 8, 16, false, true>` from the demangled name rocprofv3
    writes (return type, namespaces and parameter list dropped)."""
    s = full.replace("void ", "", 1) if full.startswith("void ") else full
    s = s.replace("(anonymous namespace)::", "")
    s = s.split("(")[0].strip()
    head, sep, tail = s.partition("<")
    return head.split("::")[-1] + sep + tail


def read_counters(path: Path) -> dict:
    """Per-kernel mean of each counter per dispatch, from a rocprofv3
    counter-collection CSV (KiB for FETCH_SIZE / WRITE_SIZE)."""
    import csv
    acc_v, disp = {}, {}
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            name = kernel_short_name(row["Kernel_Name"])
            key = (name, row["Counter_Name"])
            acc_v[key] = acc_v.get(key, 0.0) + float(row["Counter_Value"])
            disp.setdefault(key, set()).add(row["Dispatch_Id"])
    return {k: (v / len(disp[k]), len(disp[k])) for k, v in acc_v.items()}


def summarize_traffic(d: Path, workload: str, K: int, form: str) -> dict:
    """The traffic JSON of one SpMM launch from four rocprofv3 --pmc runs
    under `d` (fetch/write over `bench.py --profile-leg`, cfetch/cwrite over
    `bench.py --calibrate-copy`): FETCH_SIZE and WRITE_SIZE of the SpMM
    kernel and the split-window reduction per dispatch, scaled by what the
    same counters read for a 1 GiB device copy."""
    def find(tag):
        hits = sorted(d.rglob(f"{tag}_counter_collection.csv"))
        if not hits:
            raise SystemExit(f"no {tag}_counter_collection.csv under {d}")
        return read_counters(hits[0])
    fetch, write, cf, cw = find("fetch"), find("write"), find("cfetch"), find("cwrite")
    copy = "__amd_rocclr_copyBuffer"
    cfk = max(v[0] for (n, c), v in cf.items() if n.startswith(copy) and c == "FETCH_SIZE")
    cwk = max(v[0] for (n, c), v in cw.items() if n.startswith(copy) and c == "WRITE_SIZE")
    gib = float(1 << 30)
    ff, wf = gib / (cfk * 1024), gib / (cwk * 1024)
    kern = {}
    for src, cname in ((fetch, "FETCH_SIZE"), (write, "WRITE_SIZE")):
        for (n, c), (v, nd) in src.items():
            if c == cname and n.startswith(("spmm_kernel", "k_reduce_split", "convert_b")):
                kern.setdefault(n, {})[cname] = round(v, 2)
                kern[n]["dispatches"] = nd
    if not any(n.startswith("spmm_kernel") for n in kern):
        raise SystemExit(f"no spmm_kernel dispatch in the counter runs under {d}")

    def hbm(names):
        return int(sum(k.get("FETCH_SIZE", 0) * 1024 * ff + k.get("WRITE_SIZE", 0) * 1024 * wf
                       for n, k in kern.items() if n.startswith(names)))
    total = hbm(("spmm_kernel", "k_reduce_split"))
    return {"workload": workload, "K": K, "form": form,
            "source_stamp": kernel_source_stamp(),
            "launch": "one SpMM step of the head form at the library tile "
                      "(bench.py --profile-leg)",
            "counters_kib_per_dispatch": kern,
            "calibration": {"kernel": copy + ", 1 GiB device-to-device copy",
                            "FETCH_SIZE_kib": cfk, "WRITE_SIZE_kib": cwk,
                            "fetch_factor": round(ff, 4), "write_factor": round(wf, 4)},
            "hbm_bytes_per_launch": total,
            "hbm_bytes_per_step": hbm(("spmm_kernel", "k_reduce_split", "convert_b")),
            "formula": "per launch: sum over the SpMM kernel and k_reduce_split of FETCH_SIZE "
                       "KiB * 1024 * fetch_factor + WRITE_SIZE KiB * 1024 * write_factor; per "
                       "step: the same with the B conversion pass (convert_b) added",
            "commands": [
                f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir> -o fetch -- "
                f"python3 bench.py --workload {workload} --K {K} --forms {form} --profile-leg",
                f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d <dir> -o write -- "
                f"python3 bench.py --workload {workload} --K {K} --forms {form} --profile-leg",
                "rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir> -o cfetch -- "
                "python3 bench.py --calibrate-copy",
                "rocprofv3 --pmc WRITE_SIZE --output-format csv -d <dir> -o cwrite -- "
                "python3 bench.py --calibrate-copy",
                f"python3 bench.py --workload {workload} --K {K} --forms {form} "
                f"--traffic-from <dir>"]}


def time_form(g, B, Cd, form, tile, steps, warmup):
    import torch
    for _ in range(warmup):
        g.prepare_b(B, form, tile)
        g.spmm_prepared(B, Cd, form, tile)
    torch.cuda.synchronize()
    e0 = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    e1 = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    e2 = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        e0[i].record()
        g.prepare_b(B, form, tile)
        e1[i].record()
        g.spmm_prepared(B, Cd, form, tile)
        e2[i].record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    conv = [a.elapsed_time(b) for a, b in zip(e0, e1)]
    kern = [a.elapsed_time(b) for a, b in zip(e1, e2)]
    return {
        "ms_per_step": (t1 - t0) * 1e3 / steps,
        "kernel_ms_mean": statistics.fmean(kern),
        "kernel_ms_median": statistics.median(kern),
        "convert_ms_mean": statistics.fmean(conv),
    }


def format_diagnostics(g, K: int, form: int) -> dict:
    """Two ungraded diagnostics beside the roofline: the METCF format's bytes
    4(W+1) + 4(T+1) + 8T*RW/8 + 32T + s_A*nnz, and the window-gather bytes
    sum_w uniq_w * K * s_B (uniq_w = the window's condensed columns, i.e.
    the non-padding sparseA2B entries), the B traffic L2/MALL must serve."""
    import ctypes

    import numpy as np

    import accspmm_amd as acc
    inf = g.info
    s = 2 if form == acc.FORM_BF16 else 4
    a2b = np.empty(inf.num_tiles * 8, np.uint32)
    null = ctypes.c_void_p()
    acc._check(acc.lib().accspmm_metcf_export(g.handle, null, null, null,
                                              a2b.ctypes.data, null, null, null),
               "metcf_export")
    uniq = int(np.count_nonzero(a2b != 0xFFFFFFFF))
    T, W = inf.num_tiles, inf.num_windows
    return {"format_bytes": 4 * (W + 1) + 4 * (T + 1) + 8 * T * (inf.row_window // 8) +
                            32 * T + s * inf.nnz,
            "gather_bytes": uniq * K * s, "gathered_tile_columns": uniq}


def copy_bandwidth() -> float:
    """Device-to-device copy rate of a 1 GiB buffer (read + write bytes per
    second, GB/s), the measured counterpart of the 8 TB/s HBM peak."""
    import torch
    n = 1 << 28
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(10):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 10
    del a, b
    torch.cuda.empty_cache()
    return 2.0 * 4 * n / (ms * 1e-3) / 1e9


def cpu_baseline(csr, B, K, min_seconds=10.0, max_runs=60):
    """The oracle's OpenMP FP32 CSR SpMM on the host cores (a baseline only):
    one untimed run, then timed runs until at least min_seconds of CPU work
    (at least 3, at most max_runs); the median run is reported."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    out, used = O.spmm_f32_omp(csr.row_ptr, csr.col_idx, csr.data, B, threads=threads)
    times = []
    t_start = time.perf_counter()
    while len(times) < max_runs and (time.perf_counter() - t_start < min_seconds
                                     or len(times) < 3):
        t0 = time.perf_counter()
        O.spmm_f32_omp(csr.row_ptr, csr.col_idx, csr.data, B, out=out, threads=threads)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    return {"value": 2.0 * csr.nnz * K / med / 1e9, "unit": "GFLOP/s", "cores": used,
            "kind": "port", "runs": len(times), "seconds_per_run": med}


def check_errors(csr, B, results):
    """Max-rel (componentwise vs |A||B|), Frobenius and plain elementwise
    max-rel (over entries with |C64| >= 1e-3 max|C64|) errors vs the FP64
    oracle, for every form's C."""
    import numpy as np

    from oracle import oracle as O
    C64 = O.spmm_f64(csr.row_ptr, csr.col_idx, csr.data, B)
    ab = O.spmm_abs_f64(csr.row_ptr, csr.col_idx, csr.data, B)
    big = np.abs(C64) >= 1e-3 * np.abs(C64).max()
    errs = {}
    for name, Cm in results.items():
        mx, fr = O.errors(Cm, C64, ab)
        plain = float(np.max(np.abs(Cm[big] - C64[big]) / np.abs(C64[big]))) if big.any() else 0.0
        errs[name] = (mx, fr, plain)
    # the reference's precision class for context: A and B rounded to TF32
    # (round to nearest, ties away) as its kernels do, accumulation exact
    Ct = O.spmm_tf32_f64(csr.row_ptr, csr.col_idx, csr.data, B).astype(np.float32)
    mx, fr = O.errors(Ct, C64, ab)
    errs["_reference_tf32"] = (mx, fr, None)
    return errs


def ingest_bench(workload: str) -> dict:
    """File to resident graph: the workload graph's pattern with symmetric
    values (a hash of (min(r,c), max(r,c)) mapped to 24-bit fractions in
    (0,1], printed with 9 digits) written as a symmetric Matrix Market file
    (lower triangle, column-major), read back canonicalised by the sequential
    tokenizer and by the line-parallel reader (both must return that CSR bit
    for bit), then the METCF build; and the device canonicaliser on the same
    CSR with every row reversed (must restore it) and on the canonical CSR
    (check pass only)."""
    import shutil
    import tempfile

    import numpy as np
    import torch

    import accspmm_amd as acc
    pattern = make_graph(WORKLOADS[workload])
    rows_of = np.repeat(np.arange(pattern.rows, dtype=np.uint64),
                        np.diff(pattern.row_ptr.astype(np.int64)))
    cols64 = pattern.col_idx.astype(np.uint64)
    lo, hi = np.minimum(rows_of, cols64), np.maximum(rows_of, cols64)
    del rows_of, cols64
    h = (lo * np.uint64(0x9E3779B1) + hi * np.uint64(0x85EBCA77)) & np.uint64(0xFFFFFF)
    del lo, hi
    vals = ((h + np.uint64(1)).astype(np.float64) / float(1 << 24)).astype(np.float32)
    del h
    csr = acc.CSR(pattern.rows, pattern.cols, pattern.row_ptr, pattern.col_idx, vals)
    del pattern
    tmp = tempfile.mkdtemp(prefix="accspmm_ingest_")
    out = {"metric": "mtx_ingest_seconds", "unit": "s", "higher_is_better": False,
           "data": "synthetic", "config": {"workload": workload, "rows": csr.rows,
                                           "nnz": csr.nnz}}
    try:
        path = os.path.join(tmp, "graph.mtx")
        t0 = time.perf_counter()
        csr.write_mtx(path, symmetric=True)
        out["write_s"] = round(time.perf_counter() - t0, 3)
        out["file_bytes"] = os.path.getsize(path)
        log(f"bench: wrote {out['file_bytes']} bytes in {out['write_s']} s")
        reads, exact = {}, True
        saved = os.environ.get("ACCSPMM_HOST_THREADS")
        for th in ("1", ""):
            if th:
                os.environ["ACCSPMM_HOST_THREADS"] = th
            else:
                os.environ.pop("ACCSPMM_HOST_THREADS", None)
            t0 = time.perf_counter()
            got = acc.read_mtx(path, canonicalize=True)
            dt = time.perf_counter() - t0
            used = acc.lib().accspmm_read_mtx_threads_used()
            reads[str(used)] = round(dt, 3)
            exact = exact and np.array_equal(got.row_ptr, csr.row_ptr) and \
                np.array_equal(got.col_idx, csr.col_idx) and \
                np.array_equal(got.data.view(np.uint32), csr.data.view(np.uint32))
            log(f"bench: read_mtx on {used} thread(s): {dt:.3f} s, exact={exact}")
            del got
        if saved is None:
            os.environ.pop("ACCSPMM_HOST_THREADS", None)
        else:
            os.environ["ACCSPMM_HOST_THREADS"] = saved
        out["read_canonical_s"] = reads
        out["round_trip_exact"] = bool(exact)
        torch.cuda.init()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g = acc.DeviceGraph(csr)
        torch.cuda.synchronize()
        out["metcf_build_s"] = round(time.perf_counter() - t0, 3)
        out["file_to_graph_s"] = round(min(reads.values()) + out["metcf_build_s"], 3)
        del g
        rp, ci, va = csr.to_device()
        lens = (rp[1:] - rp[:-1]).long()
        row_of = torch.repeat_interleave(torch.arange(csr.rows, device="cuda"), lens)
        pos = torch.arange(csr.nnz, device="cuda")
        rev = rp.long()[row_of] + rp.long()[row_of + 1] - 1 - pos
        ci_r, va_r, rp_r = ci[rev].contiguous(), va[rev].contiguous(), rp.clone()
        del row_of, pos, rev, lens
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n = acc.canonicalize_device(csr.rows, csr.cols, rp_r, ci_r, va_r)
        out["gpu_canonicalize_reversed_s"] = round(time.perf_counter() - t0, 4)
        restored = n == csr.nnz and torch.equal(rp_r, rp) and torch.equal(ci_r, ci) and \
            torch.equal(va_r, va)
        out["gpu_canonicalize_exact"] = bool(restored)
        t0 = time.perf_counter()
        acc.canonicalize_device(csr.rows, csr.cols, rp_r, ci_r, va_r)
        out["gpu_canonicalize_check_s"] = round(time.perf_counter() - t0, 4)
        log(f"bench: device canonicalize {out['gpu_canonicalize_reversed_s']} s "
            f"(reversed rows, exact={restored}), {out['gpu_canonicalize_check_s']} s (check)")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return out


ORDERINGS = (
    ("input", dict(reorder_rows=0)),
    ("rows_by_length", dict(reorder_rows=1)),
    ("degree_symmetric", dict(symmetric_order=2)),
    ("rcm", dict(symmetric_order=1)),
)


def ordering_bench(workload: str, K: int, form_name: str, steps: int, warmup: int) -> dict:
    """The locality orderings on one workload (automatic row window each):
    tiles T, tile fill nnz / (T * RW * 8), window-gather bytes, build time
    and the SpMM step time of `form_name` at its library tile."""
    import torch

    import accspmm_amd as acc
    from accspmm_amd.workloads import make_graph_by_name

    name_to_form = {"fp32": acc.FORM_FP32, "split_bf16": acc.FORM_SPLIT_BF16,
                    "bf16": acc.FORM_BF16}
    form = name_to_form[form_name]
    csr, bseed = make_graph_by_name(workload)
    B = torch.from_numpy(acc.gen_dense(bseed + K, csr.cols, K)).cuda()
    Cd = torch.empty((csr.rows, K), dtype=torch.float32, device="cuda")
    out = {"workload": workload, "K": K, "form": form_name, "rows": csr.rows, "nnz": csr.nnz}
    for label, kw in ORDERINGS:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g = acc.DeviceGraph(csr, acc.default_build_opts(0, **kw))
        torch.cuda.synchronize()
        build_ms = (time.perf_counter() - t0) * 1e3
        inf = g.info
        diag = format_diagnostics(g, K, form)
        tile = g.tile_select(K, form)
        r = time_form(g, B, Cd, form, tile, steps, warmup)
        row = {"row_window": inf.row_window, "tiles": inf.num_tiles,
               "tile_fill": sig(inf.nnz / (inf.num_tiles * inf.row_window * 8.0)),
               "gather_gb": sig(diag["gather_bytes"] / 1e9),
               "build_ms": round(build_ms, 1), "ms_per_step": sig(r["ms_per_step"]),
               "kernel_ms": sig(r["kernel_ms_mean"]),
               "gflops": round(2.0 * csr.nnz * K / (r["ms_per_step"] * 1e-3) / 1e9, 1),
               "tile": list(tile.as_tuple())}
        out[label] = row
        log(f"ordering: {workload} {label}: RW {row['row_window']}, {row['tiles']} tiles, "
            f"fill {row['tile_fill']}, gathers {row['gather_gb']} GB, build {build_ms:.0f} ms, "
            f"{row['ms_per_step']} ms/step ({row['gflops']} GFLOP/s)")
        del g
    return out


def batch_bench(ngraphs: int, steps: int, warmup: int, tune: bool = True):
    """Batching: ngraphs Chung-Lu graphs (500-50,000 nodes log-uniform, mean
    degree U[4,64], exponent U[2.0,2.5], U(0,1] values, graph seed 1000+i),
    K uniform in {16, 64, 128}, B U[-1,1] (seed 2000+i), split-BF16.  With
    tune, the batch autotuner picks the batch-wide NB cap and short-window
    packing (the library-tile batch is timed too).  Times one persistent
    launch over all of them (eager and hipGraph replay) against the same
    problems as back-to-back standalone calls, and checks that every batched
    C equals its standalone C bit for bit."""
    import numpy as np
    import torch

    import accspmm_amd as acc
    probs, flops = [], 0.0
    for sp in batch_specs(ngraphs):
        K = sp["K"]
        csr = acc.gen_chung_lu(sp["n"], sp["deg"], sp["gamma"], sp["seed"], 0)
        g = acc.DeviceGraph(csr, acc.default_build_opts(8, build_adpbme=0))
        B = torch.from_numpy(acc.gen_dense(sp["bseed"], csr.cols, K)).cuda()
        probs.append((g, B, torch.empty((csr.rows, K), device="cuda"),
                      torch.empty((csr.rows, K), device="cuda")))
        flops += 2.0 * csr.nnz * K
    form = acc.FORM_SPLIT_BF16
    # library tiles first, then the batch autotuner's pick (batch-wide NB cap
    # and short-window packing, which leave every C bit as it is)
    batch = acc.Batch([(g, B, Cb) for g, B, Cb, _ in probs], form)
    untuned_ms, tuned_pick_ms = None, None
    if tune:
        batch.run()
        for _ in range(warmup):
            batch.run()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            batch.run()
        e1.record()
        torch.cuda.synchronize()
        untuned_ms = e0.elapsed_time(e1) / steps
        t0 = time.perf_counter()
        tuned_pick_ms = batch.autotune()
        log(f"batch: autotuned in {time.perf_counter() - t0:.2f} s ({tuned_pick_ms:.3f} ms)")

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    def sequential():
        for g, B, _, Cs in probs:
            g.spmm(B, Cs, form)

    seq_ms = timed(sequential)
    eager_ms = timed(lambda: batch.run())
    graph_ms = timed(lambda: batch.run(graph=True))
    torch.cuda.synchronize()
    same = all(torch.equal(Cb, Cs) for _, _, Cb, Cs in probs)
    nnz = sum(int(g.info.nnz) for g, _, _, _ in probs)
    tasks = batch.num_tasks
    batch.close()
    out = {"graphs": ngraphs, "nnz": nnz, "form": "split_bf16", "tasks": tasks,
           "sequential_ms": round(seq_ms, 4), "batched_ms": round(eager_ms, 4),
           "graph_replay_ms": round(graph_ms, 4),
           "speedup_vs_sequential": round(seq_ms / min(eager_ms, graph_ms), 3),
           "tiles": "batch autotuner (NB cap, short-window packing)" if tune
                    else "library default",
           "batched_ms_library_tiles": round(untuned_ms, 4) if untuned_ms else None,
           "batched_gflops": round(flops / (min(eager_ms, graph_ms) * 1e-3) / 1e9, 1),
           "bitwise_equal_to_standalone": bool(same)}
    log(f"batch: {ngraphs} graphs, {nnz} nnz: sequential {seq_ms:.3f} ms, one launch "
        f"{eager_ms:.3f} ms, graph replay {graph_ms:.3f} ms, equal={same}")
    return out


# serving HBM budget beyond the resident graphs: in-flight workspaces (B
# conversion and split partials) must fit in it, so at most two requests
# on a 2M-node graph at K = 128 (1.03 GB of converted B each) run at once
SERVE_WORKSPACE_BUDGET = 3 << 30


def serve_bench(requests: int, seed: int = 4000, big_nodes: int = 2_000_000,
                row_window: int = 0, reorder: int = 2):
    """Serving: Poisson arrivals over a pool of 64 resident Chung-Lu graphs
    (500-50,000 nodes, mean degree U[4,64], exponent U[2.0,2.5]) plus two
    2M-node mean-degree-32 graphs; each request is one graph (uniform) and K
    uniform in {32, 128}, split-BF16.  Runs the continuous-batching server and
    the one-launch-per-request baseline: saturation throughput of each, then
    p50/p99 latency of both at the same fixed QPS.  Both servers admit
    against an HBM budget of the resident graphs plus SERVE_WORKSPACE_BUDGET.
    Then the same request stream split by class (requests on the two 2M-node
    graphs; on the 64 small ones): each class's saturation time per request in
    both modes, the measured cap of the throughput ratio.  Graphs are
    registered with the given row window and reorder_rows build options."""
    import numpy as np
    import torch

    import accspmm_amd as acc
    opts = acc.default_build_opts(row_window, reorder_rows=reorder)
    pool = serving_pool(big_nodes=big_nodes)
    log(f"serve: pool of {len(pool)} graphs, {sum(c.nnz for c in pool)} nnz")
    bufs = {}
    for gi, c in enumerate(pool):
        for K in SERVING_KS:
            B = torch.rand((c.cols, K), device="cuda") * 2 - 1
            Cs = [torch.empty((c.rows, K), device="cuda") for _ in range(2)]
            bufs[(gi, K)] = (B, Cs)
    gsel, ksel = serving_requests(len(pool), requests, seed)
    flops = np.array([2.0 * pool[g].nnz * k for g, k in zip(gsel, ksel)])
    big = [gi for gi, c in enumerate(pool) if c.rows >= big_nodes]
    is_big = np.isin(gsel, big)

    def run(server, qps, sel=None):
        idx = np.arange(requests) if sel is None else np.flatnonzero(sel)
        gaps = poisson_gaps(len(idx), qps)
        rids = []
        t0 = time.perf_counter()
        t_next = t0
        for j, i in enumerate(idx):
            t_next += gaps[j]
            while qps and time.perf_counter() < t_next:
                pass
            B, Cs = bufs[(int(gsel[i]), int(ksel[i]))]
            rids.append(server.submit(gids[int(gsel[i])], B, Cs[j % 2]))
        server.drain()
        t1 = time.perf_counter()
        lats = np.array([server.poll(r) for r in rids])
        return t1 - t0, lats

    def server(mode):
        s = acc.Server(hbm_budget_bytes=budget, max_batch=64, max_inflight=3,
                       form=acc.FORM_SPLIT_BF16, one_per_launch=(mode == "one_per_launch"))
        return s, [s.register(c, opts) for c in pool]

    # the budget: what the pool holds resident (measured once) plus the
    # in-flight workspace allowance
    probe = acc.Server(hbm_budget_bytes=1 << 50, form=acc.FORM_SPLIT_BF16)
    for c in pool:
        probe.register(c, opts)
    resident = int(probe.stats().resident_bytes)
    probe.close()
    budget = resident + SERVE_WORKSPACE_BUDGET
    log(f"serve: {resident / 1e9:.2f} GB resident, budget {budget / 1e9:.2f} GB")

    out = {}
    for mode in ("one_per_launch", "batched"):
        srv, gids = server(mode)
        run(srv, 0)  # warm-up: schedules, workspaces
        wall, lats = run(srv, 0)
        st = srv.stats()
        out[mode] = {"saturation_req_per_s": requests / wall,
                     "saturation_gflops": flops.sum() / wall / 1e9,
                     "launches": int(st.launches),
                     "peak_inflight_bytes": int(st.peak_inflight_bytes)}
        srv.close()
        log(f"serve: {mode}: {requests / wall:.0f} req/s, {flops.sum() / wall / 1e9:.1f} GFLOP/s "
            f"at saturation, peak in-flight workspaces {st.peak_inflight_bytes / 1e9:.2f} GB")
    qps = 0.7 * out["one_per_launch"]["saturation_req_per_s"]
    for mode in ("one_per_launch", "batched"):
        srv, gids = server(mode)
        run(srv, 0)
        wall, lats = run(srv, qps)
        srv.close()
        out[mode].update({"qps": round(qps, 1), "p50_ms": float(np.percentile(lats, 50) * 1e3),
                          "p99_ms": float(np.percentile(lats, 99) * 1e3),
                          "gflops_at_qps": flops.sum() / wall / 1e9})
        log(f"serve: {mode} at {qps:.0f} QPS: p50 {out[mode]['p50_ms']:.3f} ms, "
            f"p99 {out[mode]['p99_ms']:.3f} ms")
    out["throughput_ratio"] = (out["batched"]["saturation_req_per_s"] /
                               out["one_per_launch"]["saturation_req_per_s"])
    # per class: microseconds of saturated server time per request
    per = {}
    for mode in ("one_per_launch", "batched"):
        srv, gids = server(mode)
        run(srv, 0)
        for cls, sel in (("big", is_big), ("small", ~is_big)):
            wall, _ = run(srv, 0, sel)
            per.setdefault(cls, {})[mode + "_us"] = round(wall / max(1, int(sel.sum())) * 1e6, 2)
        srv.close()
    share = float(is_big.mean())
    t_opl = share * per["big"]["one_per_launch_us"] + (1 - share) * per["small"]["one_per_launch_us"]
    t_bat = share * per["big"]["batched_us"] + (1 - share) * per["small"]["batched_us"]
    out["breakdown"] = {
        "big_share_of_requests": round(share, 4), "per_request": per,
        "modelled_ratio": round(t_opl / t_bat, 3),
        # even with free small requests the ratio stays below this: the
        # 2M-node graphs' kernel time is the same in both modes
        "ratio_cap_small_free": round(t_opl / (share * per["big"]["batched_us"]), 3)}
    log(f"serve: per request (us): {per}; ratio model {t_opl / t_bat:.2f}, cap with free "
        f"small requests {out['breakdown']['ratio_cap_small_free']:.2f}")
    out["hbm_budget_bytes"] = budget
    out["resident_bytes"] = resident
    out["requests"] = requests
    out["row_window"] = row_window
    out["reorder_rows"] = reorder
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="chung_lu_233k", choices=sorted(WORKLOADS))
    ap.add_argument("--K", type=int, default=128)
    ap.add_argument("--forms", default="split_bf16,fp32,bf16")
    ap.add_argument("--tile", default="", help="nb,wpb,split,xcd (default: library pick)")
    ap.add_argument("--row-window", type=int, default=0, choices=(0, 8, 16),
                    help="METCF row window of the timed graph (0: the library's automatic "
                         "choice)")
    ap.add_argument("--reorder", type=int, default=2, choices=(0, 1, 2),
                    help="build option reorder_rows (1: rows by decreasing length, C still "
                         "in input row order; 2: the library's automatic choice)")
    ap.add_argument("--tune", action="store_true", help="autotune each form first")
    ap.add_argument("--no-check", action="store_true", help="skip errors and CPU baseline")
    ap.add_argument("--serve-requests", type=int, default=10000,
                    help="requests per serving run (0 skips the serving benchmark)")
    ap.add_argument("--batch-graphs", type=int, default=256,
                    help="graphs in the batched-launch benchmark (0 skips it)")
    ap.add_argument("--ingest", action="store_true",
                    help="only measure Matrix Market ingest and canonicalisation of the "
                         "workload graph (prints its own JSON line)")
    ap.add_argument("--profile-leg", action="store_true",
                    help="one warm-up and one timed step of the head form only "
                         "(the run rocprofv3 --pmc profiles)")
    ap.add_argument("--traffic-from", default="",
                    help="write the traffic JSON under profiles/ from rocprofv3 counter runs "
                         "in this directory (no GPU work)")
    ap.add_argument("--k-sweep", default="32,64,128,256,512",
                    help="K values timed on the workload per form (empty: none)")
    ap.add_argument("--stencil", type=int, default=1,
                    help="also time the 100^3 stencil at K (1) or not (0)")
    ap.add_argument("--batch-tune", type=int, default=1,
                    help="run the batch autotuner in the batch leg (1) or use library tiles (0)")
    ap.add_argument("--ordering", default="",
                    help="only compare the locality orderings (input, rows by length, "
                         "degree symmetric, reverse Cuthill-McKee) on these comma-separated "
                         "workloads for the head form, and print their JSON line")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run three 1 GiB device-to-device copies (the known-size "
                         "kernel that scales the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()

    if args.traffic_from:
        head = [f for f in args.forms.split(",") if f][0]
        out = summarize_traffic(Path(args.traffic_from), args.workload, args.K, head)
        dst = traffic_file(args.workload, args.K, head)
        dst.write_text(json.dumps(out, indent=1) + "\n")
        log(f"bench: wrote {dst} ({out['hbm_bytes_per_launch'] / 1e9:.2f} GB per launch)")
        return

    if args.calibrate_copy:
        import torch
        n = 1 << 28  # 1 GiB of float32
        a = torch.ones(n, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"bench: copied {4 * n} bytes three times")
        return

    if args.ingest:
        print(json.dumps(ingest_bench(args.workload)), flush=True)
        return

    if args.ordering:
        head = [f for f in args.forms.split(",") if f][0]
        res = [ordering_bench(w, args.K, head, args.steps, args.warmup)
               for w in args.ordering.split(",") if w]
        print(json.dumps({"ordering": res}), flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(it would shard by row windows, each GPU holding its METCF slice and "
              "the B rows it gathers); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import accspmm_amd as acc

    name_to_form = {"fp32": acc.FORM_FP32, "split_bf16": acc.FORM_SPLIT_BF16,
                    "bf16": acc.FORM_BF16}
    forms = [f for f in args.forms.split(",") if f]
    spec = WORKLOADS[args.workload]
    K = args.K
    log(f"bench: generating {args.workload} ...")
    t0 = time.perf_counter()
    csr = make_graph(spec)
    gen_s = time.perf_counter() - t0
    log(f"bench: {csr.rows} rows, {csr.nnz} nnz, generated in {gen_s:.1f}s")
    Bh = acc.gen_dense(spec["bseed"] + K, csr.cols, K)
    torch.cuda.init()
    torch.cuda.synchronize()
    opts = acc.default_build_opts(args.row_window, reorder_rows=args.reorder)
    if args.tune and args.row_window == 0 and args.reorder == 2:
        # the layout autotuner picks row window and row order for the head form
        Bt = torch.from_numpy(Bh).cuda()
        Ct = torch.empty((csr.rows, K), dtype=torch.float32, device="cuda")
        opts, lay_ms = acc.autotune_layout(csr, Bt, Ct, name_to_form[forms[0]])
        log(f"bench: layout autotuner: row window {opts.row_window}, reorder_rows "
            f"{opts.reorder_rows} ({lay_ms:.3f} ms)")
        del Bt, Ct
    t0 = time.perf_counter()
    g = acc.DeviceGraph(csr, opts)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    inf = g.info
    log(f"bench: METCF built on GPU in {build_ms:.1f} ms (row window {inf.row_window}, "
        f"reorder_rows {inf.reordered}): {inf.num_windows} windows, "
        f"{inf.num_tiles} tiles, {inf.num_groups} AdpBME groups, ibd={inf.ibd}")
    B = torch.from_numpy(Bh).cuda()
    Cd = torch.empty((csr.rows, K), dtype=torch.float32, device="cuda")

    if args.profile_leg:
        form = name_to_form[forms[0]]
        tile = (acc.Tile(*[int(x) for x in args.tile.split(",")]) if args.tile
                else g.tile_select(K, form))
        time_form(g, B, Cd, form, tile, 1, 1)
        log(f"bench: profile leg {forms[0]} tile={tile.as_tuple()}")
        return

    per_form = {}
    outputs = {}
    for fname in forms:
        form = name_to_form[fname]
        if args.tile:
            tile = acc.Tile(*[int(x) for x in args.tile.split(",")])
        elif args.tune:
            tile, _ = g.autotune(B, Cd, form)
        else:
            tile = g.tile_select(K, form)
        log(f"bench: form={fname} tile={tile.as_tuple()}")
        r = time_form(g, B, Cd, form, tile, args.steps, args.warmup)
        flops = 2.0 * csr.nnz * K
        r["gflops"] = flops / (r["ms_per_step"] * 1e-3) / 1e9
        r["kernel_gflops"] = flops / (r["kernel_ms_mean"] * 1e-3) / 1e9
        ab = alg_bytes(csr.rows, csr.cols, csr.nnz, K, form)
        r["hbm_frac_kernel"] = ab / (r["kernel_ms_mean"] * 1e-3) / 1e9 / HBM_PEAK_GBS
        r["tile"] = list(tile.as_tuple())
        per_form[fname] = r
        log(f"bench: {fname}: {r['ms_per_step']:.3f} ms/step, kernel "
            f"{r['kernel_ms_mean']:.3f} ms, convert {r['convert_ms_mean']:.3f} ms, "
            f"{r['gflops']:.1f} GFLOP/s, {100 * r['hbm_frac_kernel']:.1f}% HBM roof")
        if not args.no_check:
            torch.cuda.synchronize()
            outputs[fname] = Cd.cpu().numpy()

    k_sweep = None
    if args.k_sweep:
        k_sweep = {}
        for K2 in [int(x) for x in args.k_sweep.split(",") if x]:
            B2 = torch.from_numpy(acc.gen_dense(spec["bseed"] + K2, csr.cols, K2)).cuda()
            C2 = torch.empty((csr.rows, K2), dtype=torch.float32, device="cuda")
            row = {}
            for fname in forms:
                form = name_to_form[fname]
                tile = g.tile_select(K2, form)
                # the headline's step and warm-up counts: a short wall-clock
                # run after two warm-ups read up to 10% slow at K = 32
                r = time_form(g, B2, C2, form, tile, max(10, args.steps), max(3, args.warmup))
                row[fname] = {"ms_per_step": sig(r["ms_per_step"]),
                              "gflops": round(2.0 * csr.nnz * K2 / (r["ms_per_step"] * 1e-3)
                                              / 1e9, 1),
                              "tile": list(tile.as_tuple())}
            if "split_bf16" in row and "fp32" in row:
                row["split_over_fp32"] = round(row["split_bf16"]["gflops"] /
                                               row["fp32"]["gflops"], 3)
            k_sweep[str(K2)] = row
            log(f"bench: K={K2}: " + ", ".join(f"{f} {v['gflops']:.0f}" for f, v in row.items()
                                                 if isinstance(v, dict)) + " GFLOP/s")
            del B2, C2

    diag = format_diagnostics(g, K, name_to_form[forms[0]])
    log(f"bench: format {diag['format_bytes'] / 1e9:.3f} GB, window gathers "
        f"{diag['gather_bytes'] / 1e9:.2f} GB per launch")
    cpu = None
    ref_tf32 = None
    if not args.no_check:
        log("bench: FP64 oracle check and CPU baseline ...")
        errs = check_errors(csr, Bh, outputs)
        ref_mx, ref_fr, _ = errs.pop("_reference_tf32")
        ref_tf32 = {"max_rel": ref_mx, "fro_rel": ref_fr,
                    "what": "A and B rounded to TF32 (round to nearest, ties away) as the "
                            "reference's kernels do, FP64 accumulation"}
        log(f"bench: reference TF32 input rounding: max_rel={ref_mx:.3e} fro_rel={ref_fr:.3e}")
        for fname, (mx, fr, plain) in errs.items():
            per_form[fname]["max_rel"] = mx
            per_form[fname]["fro_rel"] = fr
            per_form[fname]["plain_max_rel_large"] = plain
            log(f"bench: {fname}: max_rel={mx:.3e} fro_rel={fr:.3e} "
                f"plain max-rel (|C| >= 1e-3 max) {plain:.3e}")
        cpu = cpu_baseline(csr, Bh, K)
        cpu["sample"] = (f"full {args.workload} matrix ({csr.nnz} nnz), K={K}, "
                         f"median of {cpu.pop('runs')} runs")
        log(f"bench: CPU baseline {cpu['value']:.1f} GFLOP/s on {cpu['cores']} threads")

    # the head form on the reference's format options (row window 8, rows in
    # input order), for comparison
    input_order = None
    if inf.reordered or inf.row_window != 8:
        del g
        torch.cuda.empty_cache()
        g = acc.DeviceGraph(csr)
        form = name_to_form[forms[0]]
        tile = g.tile_select(K, form)
        r = time_form(g, B, Cd, form, tile, args.steps, args.warmup)
        input_order = {"row_window": 8, "reorder_rows": 0, "form": forms[0],
                       "tiles": g.info.num_tiles, "tile": list(tile.as_tuple()),
                       "ms_per_step": round(r["ms_per_step"], 4),
                       "kernel_ms": round(r["kernel_ms_mean"], 4),
                       "gflops": round(2.0 * csr.nnz * K / (r["ms_per_step"] * 1e-3) / 1e9, 2)}
        log(f"bench: {forms[0]} on input-order rows, row window 8: "
            f"{r['ms_per_step']:.3f} ms/step, {input_order['gflops']:.1f} GFLOP/s")

    del g, B, Cd, outputs
    torch.cuda.empty_cache()
    copy_gbs = copy_bandwidth()
    log(f"bench: device copy {copy_gbs:.0f} GB/s (read + write)")
    stencil = None
    if args.stencil and args.workload != "stencil27_100":
        st_spec = WORKLOADS["stencil27_100"]
        sc = make_graph(st_spec)
        sg = acc.DeviceGraph(sc, acc.default_build_opts(args.row_window,
                                                         reorder_rows=args.reorder))
        SB = torch.from_numpy(acc.gen_dense(st_spec["bseed"] + K, sc.cols, K)).cuda()
        SC = torch.empty((sc.rows, K), dtype=torch.float32, device="cuda")
        stencil = {"rows": sc.rows, "nnz": sc.nnz, "K": K, "row_window": sg.info.row_window,
                   "reorder_rows": sg.info.reordered}
        for fname in forms:
            form = name_to_form[fname]
            tile = sg.tile_select(K, form)
            r = time_form(sg, SB, SC, form, tile, args.steps, args.warmup)
            stencil[fname] = {"ms_per_step": sig(r["ms_per_step"]),
                              "kernel_ms_mean": sig(r["kernel_ms_mean"]),
                              "gflops": round(2.0 * sc.nnz * K / (r["ms_per_step"] * 1e-3)
                                              / 1e9, 1),
                              "hbm_frac_kernel": sig(alg_bytes(sc.rows, sc.cols, sc.nnz, K, form)
                                                     / (r["kernel_ms_mean"] * 1e-3) / 1e9
                                                     / HBM_PEAK_GBS),
                              "tile": list(tile.as_tuple())}
            log(f"bench: stencil27_100 {fname}: {r['ms_per_step']:.3f} ms/step, "
                f"{stencil[fname]['gflops']:.0f} GFLOP/s")
        if "split_bf16" in stencil and "fp32" in stencil:
            stencil["split_over_fp32"] = round(stencil["split_bf16"]["gflops"] /
                                               stencil["fp32"]["gflops"], 3)
        del sg, SB, SC, sc
        torch.cuda.empty_cache()
    batching = None
    if args.batch_graphs > 0:
        batching = batch_bench(args.batch_graphs, max(3, args.steps // 2), 2,
                               tune=bool(args.batch_tune))
        torch.cuda.empty_cache()
    serving = None
    if args.serve_requests > 0:
        serving = serve_bench(args.serve_requests, row_window=args.row_window,
                              reorder=args.reorder)

    head = forms[0]
    hr = per_form[head]
    head_form = name_to_form[head]
    ab = alg_bytes(csr.rows, csr.cols, csr.nnz, K, head_form)
    achieved = ab / (hr["kernel_ms_mean"] * 1e-3) / 1e9
    # the measured ceiling of the kernels' B access pattern: random rows of
    # the head form's gathered row size from an L2-sized region and from a
    # region the size of the gathered B
    gather = None
    row_bytes = K * (2 if head_form == acc.FORM_BF16 else 4)
    if row_bytes in (256, 512):
        gather = {"row_bytes": row_bytes,
                  "l2_region_bytes": 4096 * row_bytes,
                  "l2_region_gbs": round(acc.gather_ceiling(4096, row_bytes), 1),
                  "b_region_bytes": (csr.cols + 1) * row_bytes,
                  "b_region_gbs": round(acc.gather_ceiling(csr.cols + 1, row_bytes), 1)}
        log(f"bench: gather ceiling ({row_bytes}-byte rows): {gather['l2_region_gbs']:.0f} GB/s "
            f"from L2, {gather['b_region_gbs']:.0f} GB/s from a {gather['b_region_bytes'] / 1e6:.0f}"
            f" MB region")
    # the profiled traffic counts only for the kernels it was measured on:
    # the profile carries the digest of their sources, and the tile
    traffic, traffic_note = None, "no profile"
    tfile = traffic_file(args.workload, K, head)
    if tfile.exists() and not args.tile and not args.tune and args.row_window == 0 \
            and args.reorder == 2:
        prof = json.loads(tfile.read_text())
        if prof.get("source_stamp") == kernel_source_stamp():
            traffic = prof.get("hbm_bytes_per_launch")
            traffic_note = f"{tfile.relative_to(REPO)} (source stamp {prof['source_stamp']})"
            # the SpMM kernel's reads from beyond L2 against the gather ceiling
            ff = prof.get("calibration", {}).get("fetch_factor", 1.0)
            fetch = sum(v["FETCH_SIZE"] * 1024.0 * ff
                        for k, v in prof.get("counters_kib_per_dispatch", {}).items()
                        if k.startswith("spmm_kernel") and "FETCH_SIZE" in v)
            if gather and fetch:
                miss_gbs = fetch / (hr["kernel_ms_mean"] * 1e-3) / 1e9
                gather["kernel_miss_read_bytes"] = round(fetch)
                gather["kernel_miss_read_gbs"] = round(miss_gbs, 1)
                gather["kernel_miss_frac_of_b_region"] = round(miss_gbs / gather["b_region_gbs"], 4)
        else:
            traffic_note = (f"stale: {tfile.relative_to(REPO)} was measured at source stamp "
                            f"{prof.get('source_stamp')}, this build is {kernel_source_stamp()}")
    out = {
        "metric": "SpMM GFLOP/s (2*nnz*K/t)",
        "value": round(hr["gflops"], 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(hr["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16" if head_form != acc.FORM_FP32 else "f32",
        "data": "synthetic (seeded generators, A and B resident in HBM)",
        "config": {"workload": args.workload, "rows": csr.rows, "nnz": csr.nnz, "K": K,
                   "form": head, "row_window": inf.row_window,
                   "reorder_rows": inf.reordered,
                   "cache": "warm: A, B and C resident in HBM, timed steps after warm-up "
                            "steps on the same B",
                   "tiles": inf.num_tiles, "tile": hr["tile"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_profile": traffic_note,
                     "kernel": "spmm_kernel+k_reduce_split",
                     "stream_copy_gbs": round(copy_gbs, 1),
                     # the profiled HBM/MALL bytes per launch over this run's
                     # kernel time: the rate the gathers actually move
                     "traffic_gbs": (round(traffic / (hr["kernel_ms_mean"] * 1e-3) / 1e9, 1)
                                     if traffic else None),
                     "traffic_frac": (round(traffic / (hr["kernel_ms_mean"] * 1e-3) / 1e9 /
                                            HBM_PEAK_GBS, 4) if traffic else None),
                     "kernel_ms": round(hr["kernel_ms_mean"], 4),
                     "alg_bytes_per_launch": ab,
                     "gather_ceiling": gather},
        "diagnostics": diag,
        "reference_tf32_precision": ref_tf32,
        "cpu_baseline": cpu,
        "batching": batching,
        "serving": serving,
        "build_ms": round(build_ms, 1),
        "input_order": input_order,
        "forms": {k: {kk: (sig(vv) if isinstance(vv, float) else vv)
                      for kk, vv in v.items()} for k, v in per_form.items()},
        "k_sweep": k_sweep,
        "stencil27_100": stencil,
    }
    if "fp32" in per_form and head != "fp32":
        out["vs_fp32_form"] = round(hr["gflops"] / per_form["fp32"]["gflops"], 3)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
