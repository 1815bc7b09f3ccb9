#!/usr/bin/env python3
"""All-pairs Meta-Storms benchmark on one MI355X.

Workload (one step = one pass of the hot path over the batch):
  workload 1 of the benchmark plan: seeded Yule tree with 99,322 leaves
  (branch lengths ~ Exp(0.02) clipped to [0, 0.99]), 10,000 samples with
  200-2,000 nonzero OTUs (Zipf 1.1 popularity, NegBinomial(40, 0.4) counts,
  copy-number correction), all 49,995,000 sample pairs.

Timed step (inputs resident in HBM as CSR abundances): rebuild the kernel's
device layout from the samples (walk layout, fixed-point embedding, block
bitmaps) and compute every pair into a device buffer. The headline is the
shared-minimum kernel (exact identity of the walk, fixed point, <= 1e-5 vs
FP64); the bit-exact FP32 walk drop-in is timed beside it.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import importlib
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12  # FP32/INT32 VALU lane-ops per second, non-packed
HBM_PEAK_GBS = 8000.0
# shared-minimum tiles: one v_sad_u32 per pair-entry plus one column-sum add
# per staged row and lane (a lane holds 16 pairs)
SM_OPS_PER_PAIR_ENTRY = 17.0 / 16.0
METRIC = "sample-pairs/sec & pair·node updates/sec on 1 MI355X at ≤1e-5 abs err vs FP64"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed_steps(ctx, steps, warmup):
    """Warmup, then exactly `steps` passes bracketed by device synchronisation."""
    for _ in range(warmup):
        ctx.prepare(force=True)
        ctx.all_pairs_device()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.prepare(force=True)
        ctx.all_pairs_device()
    ctx.synchronize()
    return (time.perf_counter() - t0) / steps


def cpu_model() -> str:
    """The host CPU's model name and logical CPU count (/proc/cpuinfo)."""
    try:
        names = [ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo") if ln.startswith("model name")]
        return f"{names[0]} ({len(names)} logical CPUs visible, {len(os.sched_getaffinity(0))} usable)"
    except (OSError, IndexError):
        return "unknown"


def cpu_baseline(cms, tree, cp, counts, abd, gpu_sm, gpu_walk, n_all, threads, sample_n, pairs):
    """Reference Calc_sim timed on the host cores over a bounded pair sample
    (every pair of a seeded random subset of sample_n samples), plus the
    accuracy spot-check of the GPU results on pairs of that subset."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    pick = np.sort(np.random.default_rng(5).choice(n_all, sample_n, replace=False))
    sub = counts.subset(pick)
    res = {}
    drv = orc.ref_binary("ref_driver_O3")
    if drv:
        with tempfile.TemporaryDirectory() as d:
            cms.write_db(d, "bench", tree, cp)
            table = os.path.join(d, "table.txt")
            cms.write_table(table, sub, tree.leaf_n)
            env = dict(os.environ, ParallelMETA=d)
            r = subprocess.run([drv, "-D", "G", "-T", table, "--time", str(pairs), str(threads)], env=env,
                               capture_output=True, text=True, timeout=600)
            if r.returncode == 0:
                j = json.loads(r.stdout.strip().splitlines()[-1])
                res = {"value": j["pairs_per_sec"], "unit": "sample-pairs/s", "cores": threads, "kind": "reference",
                       "sample": f"{j['pairs']} pairs: the upper triangle of {sample_n} samples drawn at random "
                                 f"(seed 5) from the {n_all} (99,322-leaf tree); reference _Comp_Tree::Calc_sim built "
                                 f"-O3 -ffp-contract=off, OpenMP dynamic,1 on {threads} threads, wall clock "
                                 f"{j['seconds']:.2f} s"}
                # the reference as shipped (its Makefile builds -O0), on a tenth of the pairs
                drv0 = orc.ref_binary("ref_driver_O0")
                if drv0:
                    r0 = subprocess.run([drv0, "-D", "G", "-T", table, "--time", str(max(pairs // 10, 1)),
                                         str(threads)], env=env, capture_output=True, text=True, timeout=600)
                    if r0.returncode == 0:
                        j0 = json.loads(r0.stdout.strip().splitlines()[-1])
                        res["as_shipped_O0"] = {"value": j0["pairs_per_sec"], "pairs": j0["pairs"],
                                                "seconds": j0["seconds"]}
            else:
                log(f"reference driver failed: {r.stderr[-500:]}")
    if not res:
        dense = sub.dense(tree.leaf_n)
        t0 = time.perf_counter()
        got = orc.all_pairs_f32(tree, dense, threads, pairs)
        sec = time.perf_counter() - t0
        res = {"value": len(got) / sec, "unit": "sample-pairs/s", "cores": threads, "kind": "port",
               "sample": f"{len(got)} pairs of {sample_n} random samples, oracle FP32 walk, {threads} threads"}
    res["cpu_model"] = cpu_model()
    # accuracy of the GPU outputs on 2,000 pairs of the subset vs the oracle
    rng = np.random.default_rng(7)
    ii = rng.integers(0, sample_n, 4000)
    jj = rng.integers(0, sample_n, 4000)
    keep = ii != jj
    ii, jj = np.minimum(ii, jj)[keep][:2000].astype(np.int32), np.maximum(ii, jj)[keep][:2000].astype(np.int32)
    dense = abd.subset(pick).dense(tree.leaf_n)
    f64 = orc.pairs_f64(tree, dense, ii, jj, threads=threads)
    f32 = orc.pairs_f32(tree, dense, ii, jj, threads=threads)
    p = cms.packed_index(pick[ii], pick[jj], n_all)
    acc = {}
    if gpu_sm is not None:
        e = np.abs(gpu_sm[p].astype(np.float64) - f64)
        acc["sharedmin_max_abs_err_vs_fp64"] = float(e.max())
        acc["sharedmin_rms_err_vs_fp64"] = float(np.sqrt((e ** 2).mean()))
    if gpu_walk is not None:
        acc["walk_bitexact_vs_reference_fp32"] = bool(np.array_equal(gpu_walk[p].view(np.uint32),
                                                                     f32.view(np.uint32)))
        acc["walk_max_abs_err_vs_fp64"] = float(np.abs(gpu_walk[p].astype(np.float64) - f64).max())
    acc["reference_fp32_max_abs_err_vs_fp64"] = float(np.abs(f32.astype(np.float64) - f64).max())
    acc["pairs_checked"] = int(len(ii))
    return res, acc


def exact_pair_entries(tree, abd, sample_n, seed=9):
    """Mean number of entries (child edges) where both samples of a pair have
    mass, over every pair of a seeded random subset of sample_n samples: the
    work an exact sparse intersection would do per pair. Sample sets of the
    entries are propagated up the schedule as bitsets."""
    import numpy as np
    n = abd.n
    pick = np.sort(np.random.default_rng(seed).choice(n, min(sample_n, n), replace=False))
    sub = abd.subset(pick)
    m = sub.n
    words = (m + 63) // 64
    leaf = np.zeros((tree.leaf_n, words), np.uint64)
    rows = np.repeat(np.arange(m), np.diff(sub.rowptr))
    np.bitwise_or.at(leaf, (sub.col, rows // 64), np.uint64(1) << (rows % 64).astype(np.uint64))
    pop = np.array([bin(i).count("1") for i in range(1 << 16)], np.int64)

    def popcount(bs):
        v = bs.view(np.uint16)
        return int(pop[v].sum())

    slots = {}
    pairs_total = 0
    for i in range(tree.order_n):
        ch = []
        for c in (int(tree.order_1[i]), int(tree.order_2[i])):
            ch.append(leaf[c] if c >= 0 else slots.pop(c))
        for bs in ch:
            k = popcount(bs)
            pairs_total += k * (k - 1) // 2
        slots[int(tree.order_d[i])] = ch[0] | ch[1]
    return pairs_total / (m * (m - 1) / 2), m


def _wall(cmd, env):
    t0 = time.perf_counter()
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900)
    return time.perf_counter() - t0, r


def _cuda_comp_times(stdout):
    """The '*** load ... s, gpu ... s ... write ... s' line cuda-comp prints."""
    import re
    m = re.search(r"\*\*\* load ([0-9.]+) s, gpu(?: \+ write)? ([0-9.]+) s \(pair kernel ([0-9.]+) ms\)"
                  r"(?:, write ([0-9.]+) s)?", stdout)
    if not m:
        return {}
    out = {"load_s": float(m.group(1)), "gpu_s": float(m.group(2)), "pair_kernel_ms": float(m.group(3))}
    if m.group(4):
        out["write_s"] = float(m.group(4))
    return out


def end_to_end(cms, tree, cp, counts, threads, cpu_samples, with_cpu):
    """cuda-comp run as a program, files in and out, wall clock: the workload's
    10,000 samples from a sparse table file (parse, copy-number correction,
    upload, embedding, pair kernel, download, .dist text or packed binary
    written); and, on one random subset of cpu_samples samples in the
    reference's text table, cuda-comp beside the reference `comp` as its
    Makefile builds it (-O0; the -O3 build of comp_sam.cpp crashes at start)
    (parse, pair loop on the host cores, .dist written)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    exe = os.path.join(cms.BIN_DIR, "cuda-comp")
    res = {}
    with tempfile.TemporaryDirectory() as d:
        cms.write_db(d, "bench", tree, cp)
        env = dict(os.environ, ParallelMETA=d)
        sparse = os.path.join(d, "table.cmst")
        cms.write_table_csr(sparse, counts, tree.leaf_n)
        gpu = {"samples": counts.n, "input": "sparse table file", "input_bytes": os.path.getsize(sparse)}
        for fmt, ext in (("dist", "dist"), ("binary", "bin")):
            out = os.path.join(d, "out." + ext)
            sec, r = _wall([exe, "-T", sparse, "-D", "G", "-o", out, "--kernel", "sharedmin", "--format", fmt], env)
            if r.returncode != 0:
                log(f"cuda-comp {fmt} failed: {r.stderr[-300:]}")
                continue
            gpu[fmt] = {"wall_s": sec, "output_bytes": os.path.getsize(out), **_cuda_comp_times(r.stdout)}
            os.remove(out)
            log(f"end to end, cuda-comp --format {fmt}: {sec:.2f} s ({gpu[fmt]})")
        res["cuda_comp_sharedmin"] = gpu
        if with_cpu:
            comp = orc.ref_binary("comp_O0")
            pick = np.sort(np.random.default_rng(6).choice(counts.n, cpu_samples, replace=False))
            table = os.path.join(d, "small.txt")
            cms.write_table(table, counts.subset(pick), tree.leaf_n)
            side = {"samples": cpu_samples, "input": "text table", "input_bytes": os.path.getsize(table),
                    "sample": f"{cpu_samples} samples drawn at random (seed 6)"}
            for name, cmd in (("cuda_comp_walk", [exe, "-T", table, "-D", "G", "-o", os.path.join(d, "g.dist")]),
                              ("reference_comp_as_shipped_O0",
                               [comp, "-T", table, "-D", "G", "-o", os.path.join(d, "c.dist"), "-M", "0", "-t",
                                str(threads)] if comp else None)):
                if cmd is None:
                    continue
                sec, r = _wall(cmd, env)
                if r.returncode == 0:
                    side[name] = {"wall_s": sec, **(_cuda_comp_times(r.stdout) if "cuda" in name else {})}
                    log(f"end to end, {name} on {cpu_samples} samples: {sec:.2f} s")
                else:
                    log(f"end to end, {name} failed ({r.returncode}): {r.stderr[-300:]}")
            if "reference_comp_as_shipped_O0" in side and "cuda_comp_walk" in side:
                side["reference_threads"] = threads
                side["outputs_identical"] = open(os.path.join(d, "g.dist"), "rb").read() == \
                    open(os.path.join(d, "c.dist"), "rb").read()
            res["same_text_table"] = side
    return res


def roofline(st, kernel, traffic=None):
    """Roofline entry of a shared-minimum pair kernel from its work counters:
    VALU (algorithmic lane-ops executed: 17/16 per pair of a staged 32 x 32
    tile-row, one v_sad_u32 each plus one column-sum add per staged row and
    lane; one minimum and one add per pair with mass of a sparse or half
    entry) and HBM (algorithmic bytes: 2 x 32 x 4 B of Y per staged row, two
    16-byte records per sparse entry, 32 x 4 B of Y and one record per half
    entry), each over the kernel time from HIP events; `bound` is the one
    closer to its peak."""
    t = st.kernel_ms * 1e-3
    valu = st.valu_ops / t / 1e12 if t > 0 else 0.0
    hbm = st.bytes_moved / t / 1e9 if t > 0 else 0.0
    fv, fh = valu / VALU_PEAK_TOPS, hbm / HBM_PEAK_GBS
    main = ({"bound": "valu", "achieved": valu, "peak": VALU_PEAK_TOPS, "unit": "TOP/s", "frac": fv} if fv >= fh else
            {"bound": "hbm", "achieved": hbm, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": fh})
    main.update({"traffic": traffic, "kernel": kernel, "kernel_ms": st.kernel_ms,
                 "valu": {"achieved_tops": valu, "frac": fv, "ops": st.valu_ops},
                 "hbm": {"achieved_gbs": hbm, "frac": fh, "algorithmic_bytes": st.bytes_moved},
                 "pair_entries": st.pair_node_updates, "ops_per_staged_pair_entry": SM_OPS_PER_PAIR_ENTRY,
                 "tune_ms": list(st.tune_ms)})
    return main


def traffic_from_profiles():
    """HBM bytes per launch of the pair kernel, as measured by the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE runs (profiles/traffic_workload1.txt:
    a line 'traffic_bytes_per_launch <bytes>'), or None."""
    p = os.path.join(ROOT, "profiles", "traffic_workload1.txt")
    try:
        with open(p) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 2 and parts[0] == "traffic_bytes_per_launch":
                    return float(parts[1])
    except OSError:
        pass
    return None


def run_batched(cms, a):
    """Workload 3: 100,000 samples, nonzero OTUs log-uniform in [50, 20,000]
    (Zipf 1.0, NegBinomial(30, 0.3)), 99,322-leaf tree; all ~5e9 pairs
    streamed to pinned host memory in bands (binned samples, persistent
    kernel, autotuned variant, hipGraph-replayed band launches)."""
    import numpy as np
    t0 = time.perf_counter()
    tree, cp, counts, abd = cms.workload(3, samples=a.samples)
    del counts
    n = abd.n
    npairs = n * (n - 1) // 2
    log(f"workload 3: {n} samples, nnz {abd.rowptr[-1]}, generated in {time.perf_counter() - t0:.1f} s")
    ctx = cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, bin=a.bin, autotune=a.autotune, variant=a.variant,
                      use_graph=a.graph)
    ctx.set_tree(tree)
    ctx.set_samples_csr(abd)
    t1 = time.perf_counter()
    ctx.prepare(force=True)
    ctx.synchronize()
    prep_s = time.perf_counter() - t1
    got = {"pairs": 0, "bands": 0, "sum": 0.0}

    def fn(first, count, sims):
        got["pairs"] += count
        got["bands"] += 1
        got["sum"] += float(sims[:: max(1, count // 1024)].sum())
        log(f"band {got['bands']}: {got['pairs'] / npairs * 100:.1f} % of pairs")

    t2 = time.perf_counter()
    ctx.all_pairs_stream(fn, a.band_pairs)
    stream_s = time.perf_counter() - t2
    st = ctx.stats()
    # the same stream with the other band-launch mode (hipGraph replay or
    # direct launches), the tuned variant kept: what the graph buys
    other = None
    if a.graph_compare:
        ctx.set_opts(cms.KERNEL_SHAREDMIN, bin=a.bin, autotune=a.autotune, variant=st.variant,
                     use_graph=0 if a.graph else 1)
        got2 = {"pairs": 0}

        def fn2(first, count, sims):
            got2["pairs"] += count

        t3 = time.perf_counter()
        ctx.all_pairs_stream(fn2, a.band_pairs)
        other = {"hipgraph": not bool(a.graph), "stream_s": time.perf_counter() - t3,
                 "kernel_ms": ctx.stats().kernel_ms, "pairs_streamed": got2["pairs"]}
        log(f"stream with hipgraph={int(not a.graph)}: {other['stream_s']:.2f} s (kernel {other['kernel_ms']:.0f} ms)")
    res = {"metric": METRIC, "value": npairs / (stream_s + prep_s), "unit": "sample-pairs/s", "n_gpus": 1,
           "steps": 1, "warmup": 0, "ms_per_step": (stream_s + prep_s) * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "u32 fixed-point (FP64 embedding)",
           "data": "synthetic (seeded generator; tree 1003, samples 2003, copy numbers 3003)",
           "config": {"workload": "workload 3: 100,000 samples (log-uniform 50-20,000 OTUs) x 99,322-leaf tree, "
                                  "streamed to host", "samples": n, "pairs": npairs, "binned": bool(a.bin),
                      "variant": st.variant, "bands": st.bands, "band_pairs": a.band_pairs,
                      "hipgraph": bool(a.graph)},
           "prep_s": prep_s, "stream_s": stream_s, "pairs_streamed": got["pairs"],
           "kernel_ms": st.kernel_ms, "other_launch_mode": other,
           "roofline": roofline(st, "sm_persistent_split_kernel")}
    ctx.close()
    print(json.dumps(res), flush=True)


def run_precision(cms, a):
    """Workload 2: 50,000 samples of the workload-1 generator. The walk with
    abundances stored BF16 and F16 (per-sample power-of-two scale, FP32
    propagation, Kahan-compensated FP32 total, FP64 final division): all
    ~1.25e9 pairs timed, then max / RMS error vs the CPU FP64 oracle on random
    pairs (the oracle runs here only as the checker)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    t0 = time.perf_counter()
    tree, cp, counts, abd = cms.workload(2, samples=a.samples)
    del counts
    n = abd.n
    npairs = n * (n - 1) // 2
    log(f"workload 2: {n} samples, generated in {time.perf_counter() - t0:.1f} s")
    rng = np.random.default_rng(11)
    m = a.error_pairs
    ii = rng.integers(0, n, 2 * m)
    jj = rng.integers(0, n, 2 * m)
    keep = ii != jj
    ii, jj = np.minimum(ii, jj)[keep][:m].astype(np.int32), np.maximum(ii, jj)[keep][:m].astype(np.int32)
    # FP64 oracle over the sampled pairs, from the samples they touch
    used = np.unique(np.concatenate([ii, jj]))
    remap = np.full(n, -1, np.int64)
    remap[used] = np.arange(len(used))
    t1 = time.perf_counter()
    dense = abd.subset(used).dense(tree.leaf_n)
    f64 = orc.pairs_f64(tree, dense, remap[ii].astype(np.int32), remap[jj].astype(np.int32), threads=a.cpu_threads)
    del dense
    log(f"FP64 oracle on {len(ii)} pairs in {time.perf_counter() - t1:.1f} s")
    runs = {}
    for name, storage in (("bf16", cms.STORE_BF16), ("f16", cms.STORE_F16), ("f32", cms.STORE_F32)):
        with cms.Context(0, kernel=cms.KERNEL_WALK, storage=storage) as ctx:
            ctx.set_tree(tree)
            ctx.set_samples_csr(abd)
            sec = timed_steps(ctx, 1, 0)
            st = ctx.stats()
            sims = ctx.sim_pairs(ii, jj)
        e = np.abs(sims.astype(np.float64) - f64)
        runs[name] = {"ms_per_step": sec * 1e3, "pairs_per_s": npairs / sec, "kernel_ms": st.kernel_ms,
                      "max_abs_err_vs_fp64": float(e.max()), "rms_err_vs_fp64": float(np.sqrt((e ** 2).mean()))}
        log(f"walk {name}: {sec:.1f} s, max err {e.max():.3e}, rms {np.sqrt((e ** 2).mean()):.3e}")
    with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, bin=a.bin, autotune=a.autotune) as ctx:
        ctx.set_tree(tree)
        ctx.set_samples_csr(abd)
        sec = timed_steps(ctx, 1, 1)
        sims = ctx.sim_pairs(ii, jj)
    e = np.abs(sims.astype(np.float64) - f64)
    runs["sharedmin_u32"] = {"ms_per_step": sec * 1e3, "pairs_per_s": npairs / sec,
                             "max_abs_err_vs_fp64": float(e.max()), "rms_err_vs_fp64": float(np.sqrt((e ** 2).mean()))}
    log(f"sharedmin: {sec:.2f} s, max err {e.max():.3e}")
    res = {"metric": "max/RMS abs error vs FP64 on random pairs, and pairs/s", "workload": "workload 2: 50,000 samples "
           "x 99,322-leaf tree (~1.25e9 pairs)", "samples": n, "pairs": npairs, "error_pairs": int(len(ii)),
           "runs": runs}
    print(json.dumps(res), flush=True)


def run_serving(cms, a):
    """Workload 4: 200,000 resident references (workload-3 generator) on the
    99,322-leaf tree; query sets of 1-512 samples (log-uniform sizes, drawn
    from a pool of the same generator with another seed) arrive as a Poisson
    process and are compared with every reference through the serving loop
    (continuous batching over a band cycle, FIFO admission by query slots and
    device memory). Reports latency percentiles, achieved vs offered pairs/s
    and the cross kernel's share of VALU peak; a few pairs are checked
    against the FP64 oracle."""
    import threading
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    t0 = time.perf_counter()
    tree, cp, counts, refs = cms.workload(4, samples=a.samples)
    del counts
    w = cms.WORKLOADS[4]
    pool = cms.counts_to_abundance(cms.synth_counts(6004, a.query_pool, w["leaves"], w["nnz"][0], w["nnz"][1],
                                                    w["zipf"], w["nb"][0], w["nb"][1], w["log_uniform"]), cp)
    n_ref = refs.n
    log(f"workload 4: {n_ref} references (nnz {refs.rowptr[-1]}), query pool {pool.n}, generated in "
        f"{time.perf_counter() - t0:.1f} s")
    ctx = cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, bin=1)
    ctx.set_tree(tree)
    ctx.set_samples_csr(refs)
    t1 = time.perf_counter()
    srv = cms.Server(ctx, slots=a.slots, bands=a.bands, memory_budget=a.budget)
    prep_s = time.perf_counter() - t1
    log(f"references embedded and resident in {prep_s:.1f} s (slots {a.slots}, bands {a.bands})")
    # warm-up set (not measured)
    srv.submit(pool.subset(np.arange(64))).wait()
    st0 = srv.stats()

    rng = np.random.default_rng(4004)
    n_req = max(1, int(a.rate * a.serve_seconds))
    gaps = rng.exponential(1.0 / a.rate, n_req)
    sizes = np.minimum(512, np.floor(np.exp(rng.uniform(0.0, np.log(513.0), n_req)))).astype(np.int64)
    starts = rng.integers(0, pool.n, n_req)
    subsets = [pool.subset((starts[k] + np.arange(sizes[k])) % pool.n) for k in range(n_req)]
    reqs = [None] * n_req
    submitted = threading.Event()
    state = {"stopped_at": None}

    def submitter():
        t_start = time.perf_counter()
        t_next = t_start
        for k in range(n_req):
            t_next += gaps[k]
            dt = t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            outstanding = sum(1 for r in reqs[:k] if r is not None and r.info is None)
            if outstanding >= a.max_outstanding:
                state["stopped_at"] = k  # saturated: stop offering load
                break
            reqs[k] = srv.submit(subsets[k])
        submitted.set()

    th = threading.Thread(target=submitter)
    t_run = time.perf_counter()
    th.start()
    lat, queue_s, service_s, compute_s, deliver_s, pairs_done = [], [], [], [], [], 0
    checked = []
    k = 0
    while True:
        if k < n_req and reqs[k] is not None:
            r = reqs[k]
            r.wait()
            i = r.info
            lat.append(i.t_done - i.t_submit)
            queue_s.append(i.t_join - i.t_submit)
            service_s.append(i.t_done - i.t_join)
            if i.t_computed >= 0:
                compute_s.append(i.t_computed - i.t_join)
                deliver_s.append(i.t_done - i.t_computed)
            pairs_done += int(i.nq) * n_ref
            if len(checked) < 3:
                checked.append((k, r.out))
            else:
                r.out = None
            k += 1
            continue
        if submitted.is_set() and (k >= n_req or reqs[k] is None):
            break
        time.sleep(0.001)
    th.join()
    wall = time.perf_counter() - t_run
    st = srv.stats()
    kernel_ms = st.kernel_ms - st0.kernel_ms
    entries = st.pair_entries - st0.pair_entries
    staged = st.entries_staged - st0.entries_staged
    sparse = st.entries_sparse - st0.entries_sparse
    half = st.entries_half - st0.entries_half
    light_pairs = (st.sparse_pairs - st0.sparse_pairs) + (st.half_pairs - st0.half_pairs)
    # FP64 spot check of a few pairs per checked request
    errs = []
    for kk, out in checked:
        q = subsets[kk]
        qi = rng.integers(0, q.n, 16)
        ri = rng.integers(0, n_ref, 16)
        both = np.vstack([q.subset(qi).dense(tree.leaf_n), refs.subset(ri).dense(tree.leaf_n)])
        f64 = orc.pairs_f64(tree, both, np.arange(16, dtype=np.int32), np.arange(16, 32, dtype=np.int32))
        errs.append(float(np.abs(out[qi, ri].astype(np.float64) - f64).max()))
    srv.close()
    ctx.close()
    lat = np.array(lat)
    served = len(lat)
    mean_q = float(sizes[:served].mean()) if served else 0.0
    # algorithmic work of the cross tiles, counted as for the all-pairs
    # tiles: staged entries 17/16 ops per pair (1,024 pairs) and 2 x 128 B;
    # sparse entries 2 x 16 B, half entries 128 + 16 B; a minimum and an add
    # per pair of the light entries
    valu_ops = staged * 1024 * SM_OPS_PER_PAIR_ENTRY + 2.0 * light_pairs
    algo_bytes = staged * 256 + sparse * 32 + half * 144
    valu_tops = valu_ops / (kernel_ms * 1e-3) / 1e12 if kernel_ms > 0 else 0.0
    hbm_gbs = algo_bytes / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else 0.0
    res = {"metric": "serving: p50/p99 latency and query x reference pairs/s on 1 MI355X",
           "value": pairs_done / wall, "unit": "pairs/s", "n_gpus": 1, "higher_is_better": True,
           "dtype": "u32 fixed-point (FP64 embedding)",
           "data": "synthetic (seeded generator; references tree 1004 / samples 2004, queries 6004)",
           "config": {"workload": "workload 4: 200,000 resident references (log-uniform 50-20,000 OTUs) x "
                                  "99,322-leaf tree; query sets of 1-512 (log-uniform), Poisson arrivals",
                      "references": n_ref, "rate_per_s": a.rate, "seconds": a.serve_seconds, "slots": a.slots,
                      "bands": a.bands, "memory_budget": st.arena_bytes},
           "requests_offered": n_req, "requests_served": served, "saturated": state["stopped_at"] is not None,
           "stopped_offering_at": state["stopped_at"], "mean_queries_per_set": mean_q,
           "offered_pairs_per_s": a.rate * float(sizes.mean()) * n_ref,
           "latency_s": {"p50": float(np.percentile(lat, 50)) if served else None,
                         "p99": float(np.percentile(lat, 99)) if served else None,
                         "mean": float(lat.mean()) if served else None, "max": float(lat.max()) if served else None},
           "queue_s_p50": float(np.percentile(queue_s, 50)) if served else None,
           "queue_s_p99": float(np.percentile(queue_s, 99)) if served else None,
           "service_s_p50": float(np.percentile(service_s, 50)) if served else None,
           "service_s_p99": float(np.percentile(service_s, 99)) if served else None,
           "compute_s_p50_p99": [float(np.percentile(compute_s, 50)), float(np.percentile(compute_s, 99))]
           if compute_s else None,
           "delivery_s_p50_p99": [float(np.percentile(deliver_s, 50)), float(np.percentile(deliver_s, 99))]
           if deliver_s else None,
           "wall_s": wall, "prep_s": prep_s, "steps": st.steps - st0.steps, "kernel_ms": kernel_ms,
           "kernel_pairs_per_s": pairs_done / (kernel_ms * 1e-3) if kernel_ms > 0 else None,
           "arena_peak": st.arena_peak, "slots_peak": st.slots_peak,
           "max_abs_err_vs_fp64_spot": max(errs) if errs else None,
           "roofline": {"bound": "valu", "achieved": valu_tops, "peak": VALU_PEAK_TOPS, "unit": "TOP/s",
                        "frac": valu_tops / VALU_PEAK_TOPS, "traffic": None,
                        "kernel": "sm_cross_split_kernel" if sparse + half > 0 else "sm_cross_kernel",
                        "entries_visited": entries / 1024, "entries_staged": staged, "entries_sparse": sparse,
                        "entries_half": half, "valu_ops": valu_ops, "algorithmic_bytes": algo_bytes,
                        "hbm": {"achieved_gbs": hbm_gbs, "frac": hbm_gbs / HBM_PEAK_GBS}}}
    if hbm_gbs / HBM_PEAK_GBS > valu_tops / VALU_PEAK_TOPS:
        res["roofline"].update({"bound": "hbm", "achieved": hbm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": hbm_gbs / HBM_PEAK_GBS})
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--samples", type=int, default=None, help="override the workload's sample count")
    ap.add_argument("--walk-steps", type=int, default=1, help="timed steps of the FP32 walk drop-in (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--cpu-samples", type=int, default=800)
    ap.add_argument("--cpu-pairs", type=int, default=319600, help="about 10 s of reference work on 16 threads")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--e2e", type=int, default=1, help="time cuda-comp end to end (files in and out)")
    ap.add_argument("--exact-samples", type=int, default=400,
                    help="random samples whose pairs estimate the exact shared-entry count")
    ap.add_argument("--e2e-cpu-samples", type=int, default=400,
                    help="samples of the text table on which cuda-comp and the reference comp run end to end")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="also run a plain copy kernel of this many bytes (traffic-counter reference)")
    ap.add_argument("--bin", type=int, default=1, help="bin samples by sparsity (shared-minimum path)")
    ap.add_argument("--autotune", type=int, default=0,
                    help="autotune the persistent-kernel variant (default: the library's default variant, "
                         "the one the committed profiles measure)")
    ap.add_argument("--variant", type=int, default=0, help="force a persistent-kernel variant (1..16; 7..16 take nearly empty entries out of the staged path, "
                         "13..16 share each tile among 2 or 4 waves)")
    ap.add_argument("--workload", type=int, default=1, choices=[1, 2, 3, 4],
                    help="1: headline all-pairs; 2: precision (16-bit walk storage, 50,000 samples); "
                         "3: batched stream (100,000 samples of uneven sparsity); "
                         "4: serving (200,000 resident references, Poisson query sets)")
    ap.add_argument("--rate", type=float, default=20.0, help="query sets per second (workload 4)")
    ap.add_argument("--serve-seconds", type=float, default=30.0, help="seconds of arrivals (workload 4)")
    ap.add_argument("--slots", type=int, default=1024, help="resident query slots (workload 4)")
    ap.add_argument("--bands", type=int, default=16, help="reference bands per cycle (workload 4)")
    ap.add_argument("--budget", type=int, default=0, help="serving arena bytes, 0 = default (workload 4)")
    ap.add_argument("--query-pool", type=int, default=8192, help="query samples generated (workload 4)")
    ap.add_argument("--max-outstanding", type=int, default=400,
                    help="stop offering load beyond this many unfinished sets (workload 4)")
    ap.add_argument("--band-pairs", type=int, default=1 << 28, help="pairs per streamed band (workload 3)")
    ap.add_argument("--graph", type=int, default=1, help="replay streamed bands from a hipGraph (workload 3)")
    ap.add_argument("--graph-compare", type=int, default=1,
                    help="workload 3: stream again with the other launch mode (with / without hipGraph)")
    ap.add_argument("--error-pairs", type=int, default=1000000, help="random pairs checked vs FP64 (workload 2)")
    ap.add_argument("--metrics", type=str, default="",
                    help="workload 1: also time these comp -M metrics (e.g. 1,2,3,4,5) on the same samples")
    a = ap.parse_args()
    if a.cpu_pairs < 1 or a.cpu_samples < 2:
        ap.error("--cpu-pairs must be >= 1 and --cpu-samples >= 2 (--no-cpu skips the CPU baseline)")
    if a.gpus != 1:
        print("bench.py measures one GPU: every GPU box here has one, and multi-GPU runs are not part of "
              "this build (pairs shard without communication; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    # No torch here: libcms owns its device memory, and torch would load a
    # second HIP runtime into the process.
    cms = importlib.import_module("coordinated-meta-storms_amd")
    if cms.device_count() < 1:
        raise SystemExit("bench.py needs a HIP device: " + cms.last_error())
    if a.workload == 4:
        return run_serving(cms, a)
    if a.workload == 3:
        return run_batched(cms, a)
    if a.workload == 2:
        return run_precision(cms, a)

    t0 = time.perf_counter()
    tree, cp, counts, abd = cms.workload(1, samples=a.samples)
    n = abd.n
    npairs = n * (n - 1) // 2
    log(f"workload 1: {n} samples, {tree.leaf_n} leaves, nnz {abd.rowptr[-1]}, "
        f"generated in {time.perf_counter() - t0:.1f} s")
    # headline: shared-minimum kernel
    sm = cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, bin=a.bin, autotune=a.autotune, variant=a.variant)
    sm.set_tree(tree)
    sm.set_samples_csr(abd)
    sec = timed_steps(sm, a.steps, a.warmup)
    st = sm.stats()
    gpu_sm = sm.copy_result(npairs)
    log(f"sharedmin: {sec * 1e3:.2f} ms/step, pair kernel {st.kernel_ms:.2f} ms, prep {st.prep_ms:.2f} ms, "
        f"executed pair-entries {st.pair_node_updates:.3e}")
    sm.close()
    t_ex = time.perf_counter()
    exact_per_pair, exact_n = exact_pair_entries(tree, abd, a.exact_samples)
    exact = exact_per_pair * npairs
    work = {"exact_pair_entries_est": exact,
            "exact_per_pair": exact_per_pair,
            "exact_sample": f"every pair of {exact_n} samples drawn at random (seed 9)",
            "bitmap_visited_pair_entries": float(st.entries_visited) * 1024.0,
            "executed_pair_entries": st.pair_node_updates,
            "visited_vs_exact": float(st.entries_visited) * 1024.0 / exact,
            "executed_vs_exact": st.pair_node_updates / exact,
            "entries_visited": st.entries_visited, "entries_staged": st.entries_staged,
            "entries_sparse": st.entries_sparse, "entries_half": st.entries_half,
            "sparse_pairs": st.sparse_pairs, "half_pairs": st.half_pairs}
    log(f"exact pair-entries {exact:.3e} ({exact_per_pair:.0f} per pair, {time.perf_counter() - t_ex:.1f} s); "
        f"bitmap-visited / exact {work['visited_vs_exact']:.2f}, executed / exact {work['executed_vs_exact']:.2f}")

    others = {}
    for m in [int(x) for x in a.metrics.split(",") if x.strip()]:
        with cms.Context(0, kernel=cms.KERNEL_SHAREDMIN, metric=m, bin=a.bin, autotune=a.autotune) as oc:
            oc.set_tree(tree)
            oc.set_samples_csr(abd)
            osec = timed_steps(oc, 1, 1)
            ost = oc.stats()
        others[f"M{m}"] = {"ms_per_step": osec * 1e3, "pairs_per_s": npairs / osec, "kernel_ms": ost.kernel_ms,
                           "prep_ms": ost.prep_ms}
        log(f"-M {m}: {osec * 1e3:.2f} ms/step, kernel {ost.kernel_ms:.2f} ms, prep {ost.prep_ms:.2f} ms")
        if m >= 1:
            # the same metric in the reference's FP32 operation order (the
            # unweighted walk for -M 1, the per-pair merge for -M 2..5)
            with cms.Context(0, kernel=cms.KERNEL_WALK, metric=m) as oc:
                oc.set_tree(tree)
                oc.set_samples_csr(abd)
                osec = timed_steps(oc, 1, 1)
                ost = oc.stats()
            others[f"M{m}_reference_order"] = {"ms_per_step": osec * 1e3, "pairs_per_s": npairs / osec,
                                               "kernel_ms": ost.kernel_ms, "prep_ms": ost.prep_ms}
            log(f"-M {m} reference order: {osec * 1e3:.2f} ms/step, kernel {ost.kernel_ms:.2f} ms, "
                f"prep {ost.prep_ms:.2f} ms")

    walk = None
    gpu_walk = None
    if a.walk_steps > 0:
        wc = cms.Context(0, kernel=cms.KERNEL_WALK)
        wc.set_tree(tree)
        wc.set_samples_csr(abd)
        wsec = timed_steps(wc, a.walk_steps, 0)
        wst = wc.stats()
        gpu_walk = wc.copy_result(npairs)
        wc.close()
        w_ops = wst.valu_ops / (wst.kernel_ms * 1e-3) / 1e12
        walk = {"ms_per_step": wsec * 1e3, "pairs_per_s": npairs / wsec,
                "pair_node_updates_per_s": npairs * tree.order_n / wsec, "kernel_ms": wst.kernel_ms,
                "roofline": {"bound": "valu", "achieved": w_ops, "peak": VALU_PEAK_TOPS, "unit": "TOP/s",
                             "frac": w_ops / VALU_PEAK_TOPS, "traffic": None,
                             "ops_per_pair_node": 14}}
        log(f"walk: {wsec * 1e3:.1f} ms/step, kernel {wst.kernel_ms:.1f} ms, {w_ops:.2f} TOP/s")

    res = {
        "metric": METRIC,
        "value": npairs / sec,
        "unit": "sample-pairs/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 fixed-point (FP64 embedding)",
        "data": "synthetic (seeded generator; tree 1001, samples 2001, copy numbers 3001)",
        "config": {"workload": "workload 1: 10,000 samples x 99,322-leaf tree, all 49,995,000 pairs",
                   "samples": n, "leaves": tree.leaf_n, "pairs": npairs, "kernel": "sharedmin",
                   "binned": bool(a.bin), "variant": st.variant},
        "pair_node_updates_per_s": npairs * tree.order_n / sec,
        "roofline": roofline(st, "sm_persistent_split_kernel", traffic_from_profiles()),
        "work": work,
        "prep_ms": st.prep_ms,
        "plain_fp32_walk": walk,
    }
    if others:
        res["other_metrics"] = others
    if walk:
        res["speedup_vs_plain_fp32_walk"] = walk["ms_per_step"] / (sec * 1e3)
    if a.copy_calib > 0:
        with cms.Context(0) as cc:
            ms = cc.calibrate_copy(a.copy_calib, 5)
        res["copy_calibration"] = {"bytes": a.copy_calib, "ms": ms, "GBps": 2 * a.copy_calib / ms / 1e6}
        log(f"copy calibration: {a.copy_calib} B in {ms:.3f} ms ({2 * a.copy_calib / ms / 1e6:.0f} GB/s read+write)")
    if a.e2e:
        res["end_to_end"] = end_to_end(cms, tree, cp, counts, a.cpu_threads, min(a.e2e_cpu_samples, n),
                                       not a.no_cpu)
    if not a.no_cpu:
        cb, acc = cpu_baseline(cms, tree, cp, counts, abd, gpu_sm, gpu_walk, n, a.cpu_threads,
                               min(a.cpu_samples, n), a.cpu_pairs)
        res["cpu_baseline"] = cb
        res["accuracy"] = acc
        log(f"cpu baseline: {cb['value']:.1f} pairs/s ({cb['kind']}, {cb['cores']} threads); accuracy {acc}")
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
