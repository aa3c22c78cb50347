"""Benchmark: AST nodes/s to bit-exact RISC-V words on one MI355X.

One step = one full pass of the code generator (the seven stages of
rv_generator::process, src/codegen/rv_generator.cpp:129-148) over the
large_16m workload (one synthetic Pareas program, ~16M AST nodes, 20k
functions), with the DepthTree arrays resident in HBM and the words left in
HBM.  `value` = nodes * steps / wall time of the timed steps.

The JSON line also carries:
  roofline      the dominant kernel's algorithmic bytes per launch over its
                HIP-event duration on the library's stream, against 8 TB/s,
                and (latency) the allocation walk's step count times its
                measured time per step: the walk is a chain of dependent steps;
  cpu_baseline  the CPU oracle (a one-thread restatement of the reference
                generator) timed on the same input on this host, whose output
                also checks the GPU result word for word (bit_exact);
  stages_ms     per-stage device time;
  stages_roofline  each stage's algorithmic bytes (stage-boundary elements read
                and written once) over its device time, as a fraction of the
                8 TB/s HBM peak, and the measured HBM bytes of the committed
                rocprofv3 PMC summary over the same time where there is one;
  batched       the same metric for the batched pipeline on batch_4096 (4096
                programs, ~100M nodes, one pass) and stress_1024 (1024
                register/jump stress programs of ~64k nodes), each with
                sampled programs checked word for word against the oracle, and
                for batch_4096 a multi-core host baseline (the oracle on a pool
                of host threads, one program per task, over the first ~20M
                nodes of the batch, checked against the GPU's output);
  to_asm        the text output (rv_generator::to_asm) of the compiled words,
                formatted on the device, against the HBM peak and beside the
                oracle's restatement on a sample;
  serving       the request loop (ptl_server) under Poisson arrivals at each
                rate: p50/p99 submit-to-result latency, nodes/s, batches, and
                whether every result equals the program compiled alone.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload large_16m]
                    [--batch-steps K] [--no-batch] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "AST nodes/s to bit-exact RV64, batched vs plain GPU vs AVX2 host; p50/p99 lat"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_alg_bytes(name: str, n: int, w0: int, w1: int, w2: int, f: int) -> float | None:
    """Algorithmic (compulsory) HBM bytes of one launch of a kernel: each input
    element read once, each output element written once (u8 types, u32 else)."""
    table = {
        "k_pre12": 18 * n + 6 * n,          # read the six input arrays, write types, rtypes, node data
        "k_pre3": 1 * n,                    # read input data types
        "k_scan_count": 11 * n + 4 * n,     # read types, rtypes, parents, child idx; write locations
        "k_cnt_fix": 14 * n,                # read parents, types, child idx, rtypes, locations
        "k_scan_decl": 1 * n + 4 * f,
        "k_gen_regs": 22 * n + 4 * n,       # node arrays + locations; parent slots
        "k_gen_owner": 26 * n + 8 * w0,     # node arrays + slots; owner keys
        "k_gen_write": 26 * n + 8 * w0 + 20 * w0,
        "k_gen_ties": 26 * n + 8 * w0,
        "k_opt_init": 16 * w0 + 2 * w0,
        "k_opt_iter": 12 * w0 + 1 * w0,
        "k_opt_final": 4 * w0 + 2 * w0,
        "k_ra_walk": 16 * w0 * 7 // 8 + 2 * w0 * 7 // 8,
        "k_ra_lane0": 16 * w0 // 8 + 2 * w0 // 8,
        "k_scan_ra_count": 13 * w0 + 4 * w0,
        "k_ra_emit": 24 * w0 + 2 * w0 + 20 * w1,
        "k_scan_fj_size": 4 * w1 + 4 * w1,
        "k_fj_move": 24 * w1 + 4 * w2,
    }
    return table.get(name)


STAGES = ("preprocess", "isn_cnt", "isn_gen", "optimize", "regalloc", "fix_jumps", "postprocess")


def stage_alg_bytes(n: int, w0: int, w1: int, w2: int, f: int) -> dict:
    """Stage-boundary compulsory bytes of each stage: every stage input element
    read once and every output element written once (u8 types, 4 B else)."""
    return {"preprocess": 20 * n, "isn_cnt": 18 * n + 16 * f, "isn_gen": 22 * n + 8 * f + 20 * w0,
            "optimize": 20 * w0, "regalloc": 24 * w0 + 20 * w1 + 24 * f,
            "fix_jumps": 20 * w1 + 20 * w2 + 24 * f, "postprocess": 20 * w2}


def measured_stage_traffic(workload: str) -> dict:
    """HBM bytes per compile of each stage from the committed rocprofv3 PMC
    summary of `workload` (its per-stage table), if any."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{workload}.md")
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 3 and cells[0].startswith("stage "):
            try:
                out[cells[0][6:]] = float(cells[2].replace(",", ""))
            except ValueError:
                pass
    return out


def stage_roofline(stages_ms: dict, n: int, w0: int, w1: int, w2: int, f: int, workload: str) -> dict:
    """Each stage as a fraction of the HBM peak: algorithmic bytes over its
    device time, and (where profiled) the measured HBM bytes over the same time.
    fix_jumps and postprocess run as one fused kernel: they are reported together."""
    alg = stage_alg_bytes(n, w0, w1, w2, f)
    meas = measured_stage_traffic(workload)
    out = {}
    for st in ("preprocess", "isn_cnt", "isn_gen", "optimize", "regalloc", "fix_jumps+postprocess"):
        parts = st.split("+")
        ms = sum(stages_ms[p] for p in parts)
        a = sum(alg[p] for p in parts)
        rec = {"ms": ms, "alg_bytes": a, "alg_GBps": a / (ms * 1e-3) / 1e9,
               "alg_frac": a / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        if st in meas:
            rec["traffic_bytes"] = meas[st]
            rec["traffic_GBps"] = meas[st] / (ms * 1e-3) / 1e9
            rec["traffic_frac"] = rec["traffic_GBps"] / HBM_PEAK_GBS
        out[st] = rec
    return out


def host_cpu() -> dict:
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count()}


def cpu_batch_baseline(trees, res, max_nodes: int) -> dict:
    """The oracle on a pool of host threads, one program per task (programs are
    independent; ctypes releases the GIL), over the first programs of a batch up
    to max_nodes nodes, each checked word for word against the GPU's result."""
    from concurrent.futures import ThreadPoolExecutor
    import oracle
    oracle.lib()
    k, tot = 0, 0
    while k < len(trees) and tot < max_nodes:
        tot += trees[k].n
        k += 1
    threads = max(1, min((os.cpu_count() or 2) - 1, 15))
    keys = ("words", "func_ids", "func_starts", "func_ends", "func_sizes")
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        outs = list(ex.map(oracle.compile_tree, trees[:k]))
    el = time.perf_counter() - t0
    same = all(all(np.array_equal(res[p][q], outs[p][q]) for q in keys) for p in range(k))
    return {"value": tot / el, "unit": "nodes/s", "cores": threads, "kind": "port",
            "sample": f"first {k} programs of the batch ({tot} nodes), the oracle (C restatement of "
                      f"rv_generator_avx at one thread, AVX2 build) on {threads} host threads, one program per "
                      f"task, {el:.1f}s", "bit_exact": bool(same)}


def measured_traffic(kernel: str, workload: str) -> float | None:
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    of `workload` (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), if any."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{workload}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0].split("<")[0] == kernel:
            try:
                return float(cells[4].replace(",", ""))
            except ValueError:
                return None
    return None


def measure_asm(ctx, res: dict, with_cpu: bool) -> dict:
    """Text output (rv_generator::to_asm) of the last compile's words, formatted
    on the device: the length scan and the write kernel's HIP-event times, their
    algorithmic bytes (words read once, text written once) against 8 TB/s, the
    call's wall time (text copied to host memory included), and the oracle's
    restatement on a sample of the same words."""
    import oracle
    ctx.profile(True)
    ctx.result_asm(0)  # warm
    t0 = time.perf_counter()
    text = ctx.result_asm(0)
    wall = time.perf_counter() - t0
    kt = ctx.kernel_times()
    ctx.profile(False)
    words = res["n_words"]
    dev_ns = sum(v[0] for k, v in kt.items() if k.startswith("k_asm"))
    alg = 4 * words + len(text)
    out = {"words": words, "text_bytes": len(text), "device_ms": dev_ns / 1e6, "wall_ms": wall * 1e3,
           "kernels_ms": {k: v[0] / 1e6 for k, v in kt.items()},
           "alg_bytes": alg, "achieved_GBps": alg / (dev_ns * 1e-9) / 1e9,
           "frac": alg / (dev_ns * 1e-9) / 1e9 / HBM_PEAK_GBS, "words_per_s": words / (dev_ns * 1e-9)}
    if with_cpu and len(res["words"]):
        k = min(len(res["words"]), 2_000_000)
        c0 = time.perf_counter()
        want = oracle.to_asm(res["words"][:k])
        c1 = time.perf_counter()
        out["cpu_baseline"] = {"value": k / (c1 - c0), "unit": "words/s", "cores": 1, "kind": "port",
                               "sample": f"first {k} words, the oracle's to_asm restatement, {c1 - c0:.2f}s"}
        out["bit_exact"] = text[:len(want)] == want
    log(f"to_asm: {words} words -> {len(text)} bytes, device {dev_ns / 1e6:.2f} ms "
        f"({out['achieved_GBps']:.0f} GB/s algorithmic), wall {wall * 1e3:.1f} ms"
        + (f", host {out['cpu_baseline']['value'] / 1e6:.2f} M words/s, bit-exact {out['bit_exact']}"
           if "cpu_baseline" in out else ""))
    return out


def plain_vs_batched(ctx, trees, slot: int) -> dict:
    """The same programs through the plain one-program pipeline, one compile
    each (inputs resident in HBM, words left there), and as one batch."""
    nodes = sum(t.n for t in trees)
    plain = 0.0
    for t in trees:
        ctx.upload(slot, t)
        ctx.compile_resident(slot, copy_out=False)
        a = time.perf_counter()
        ctx.compile_resident(slot, copy_out=False)
        plain += time.perf_counter() - a
    ctx.upload_batch(slot, trees)
    ctx.compile_resident_batch(slot, copy_out=False, results=False)
    a = time.perf_counter()
    ctx.compile_resident_batch(slot, copy_out=False, results=False)
    batched = time.perf_counter() - a
    out = {"programs": len(trees), "nodes": nodes, "plain_nodes_per_s": nodes / plain,
           "batched_nodes_per_s": nodes / batched, "plain_ms": plain * 1e3, "batched_ms": batched * 1e3}
    log(f"  first {len(trees)} programs ({nodes} nodes): plain pipeline one by one {plain * 1e3:.1f} ms "
        f"({out['plain_nodes_per_s'] / 1e6:.1f} M nodes/s), one batch {batched * 1e3:.1f} ms "
        f"({out['batched_nodes_per_s'] / 1e6:.1f} M nodes/s)")
    return out


def measure_batch(ctx, name: str, slot: int, steps: int, warmup: int, n_check: int, cpu_nodes: int) -> dict:
    """Batched pipeline over one named batch workload, resident in HBM."""
    import oracle
    from ptilopsis_amd import workloads as W
    t0 = time.time()
    trees = W.batch_4096() if name == "batch_4096" else W.stress_1024()
    nodes = sum(t.n for t in trees)
    log(f"{name}: {len(trees)} programs, {nodes} nodes (generated+parsed in {time.time() - t0:.1f}s)")
    ctx.upload_batch(slot, trees)
    for i in range(warmup):
        ctx.compile_resident_batch(slot, copy_out=False, results=False)
    t1 = time.perf_counter()
    for i in range(steps):
        ctx.compile_resident_batch(slot, copy_out=False, results=False)
    el = time.perf_counter() - t1
    ctx.profile(True)
    res = ctx.compile_resident_batch(slot, copy_out=True)
    kt = ctx.kernel_times()
    ctx.profile(False)
    value = nodes * steps / el
    log(f"{name}: {el * 1e3 / steps:.1f} ms/step, {value / 1e6:.1f} M nodes/s")
    log("kernels (ms total, launches): " + ", ".join(
        f"{k}={v[0] / 1e6:.3f}/{v[1]}" for k, v in sorted(kt.items(), key=lambda kv: -kv[1][0])[:12]))
    rng = np.random.default_rng(0x5EED)
    pick = sorted(set(rng.choice(len(trees), min(n_check, len(trees)), replace=False).tolist()
                      + [int(np.argmax([t.n for t in trees]))]))
    ok = True
    for p in pick:
        want = oracle.compile_tree(trees[p])
        ok &= all(np.array_equal(res[p][k], want[k]) for k in ("words", "func_ids", "func_starts", "func_ends",
                                                                  "func_sizes"))
    log(f"{name}: {len(pick)} sampled programs bit-exact: {ok}")
    stages = {k: v / 1e6 for k, v in zip(STAGES, res[0]["stage_ns"])}
    sroof = stage_roofline(stages, nodes, sum(r["W0"] for r in res), sum(r["W1"] for r in res),
                           sum(r["n_words"] for r in res), sum(r["n_funcs"] for r in res), name)
    out = {"workload": name, "value": value, "unit": "nodes/s", "ms_per_step": el * 1e3 / steps, "steps": steps,
           "programs": len(trees), "nodes": nodes, "words": int(sum(r["n_words"] for r in res)),
           "sampled_bit_exact": ok, "sampled_programs": pick, "stages_ms": stages, "stages_roofline": sroof,
           "top_kernels_ms": {k: v[0] / 1e6 for k, v in sorted(kt.items(), key=lambda kv: -kv[1][0])[:6]}}
    out["plain_vs_batched"] = plain_vs_batched(ctx, trees[:256], slot)
    if cpu_nodes:
        out["cpu_baseline"] = cpu_batch_baseline(trees, res, cpu_nodes)
        log(f"{name}: host baseline {out['cpu_baseline']['value'] / 1e6:.2f} M nodes/s on "
            f"{out['cpu_baseline']['cores']} threads ({out['cpu_baseline']['sample']}), "
            f"bit-exact: {out['cpu_baseline']['bit_exact']}")
    del trees, res
    return out


OUT_FIELDS = ("words", "func_ids", "func_starts", "func_ends", "func_sizes")


def first_difference(got: dict, want: dict):
    """None when every output array is equal, else the first differing field,
    its first differing index and the two values there (-1: past the end)."""
    for k in OUT_FIELDS:
        g, w = np.asarray(got[k]), np.asarray(want[k])
        if g.shape == w.shape and np.array_equal(g, w):
            continue
        m = min(g.shape[0], w.shape[0])
        idx = np.nonzero(g[:m] != w[:m])[0]
        i = int(idx[0]) if idx.size else m
        return {"field": k, "index": i, "got": int(g[i]) if i < g.shape[0] else -1,
                "want": int(w[i]) if i < w.shape[0] else -1, "got_len": int(g.shape[0]), "want_len": int(w.shape[0])}
    return None


def measure_serving(rates, seconds: float, pool_n: int, against_oracle: bool = False, records_rates=()) -> dict:
    """Serving stream: Poisson arrivals at each rate for `seconds`, requests
    drawn in order from a pool of programs of log-normal size, submitted to the
    library's request loop (ptl_server) from a host thread as they arrive.  Each
    result is checked against the same program compiled alone; a mismatch is
    recorded with its ticket, batch, worker, the batch's members and the first
    differing word.  At records_rates the same requests go in as parser records
    (ptl_server_submit_arena: the tree is built on the device inside the batch,
    no flattening on the host)."""
    import threading
    from ptilopsis_amd import workloads as W
    from ptilopsis_amd.generator import Context, Server
    t0 = time.time()
    threads = max(1, min(16, os.cpu_count() or 1) - 1)
    pool = W.serve_pool(pool_n, threads=threads)
    pool_rec = W.serve_pool_records(pool_n, threads=threads) if records_rates else None
    log(f"serving pool: {pool_n} programs, {sum(t.n for t in pool)} nodes (generated+parsed in {time.time() - t0:.1f}s "
        f"on {threads} threads)")
    t0 = time.time()
    if against_oracle:
        # every result is checked against the oracle's output for its program,
        # computed once per pool program on host threads
        import oracle
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(threads) as ex:
            ref = list(ex.map(lambda t: {k: v for k, v in oracle.compile_tree(t).items() if k in OUT_FIELDS}, pool))
        ref_kind = f"the oracle ({threads} host threads, {time.time() - t0:.1f}s)"
    else:
        ctx = Context(0)
        ref = [{k: v for k, v in ctx.compile(t).items() if k in OUT_FIELDS} for t in pool]
        ctx.close()
        ref_kind = "the same program compiled alone"
    log(f"serving references: {ref_kind}")
    srv = Server(0, hbm_budget=48 << 30, max_batch=4096, workers=2)
    for k in [srv.submit(t) for t in pool[:64]]:  # workers' arenas warm
        srv.wait(k)

    def run_rate(rate, records: bool) -> dict:
        arr = W.arrivals(rate, seconds)
        tickets = [None] * len(arr)

        def submitter():
            tb = time.perf_counter()
            for i, ta in enumerate(arr):
                while time.perf_counter() - tb < ta:
                    time.sleep(0.00005)
                tickets[i] = srv.submit_records(pool_rec[i % pool_n]) if records else srv.submit(pool[i % pool_n])

        st0 = srv.stats()
        th = threading.Thread(target=submitter)
        tb = time.perf_counter()
        th.start()
        lat, nodes, mism = [], 0, []
        phases = {k: [] for k in ("queue_ns", "pack_ns", "run_ns", "device_ns", "copyout_ns")}
        members = {}  # batch id -> [(ticket, pool index)]
        for i in range(len(arr)):
            while tickets[i] is None:
                time.sleep(0.00005)
            r, tm = srv.wait(tickets[i])
            lat.append(tm["total_ns"])
            for k in phases:
                phases[k].append(tm[k])
            members.setdefault(tm["batch_id"], []).append((tickets[i], i % pool_n))
            d = first_difference(r, ref[i % pool_n])
            if d is not None:
                d.update(ticket=tickets[i], request=i, pool_index=i % pool_n, nodes=pool[i % pool_n].n,
                         batch_id=tm["batch_id"], worker=tm["worker"], batch_programs=tm["batch_programs"])
                mism.append(d)
                log(f"serving MISMATCH: {d}")
            nodes += pool[i % pool_n].n
        th.join()
        el = time.perf_counter() - tb
        st1 = srv.stats()
        nb = st1["batches"] - st0["batches"]
        for d in mism:
            d["batch_members"] = members.get(d["batch_id"], [])
        pct = lambda v, q: float(np.percentile(v, q)) / 1e6
        rec = {"rate_offered": rate, "input": "records" if records else "arrays", "requests": len(arr),
               "seconds": el, "rate_achieved": len(arr) / max(1e-9, float(arr[-1]) if len(arr) else 1.0),
               "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "nodes_per_s": nodes / el,
               "batches": nb, "mean_batch": len(arr) / max(1, nb),
               "breakdown_ms": {k[:-3]: {"p50": pct(v, 50), "p99": pct(v, 99)} for k, v in phases.items()},
               "same_as_alone": not mism, "mismatches": mism}
        b = rec["breakdown_ms"]
        log(f"serving {rate} req/s ({rec['input']}): p50 {rec['p50_ms']:.2f} ms, p99 {rec['p99_ms']:.2f} ms, "
            f"{rec['nodes_per_s'] / 1e6:.1f} M nodes/s, {nb} batches (mean {rec['mean_batch']:.1f}), "
            f"mismatches: {len(mism)}; p50 queue {b['queue']['p50']:.2f} pack {b['pack']['p50']:.2f} "
            f"run {b['run']['p50']:.2f} (device {b['device']['p50']:.2f}) copyout {b['copyout']['p50']:.2f} ms")
        return rec

    out = [run_rate(rate, False) for rate in rates]
    out += [run_rate(rate, True) for rate in records_rates]
    srv.close()
    return {"pool_programs": pool_n, "pool_nodes": sum(t.n for t in pool), "seconds_per_rate": seconds,
            "workers": 2, "checked_against": ref_kind, "rates": out}


def workload_source(name: str) -> str:
    from ptilopsis_amd import workloads as W
    if name == "large_16m":
        return W.large_16m_source()
    if name == "single_50k":
        for attempt in range(64):
            src, n = W.program_source(W.SEED_SINGLE + (attempt << 40), 64)
            if 45_000 <= n <= 55_000:
                return src
    raise SystemExit(f"unknown workload {name}")


TREE_BYTES_PER_NODE = 36  # records read (types 1+1, data 4, first child 4, count 4, kid 4) + six arrays written (18)


def measure_tree(ctx, src: str, tree, steps: int, with_cpu: bool) -> dict:
    """DepthTree construction on the device from the parser's node records
    (ptl_upload_arena: records copied to HBM, the tour ranked, the six arrays
    written) against the host front end's flattening of the same records
    (what depthtree.cpp:11-100 does), and bit-exactness against the arrays the
    pipeline ran on."""
    from ptilopsis_amd.native import flatten_records, parse_arena
    t0 = time.perf_counter()
    ar = parse_arena(src)
    t1 = time.perf_counter()
    ctx.upload_arena(7, ar)
    ctx.upload_arena(7, ar)
    w0 = time.perf_counter()
    for _ in range(steps):
        ctx.upload_arena(7, ar)
    w1 = time.perf_counter()
    ctx.profile(True)
    ctx.upload_arena(7, ar)
    kt = ctx.kernel_times()
    ctx.profile(False)
    dev_ns = sum(v[0] for v in kt.values())
    got = ctx.tree_build(ar)[0]
    bit_exact = all(np.array_equal(a, b) for a, b in zip(got.arrays(), tree.arrays())) and got.max_depth == tree.max_depth
    alg = TREE_BYTES_PER_NODE * tree.n
    out = {"nodes": tree.n, "device_ms": dev_ns / 1e6, "wall_ms_with_h2d": (w1 - w0) * 1e3 / steps,
           "alg_bytes": alg, "achieved_GBps": alg / (dev_ns * 1e-9) / 1e9,
           "frac": alg / (dev_ns * 1e-9) / 1e9 / HBM_PEAK_GBS,
           "kernels_ms": {k: v[0] / 1e6 for k, v in sorted(kt.items(), key=lambda kv: -kv[1][0])},
           "host_parse_to_records_ms": (t1 - t0) * 1e3, "bit_exact": bit_exact}
    if with_cpu:
        tm = []
        flatten_records(ar, tm)
        out["cpu_baseline"] = {"value": tree.n / tm[0], "unit": "nodes/s", "cores": 1, "kind": "port",
                               "ms": tm[0] * 1e3,
                               "sample": "the same records flattened by the host front end (ptl_flatten_records, "
                                         "a restatement of depthtree.cpp:52-100), one thread"}
    log(f"tree build: {tree.n} nodes, device {out['device_ms']:.3f} ms ({out['achieved_GBps']:.0f} GB/s, "
        f"frac {out['frac']:.3f}), with H2D of the records {out['wall_ms_with_h2d']:.2f} ms"
        + (f", host flatten {out['cpu_baseline']['ms']:.1f} ms" if with_cpu else "") + f", bit-exact {bit_exact}")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="large_16m")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--batches", default="batch_4096,stress_1024",
                    help="batch workloads to measure (comma-separated; for profiling one of them)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--overlap", type=int, default=1, choices=(0, 1),
                    help="one-program allocation: walk and lockstep at the same time (1) or one after the other "
                         "(0; for counter collection, which serializes kernels)")
    ap.add_argument("--serve-rates", default="500,2000,5000,20000")
    ap.add_argument("--serve-seconds", type=float, default=2.5)
    ap.add_argument("--serve-records-rates", default="5000,20000",
                    help="rates at which the requests are also submitted as parser records (tree built on the device)")
    ap.add_argument("--serve-pool", type=int, default=384)
    ap.add_argument("--serve-oracle", action="store_true",
                    help="check every served result against the oracle (computed once per pool program) instead of "
                         "the same program compiled alone")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy kernel of BYTES bytes three times (to scale "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration // 4
        a = torch.arange(n, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 4 * n, "launches": 3}), flush=True)
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: the code generator has no multi-GPU path "
              "(it would shard by program, whole programs per device, with no collectives)", file=sys.stderr)
        return 2

    from ptilopsis_amd.generator import Context

    from ptilopsis_amd import parse_source
    t0 = time.time()
    src = workload_source(args.workload)
    tree = parse_source(src)
    log(f"workload {args.workload}: {tree.n} nodes, {tree.n_functions} functions, depth {tree.max_depth} "
        f"(generated+parsed in {time.time() - t0:.1f}s)")
    ctx = Context(0)
    ctx.set_overlap(args.overlap)
    ctx.upload(0, tree)
    for i in range(args.warmup):
        r = ctx.compile_resident(0, copy_out=False)
        log(f"warmup {i}: {sum(r['stage_ns']) / 1e6:.1f} ms device")
    t1 = time.perf_counter()
    for i in range(args.steps):
        r = ctx.compile_resident(0, copy_out=False)
    t2 = time.perf_counter()
    elapsed = t2 - t1
    value = tree.n * args.steps / elapsed
    log(f"timed {args.steps} steps: {elapsed * 1e3 / args.steps:.2f} ms/step, {value / 1e6:.1f} M nodes/s")

    # one more pass with per-kernel events and the words copied back
    ctx.profile(True)
    res = ctx.compile_resident(0, copy_out=True)
    ktimes = ctx.kernel_times()
    ctx.profile(False)
    stages = {k: v / 1e6 for k, v in zip(STAGES, res["stage_ns"])}
    sroof = stage_roofline(stages, tree.n, res["W0"], res["W1"], res["n_words"], res["n_funcs"], args.workload)
    log("stages (ms, fraction of HBM peak by algorithmic bytes): " + ", ".join(
        f"{k}={v['ms']:.2f}/{v['alg_frac']:.4f}" for k, v in sroof.items()))
    dom = max(ktimes.items(), key=lambda kv: kv[1][0])
    dname, (dns, dlaunch) = dom
    alg = kernel_alg_bytes(dname, tree.n, res["W0"], res["W1"], res["n_words"], res["n_funcs"])
    avg_ns = dns / max(1, dlaunch)
    achieved = (alg / (avg_ns * 1e-9) / 1e9) if alg else None
    roof = {"kernel": dname, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": measured_traffic(dname, args.workload),
            "avg_launch_us": avg_ns / 1e3, "launches": dlaunch, "alg_bytes_per_launch": alg}
    # the allocation lockstep is a chain of dependent steps: its second bound is
    # steps x time per step
    rows = int(ctx.dump("ra_rows").max())
    lane0 = ktimes.get("k_ra_lane0", (0, 1))
    roof["latency"] = {"kernel": "k_ra_lane0", "steps": rows, "maxsize_words": int(ctx.dump("ra_maxsize").max()),
                       "us_per_step": lane0[0] / 1e3 / max(1, rows)}
    log("kernels (ms total, launches): " + ", ".join(
        f"{k}={v[0] / 1e6:.3f}/{v[1]}" for k, v in sorted(ktimes.items(), key=lambda kv: -kv[1][0])))

    to_asm = measure_asm(ctx, res, not args.no_cpu_baseline)

    # kernel-only (sum of the stages' device times) and end-to-end (host arrays
    # in, words and function table back: ptl_compile) rates of the same compile,
    # and the device's measured copy bandwidth beside the published HBM peak
    ctx.compile(tree)
    e0 = time.perf_counter()
    ctx.compile(tree)
    e1 = time.perf_counter()
    rates = {"kernel_only_nodes_per_s": tree.n / (sum(res["stage_ns"]) * 1e-9),
             "end_to_end_nodes_per_s": tree.n / (e1 - e0), "end_to_end_ms": (e1 - e0) * 1e3}
    dev = ctx.copy_bandwidth(1 << 30, 5)
    roof["measured_copy_GBps"] = dev["GBps"]
    roof["frac_of_measured_copy"] = (achieved / dev["GBps"]) if achieved else None
    log(f"kernel-only {rates['kernel_only_nodes_per_s'] / 1e6:.1f} M nodes/s, end-to-end (H2D + compile + D2H) "
        f"{rates['end_to_end_nodes_per_s'] / 1e6:.1f} M nodes/s; device copy {dev['GBps']:.0f} GB/s "
        f"({dev['compute_units']} CUs, {dev['hbm_bytes'] / 2**30:.0f} GiB)")

    cpu = None
    bit_exact = None
    if not args.no_cpu_baseline:
        import oracle
        c0 = time.perf_counter()
        want = oracle.compile_tree(tree)
        c1 = time.perf_counter()
        cpu = {"value": tree.n / (c1 - c0), "unit": "nodes/s", "cores": 1, "kind": "port",
               "sample": f"full {args.workload} input ({tree.n} nodes), one run of the oracle (C restatement of "
                         f"rv_generator_avx at one thread, AVX2 build), {c1 - c0:.1f}s", "host": host_cpu()}
        bit_exact = all(np.array_equal(res[k], want[k]) for k in
                        ("words", "func_ids", "func_starts", "func_ends", "func_sizes"))
        log(f"cpu oracle: {c1 - c0:.1f}s; bit-exact: {bit_exact}")

    tree_build = measure_tree(ctx, src, tree, 5, not args.no_cpu_baseline)
    del src

    batched = None
    if not args.no_batch:
        batched = [measure_batch(ctx, nm, 1 + k, args.batch_steps, 1, 8,
                                 0 if args.no_cpu_baseline or nm != "batch_4096" else 20_000_000)
                   for k, nm in enumerate(x for x in args.batches.split(",") if x)]

    serving = None
    if not args.no_serve:
        serving = measure_serving([float(x) for x in args.serve_rates.split(",")], args.serve_seconds,
                                  args.serve_pool, args.serve_oracle,
                                  [float(x) for x in args.serve_records_rates.split(",") if x])

    line = {
        "metric": METRIC, "value": value, "unit": "nodes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": args.workload, "nodes": tree.n, "functions": tree.n_functions,
                   "max_depth": tree.max_depth, "words": res["n_words"], "path": "plain one-program GPU pipeline",
                   "inputs": "resident in HBM"},
        "roofline": roof, "cpu_baseline": cpu, "bit_exact": bit_exact, "stages_ms": stages,
        "stages_roofline": sroof, "rates": rates, "device": dev, "to_asm": to_asm, "tree_build": tree_build,
        "batched": batched, "serving": serving,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
