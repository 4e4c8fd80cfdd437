"""Benchmark: decompressed GB/s of one gzip member on one MI355X.

Workload (default): 1 GiB from the text generator (order-1 Markov chain over
96 printable bytes, Zipf(1.2) rows, 35 % copies of U[3,258] bytes from
log-uniform distances <= 32 KiB), compressed at level 6 into one RFC 1952
member (libdeflate stand-in encoder), decoded by speculative parallel decode
(pieces planned by decode cost, tuning picked by the autotuner).  Input and
output stay resident in HBM; a step is one full decode (find, inflate,
resolve, CRC32 + ISIZE verify).

Prints ONE JSON line (see the keys below).  `roofline` uses the dominant
kernel (the step's one k_inflate launch): the member's algorithmic bytes
C + D (compressed bytes read once, decompressed bytes written once) over its
HIP-event duration.  `cpu_baseline` is the
oracle's serial port (one core) on the same member.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "decompressed GB/s on 1 MI355X (1 GiB member, 10k-member batch); p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def host_cpu():
    """(cores this process may use, CPU model) of the host running the bench."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return cores, model


def _median_time(fn, budget_s, max_runs):
    times = []
    t_start = time.time()
    while len(times) < max_runs and (not times or time.time() - t_start < budget_s):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return sorted(times)[len(times) // 2], len(times)


def cpu_baseline(gz, raw, budget_s=6.0, max_runs=3):
    """gzippy's CPU path, restated in C (oracle/), timed on this host's cores
    on the same member, every output checked against the generator bytes:
    - value: the threaded speculative restatement (oracle/gzpar.c, the
      ParallelSM pipeline of chunk_fetcher.rs:638-1600) at all host cores;
    - t1: the serial restatement (drive_thin_t1_oracle, one core);
    - zlib and libdeflate single-thread decodes as independent references."""
    import ctypes
    import zlib

    import numpy as np
    from oracle import gzoracle
    cores, model = host_cpu()
    n_raw = raw.size
    out = np.empty(n_raw + 65536, dtype=np.uint8)
    L = gzoracle.lib()
    n = ctypes.c_size_t()
    m = ctypes.c_size_t()

    def par():
        st = L.gzo_gunzip_parallel(gz.ctypes.data, gz.size, out.ctypes.data, out.size, ctypes.byref(n), cores,
                                   512 << 10)
        assert st == 0 and n.value == n_raw

    def t1():
        st = L.gzo_gunzip(gz.ctypes.data, gz.size, out.ctypes.data, out.size, ctypes.byref(n), ctypes.byref(m))
        assert st == 0 and n.value == n_raw

    rec = {"unit": "GB/s", "cores": cores, "kind": "port", "cpu_model": model}
    t, runs = _median_time(par, budget_s, max_runs)
    assert np.array_equal(out[:n_raw], raw)
    rec["value"] = round(n_raw / t / 1e9, 4)
    rec["sample"] = (f"the full {n_raw} B member, threaded speculative restatement (oracle/gzpar.c: 512 KiB "
                     f"partitions, block finder, marker decode, window chain, parallel resolve + CRC combine) on "
                     f"{cores} threads, median of {runs} runs")
    t, runs = _median_time(t1, budget_s, max_runs)
    rec["t1"] = {"value": round(n_raw / t / 1e9, 4), "cores": 1, "runs": runs,
                 "what": "serial restatement (oracle/gzoracle.c gzo_gunzip)"}
    mv = memoryview(gz)
    t, runs = _median_time(lambda: zlib.decompress(mv, 31), budget_s, 1)
    rec["zlib"] = {"value": round(n_raw / t / 1e9, 4), "cores": 1, "runs": runs,
                   "version": zlib.ZLIB_RUNTIME_VERSION}
    try:
        ld = ctypes.CDLL("libdeflate.so.0")
        ld.libdeflate_alloc_decompressor.restype = ctypes.c_void_p
        ld.libdeflate_gzip_decompress.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                                  ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        ld.libdeflate_free_decompressor.argtypes = [ctypes.c_void_p]
        d = ld.libdeflate_alloc_decompressor()

        def ldd():
            st = ld.libdeflate_gzip_decompress(d, gz.ctypes.data, gz.size, out.ctypes.data, out.size,
                                               ctypes.byref(n))
            assert st == 0 and n.value == n_raw

        t, runs = _median_time(ldd, budget_s, max_runs)
        ld.libdeflate_free_decompressor(d)
        rec["libdeflate"] = {"value": round(n_raw / t / 1e9, 4), "cores": 1, "runs": runs}
    except OSError:
        rec["libdeflate"] = None
    return rec


def cpu_members_baseline(ins, raws, cores):
    """Member-per-core CPU baseline for a batch (MultiMemberPar, bgzf.rs:2370):
    every member decoded by the serial restatement, members spread over the
    host's cores; returns GB/s over all members."""
    import ctypes
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np
    from oracle import gzoracle
    L = gzoracle.lib()
    outs = [np.empty(r.size + 64, dtype=np.uint8) for r in raws]

    def one(i):
        n = ctypes.c_size_t()
        m = ctypes.c_size_t()
        st = L.gzo_gunzip(ins[i].ctypes.data, ins[i].size, outs[i].ctypes.data, outs[i].size, ctypes.byref(n),
                          ctypes.byref(m))
        return st == 0 and n.value == raws[i].size

    with ThreadPoolExecutor(cores) as ex:
        t0 = time.perf_counter()
        ok = all(ex.map(one, range(len(ins))))
        t = time.perf_counter() - t0
    ok = ok and all(np.array_equal(outs[i][:raws[i].size], raws[i]) for i in range(0, len(raws), 97))
    D = sum(r.size for r in raws)
    return {"value": round(D / t / 1e9, 4), "unit": "GB/s", "cores": cores, "kind": "port", "ok": bool(ok),
            "what": "member per core: oracle/gzoracle.c gzo_gunzip on each member, a thread pool of `cores`"}


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_inflate_hbm_traffic_1gib_text_l6.md")


def measured_traffic(args, kernel):
    """HBM bytes per launch of `kernel`, from the committed FETCH_SIZE /
    WRITE_SIZE summary of this exact workload (default size and level)."""
    if (args.size_mib, args.level) != (1024, 6):
        return None
    got = {}
    try:
        with open(TRAFFIC_PROFILE) as f:
            for ln in f:
                if ":" in ln and ln.split(":", 1)[0] in ("kernel", "traffic_bytes_per_launch"):
                    got[ln.split(":", 1)[0]] = ln.split(":", 1)[1].strip()
    except OSError:
        return None
    if got.get("kernel") != kernel or "traffic_bytes_per_launch" not in got:
        return None
    return {"bytes_per_launch": int(got["traffic_bytes_per_launch"]), "source": os.path.relpath(TRAFFIC_PROFILE, ROOT)}


def host_leg(gz, n, ctx, runs=3):
    """Host buffers in and out (gzgpu_decompress): H2D copy of the member,
    the decode, D2H through pinned staging chunks into a caller buffer.
    PCIe-inclusive; not the headline value."""
    import ctypes

    import numpy as np
    from gzippy_amd import _native
    a = np.ascontiguousarray(gz)
    out = np.ones(n + 64, dtype=np.uint8)  # touched: no first-touch faults inside the timing
    got = ctypes.c_size_t()
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        st = ctx.lib.gzgpu_decompress(ctx.h, a.ctypes.data, a.size, out.ctypes.data, out.size, ctypes.byref(got), None)
        times.append(time.perf_counter() - t0)
        if st != _native.OK or got.value != n:
            log(f"host leg: status {st}, {got.value} bytes")
            return None
    best = min(times)
    # the writer form: time to the first bytes at the writer, and to the end
    first, t_all = [], []
    for _ in range(runs):
        seen = []
        t0 = time.perf_counter()

        @_native.WRITE_FN
        def cb(user, ptr, ln):
            if not seen:
                seen.append(time.perf_counter())
            return 0

        tot = ctypes.c_uint64()
        st = ctx.lib.gzgpu_decompress_to_writer(ctx.h, a.ctypes.data, a.size, cb, None, ctypes.byref(tot), None)
        t_all.append(time.perf_counter() - t0)
        if st != _native.OK or tot.value != n or not seen:
            log(f"writer leg: status {st}, {tot.value} bytes")
            return None
        first.append(seen[0] - t0)
    return {"api": "gzgpu_decompress (host in, host out)", "runs": runs, "best_ms": round(best * 1e3, 2),
            "gbs_decompressed": round(n / best / 1e9, 2), "note": "includes H2D and D2H over the host link",
            "writer": {"api": "gzgpu_decompress_to_writer", "first_bytes_ms": round(min(first) * 1e3, 2),
                       "total_ms": round(min(t_all) * 1e3, 2),
                       "note": "first 8 MiB reach the writer while later output ranges still resolve"}}


def batch_leg(args, dev, ctx):
    """10,000 independent members (lognormal sizes around 64 KiB, four content
    kinds, levels 1/6/9): one batched call versus one call per member."""
    import ctypes
    import numpy as np
    import torch

    import gzippy_amd
    from gzippy_amd import workloads as W

    t0 = time.time()
    mem = W.batch_members(args.batch_members, seed=args.seed + 3)
    raws = [m[0] for m in mem]
    ins = [m[1] for m in mem]
    kinds = [m[2] for m in mem]
    levels = [m[3] for m in mem]
    in_len = np.array([x.size for x in ins], dtype=np.uint64)
    in_off = np.concatenate([[0], np.cumsum(in_len)[:-1]]).astype(np.uint64)
    out_cap = np.array([r.size for r in raws], dtype=np.uint64)
    out_off = np.concatenate([[0], np.cumsum(out_cap)[:-1]]).astype(np.uint64)
    C, D = int(in_len.sum()), int(out_cap.sum())
    log(f"batch workload: {len(mem)} members, {D} B -> {C} B in {time.time() - t0:.1f}s")
    d_in = torch.from_numpy(np.concatenate(ins)).to(dev)
    d_out = torch.empty(D + 1, dtype=torch.uint8, device=dev)
    ref = torch.from_numpy(np.concatenate(raws)).to(dev)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_members_baseline(ins, raws, host_cpu()[0])
        log(f"batch cpu baseline: {cpu}")
    del raws, ins, mem
    torch.cuda.synchronize()

    # autotuned tuning for this mix (device time, about 3 s budget)
    tuned, tuned_ms = gzippy_amd.autotune(d_in, in_off, in_len, d_out, out_off, out_cap, budget_ms=3000, ctx=ctx)
    log(f"batch autotune: chunk {tuned.chunk_bytes} ppc {tuned.pieces_per_chunk} waves/CU {tuned.waves_per_cu} "
        f"split {tuned.batch_split_bytes} -> {tuned_ms:.1f} ms")
    # batched: every member in one pipeline (one persistent inflate launch)
    d_out.zero_()
    status, olen, crcs, st = gzippy_amd.decompress_batch(d_in, in_off, in_len, d_out, out_off, out_cap, ctx=ctx)
    torch.cuda.synchronize()
    ok = bool((status == 0).all()) and bool((olen == out_cap).all()) and torch.equal(d_out[:D], ref)
    if not ok:
        log("FAILED: batched output differs")
        sys.exit(1)
    bt = []
    for _ in range(args.batch_steps):
        torch.cuda.synchronize()
        ts = time.perf_counter()
        gzippy_amd.decompress_batch(d_in, in_off, in_len, d_out, out_off, out_cap, ctx=ctx)
        torch.cuda.synchronize()
        bt.append(time.perf_counter() - ts)
    b_s = sorted(bt)[len(bt) // 2]

    # the hipGraph form on the same members: one replay (block finder, piece
    # plan, inflate launches, link checks, window chain, resolve, CRC32 and
    # trailer checks, all on the device)
    d_out.zero_()
    torch.cuda.synchronize()
    with gzippy_amd.BatchGraph(d_in, in_off, in_len, d_out, out_off, out_cap, ctx=ctx) as g:
        g_status, g_olen, _, g_st = g.launch()
        torch.cuda.synchronize()
        if not (bool((g_status == 0).all()) and torch.equal(d_out[:D], ref)):
            log("FAILED: graph output differs (batch members)")
            sys.exit(1)
        gt = []
        for _ in range(args.batch_steps):
            torch.cuda.synchronize()
            ts = time.perf_counter()
            _, _, _, g_st = g.launch()
            torch.cuda.synchronize()
            gt.append(time.perf_counter() - ts)
    g_s = sorted(gt)[len(gt) // 2]
    graph = {"graph_gbs": round(D / g_s / 1e9, 3), "graph_ms": round(g_s * 1e3, 3),
             "device_ms": round(g_st["ms_total"], 3), "pieces": g_st["pieces"], "redecoded": g_st["redecoded"],
             "candidates": g_st["candidates"]}
    log(f"batch graph: {graph}")

    # plain: one inflate launch per member (one wavefront, whole member),
    # back to back on one stream, no host synchronisation in between; on the
    # first `plain_members` members (a random sample of the same mix: at one
    # wavefront per member the whole batch would take minutes), against the
    # batched call on that same sample
    k = min(args.plain_members, len(in_off))
    s_in_len, s_out_cap = in_len[:k], out_cap[:k]
    s_in_off = in_off[:k]
    s_out_off = np.concatenate([[0], np.cumsum(s_out_cap)[:-1]]).astype(np.uint64)
    Ds = int(s_out_cap.sum())
    d_sout = torch.empty(Ds + 1, dtype=torch.uint8, device=dev)
    ref_s = ref[:Ds]  # the first k members are the first Ds reference bytes
    gzippy_amd.decompress_batch(d_in, s_in_off, s_in_len, d_sout, s_out_off, s_out_cap, ctx=ctx)  # warm buffers
    sb = []
    for _ in range(args.batch_steps):
        torch.cuda.synchronize()
        ts = time.perf_counter()
        gzippy_amd.decompress_batch(d_in, s_in_off, s_in_len, d_sout, s_out_off, s_out_cap, ctx=ctx)
        torch.cuda.synchronize()
        sb.append(time.perf_counter() - ts)
    sb_s = sorted(sb)[len(sb) // 2]
    d_sout.zero_()
    torch.cuda.synchronize()
    ts = time.perf_counter()
    status_p, _, _, st_p = gzippy_amd.decompress_members_plain(d_in, s_in_off, s_in_len, d_sout, s_out_off, s_out_cap,
                                                               ctx=ctx)
    torch.cuda.synchronize()
    p_s = time.perf_counter() - ts
    if not (bool((status_p == 0).all()) and torch.equal(d_sout[:Ds], ref_s)):
        log("FAILED: plain per-member output differs")
        sys.exit(1)
    # the same members, one full single-member call each (gzgpu_decompress_device:
    # find, plan, inflate split over the GPU, resolve, CRC; a host round trip
    # per member): the per-member form at its best
    d_sout.zero_()
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for i in range(k):
        o0, o1 = int(s_out_off[i]), int(s_out_off[i]) + int(s_out_cap[i])
        a0, a1 = int(s_in_off[i]), int(s_in_off[i]) + int(s_in_len[i])
        gzippy_amd.decompress_device(d_in[a0:a1], d_sout[o0:o1], ctx=ctx)
    torch.cuda.synchronize()
    c_s = time.perf_counter() - ts
    if not torch.equal(d_sout[:Ds], ref_s):
        log("FAILED: per-member calls output differs")
        sys.exit(1)
    del d_sout, ref_s
    return {"members": len(in_off), "compressed_bytes": C, "decompressed_bytes": D, "seed": args.seed + 3,
            "batched_gbs": round(D / b_s / 1e9, 3), "batched_ms": round(b_s * 1e3, 3),
            "plain": {"form": "one inflate launch per member (one wavefront, whole member), back to back on one "
                              "stream, no host sync between members; header parse, CRC32 and trailer checks on the "
                              "device", "members": k, "decompressed_bytes": Ds,
                      "plain_gbs": round(Ds / p_s / 1e9, 3), "plain_ms": round(p_s * 1e3, 3),
                      "plain_device_ms": round(st_p["ms_total"], 3),
                      "batched_gbs_same_members": round(Ds / sb_s / 1e9, 3), "batched_ms_same_members": round(sb_s * 1e3, 3),
                      "batched_over_plain": round(p_s / sb_s, 2),
                      "per_member_calls": {"form": "one gzgpu_decompress_device call per member (the whole "
                                                   "single-member pipeline, pieces split over the GPU, a host "
                                                   "round trip each)", "gbs": round(Ds / c_s / 1e9, 3),
                                           "ms": round(c_s * 1e3, 3), "batched_over_calls": round(c_s / sb_s, 2)}},
            "pieces": st["pieces"],
            "graph": graph,
            "autotuned": {"chunk_bytes": tuned.chunk_bytes, "pieces_per_chunk": tuned.pieces_per_chunk,
                          "waves_per_cu": tuned.waves_per_cu, "batch_split_bytes": tuned.batch_split_bytes,
                          "lut_bits": tuned.lut_bits},
            "batched_phase_ms": {k: round(st[k], 3) for k in ("ms_find", "ms_inflate", "ms_resolve", "ms_crc",
                                                              "ms_total", "ms_inflate_marker", "ms_inflate_clean")},
            "batched_launches": st["launches"],
            "cpu_baseline": cpu,
            "check": "every member's bytes equal the generator's; CRC32 and ISIZE verified per member"}


def mixed_leg(args, dev, ctx):
    """The mixed-entropy member: 2 GiB in 64 MiB segments cycling text /
    uniform random / byte runs / u32 deltas, one member at level 1 and one at
    level 9 (zlib, segments encoded in parallel pigz-style).  Device-resident,
    every step checked byte for byte on the device."""
    import torch

    import gzippy_amd
    from gzippy_amd import workloads as W

    n = args.mixed_mib << 20
    t0 = time.time()
    raw = W.generate_mixed(n, args.seed + 2, threads=16)
    d_raw = torch.from_numpy(raw).to(dev)
    out = {"decompressed_bytes": n, "segment_bytes": W.SEG, "seed": args.seed + 2,
           "encoder": "zlib-" + __import__("zlib").ZLIB_RUNTIME_VERSION + " pigz-style (8 MiB segments, 32 KiB "
                      "dictionary priming, sync flush), one member", "levels": []}
    log(f"mixed workload: {n} B generated in {time.time() - t0:.1f}s")
    d_out = torch.empty(n, dtype=torch.uint8, device=dev)
    for level in (1, 9):
        t0 = time.time()
        gz = W.gzip_compress_segments(raw, level)
        log(f"mixed L{level}: {n} B -> {gz.size} B in {time.time() - t0:.1f}s")
        d_in = torch.from_numpy(gz).to(dev)
        torch.cuda.synchronize()
        got, crc, st = gzippy_amd.decompress_device(d_in, d_out, ctx=ctx)
        torch.cuda.synchronize()
        if not (got == n and torch.equal(d_out, d_raw)):
            log(f"FAILED: mixed L{level} output differs")
            sys.exit(1)
        times, stats = [], []
        for _ in range(args.mixed_steps):
            torch.cuda.synchronize()
            ts = time.perf_counter()
            got, crc, st = gzippy_amd.decompress_device(d_in, d_out, ctx=ctx)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - ts)
            stats.append(st)
        if not torch.equal(d_out, d_raw):
            log(f"FAILED: mixed L{level} output differs after the timed steps")
            sys.exit(1)
        i = sorted(range(len(times)), key=lambda k: times[k])[len(times) // 2]
        t, st = times[i], stats[i]
        C = int(gz.size)
        k_ms = st["ms_inflate_kernel"] / max(1, st["inflate_launches"])
        rec = {"level": level, "compressed_bytes": C, "ms": round(t * 1e3, 3), "gbs": round(n / t / 1e9, 3),
               "phase_ms": {k: round(st[k], 3) for k in ("ms_find", "ms_inflate", "ms_resolve", "ms_crc", "ms_total")},
               "pieces": st["pieces"], "redecoded": st["redecoded"], "inflate_launches": st["inflate_launches"],
               "inflate_kernel_ms_per_launch": round(k_ms, 3),
               "roofline_c_plus_d": {"achieved_gbs": round((C + n) / (st["ms_inflate_kernel"] / 1e3) / 1e9, 2),
                                     "frac": round((C + n) / (st["ms_inflate_kernel"] / 1e3) / 1e9 / HBM_PEAK_GBS, 5)}}
        log(f"mixed L{level}: {rec}")
        out["levels"].append(rec)
        del d_in
    del d_raw, d_out
    torch.cuda.empty_cache()
    return out


def graph_leg(args, dev, ctx, members, max_kib):
    """A batch of small members (lognormal around 16 KiB, at most `max_kib`;
    the four kinds, levels 1/6/9): the hipGraph form (one replay: header
    parse, block finder and piece plan for the members worth splitting,
    inflate from a device-side queue, link checks, resolve, CRC32 and trailer
    checks, all on the device) against gzgpu_decompress_batch (host planning)
    on the same members."""
    import numpy as np
    import torch

    import gzippy_amd
    from gzippy_amd import workloads as W

    mem = W.batch_members(members, seed=args.seed + 5, median=16 << 10, max_size=max_kib << 10)
    raws = [m[0] for m in mem]
    ins = [m[1] for m in mem]
    in_len = np.array([x.size for x in ins], dtype=np.uint64)
    in_off = np.concatenate([[0], np.cumsum(in_len)[:-1]]).astype(np.uint64)
    out_cap = np.array([r.size for r in raws], dtype=np.uint64)
    out_off = np.concatenate([[0], np.cumsum(out_cap)[:-1]]).astype(np.uint64)
    C, D = int(in_len.sum()), int(out_cap.sum())
    d_in = torch.from_numpy(np.concatenate(ins)).to(dev)
    d_out = torch.empty(D + 1, dtype=torch.uint8, device=dev)
    ref = torch.from_numpy(np.concatenate(raws)).to(dev)
    del raws, ins, mem

    def timed(fn, steps):
        ts_all = []
        for _ in range(steps):
            torch.cuda.synchronize()
            ts = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts_all.append(time.perf_counter() - ts)
        return sorted(ts_all)[len(ts_all) // 2]

    d_out.zero_()
    with gzippy_amd.BatchGraph(d_in, in_off, in_len, d_out, out_off, out_cap, ctx=ctx) as g:
        status, olen, _, _ = g.launch()
        torch.cuda.synchronize()
        if not (bool((status == 0).all()) and torch.equal(d_out[:D], ref)):
            log("FAILED: graph output differs")
            sys.exit(1)
        g_s = timed(lambda: g.launch(), args.batch_steps)
    d_out.zero_()
    gzippy_amd.decompress_batch(d_in, in_off, in_len, d_out, out_off, out_cap, ctx=ctx)
    torch.cuda.synchronize()
    if not torch.equal(d_out[:D], ref):
        log("FAILED: batched output differs (small members)")
        sys.exit(1)
    b_s = timed(lambda: gzippy_amd.decompress_batch(d_in, in_off, in_len, d_out, out_off, out_cap, ctx=ctx),
                args.batch_steps)
    return {"members": len(in_off), "compressed_bytes": C, "decompressed_bytes": D, "seed": args.seed + 5,
            "graph_gbs": round(D / g_s / 1e9, 3), "graph_ms": round(g_s * 1e3, 3),
            "batched_gbs": round(D / b_s / 1e9, 3), "batched_ms": round(b_s * 1e3, 3),
            "check": "every member's bytes equal the generator's; CRC32 and ISIZE verified per member on the device"}


def serve_leg(args, rates):
    """Open-loop Poisson arrivals at each rate for args.serve_seconds; requests
    follow the batch mix, 10 % of them 4-64-member concatenations.  Latency is
    the server's own submit-to-result clock; every result is CRC-checked."""
    import threading
    import zlib

    import numpy as np

    import gzippy_amd
    from gzippy_amd import workloads as W

    rng = np.random.default_rng(args.seed + 4)
    base = W.batch_members(args.serve_pool, seed=args.seed + 4)
    pool = []
    for i in range(args.serve_pool):
        raw, g, _, level = base[i]
        if rng.random() < 0.10:
            # pigz-style: the request's bytes cut into k members, each its own gzip stream
            k = int(rng.integers(4, 65))
            cuts = np.linspace(0, raw.size, k + 1).astype(np.int64)
            gz = b"".join(W.gzip_compress(raw[cuts[j]:cuts[j + 1]], level).tobytes() for j in range(k))
            pool.append((gz, int(raw.size), zlib.crc32(raw), k))
        else:
            pool.append((g.tobytes(), int(raw.size), zlib.crc32(raw), 1))
    del base
    buf = np.empty(max(p[1] for p in pool) + 1, dtype=np.uint8)
    out = {"pool_requests": len(pool), "multi_member_share": round(sum(p[3] > 1 for p in pool) / len(pool), 3),
           "seconds_per_rate": args.serve_seconds, "rates": []}
    with gzippy_amd.Server(workers=2) as srv:
        for rate in rates:
            n = max(1, int(rate * args.serve_seconds))
            arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
            tickets = [None] * n
            ready = threading.Semaphore(0)
            lat, bad, bytes_out = [], [0], [0]
            t_start = time.perf_counter()

            def submitter():
                for i in range(n):
                    delay = t_start + arrivals[i] - time.perf_counter()
                    if delay > 0:
                        time.sleep(delay)
                    tickets[i] = srv.submit(pool[i % len(pool)][0])
                    ready.release()

            th = threading.Thread(target=submitter)
            th.start()
            for i in range(n):
                ready.acquire()
                info = srv.poll(tickets[i], -1)
                got = srv.fetch_into(tickets[i], buf)
                exp = pool[i % len(pool)]
                if info["status"] != 0 or got != exp[1] or zlib.crc32(memoryview(buf)[:got]) != exp[2]:
                    bad[0] += 1
                lat.append(info["latency_ms"])
                bytes_out[0] += got
            th.join()
            wall = time.perf_counter() - t_start
            lat.sort()
            rec = {"offered_req_per_s": rate, "requests": n, "wall_s": round(wall, 3),
                   "throughput_gbs": round(bytes_out[0] / wall / 1e9, 3),
                   "p50_ms": round(lat[len(lat) // 2], 3), "p99_ms": round(lat[min(len(lat) - 1, int(0.99 * len(lat)))], 3),
                   "max_ms": round(lat[-1], 3), "mismatches": bad[0]}
            log(f"serve {rate}/s: {rec}")
            if bad[0]:
                log("FAILED: served output differs")
                sys.exit(1)
            out["rates"].append(rec)
        out["server"] = srv.stats()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size-mib", type=int, default=1024, help="decompressed member size")
    ap.add_argument("--level", type=int, default=6)
    ap.add_argument("--chunk-kib", type=int, default=512)
    ap.add_argument("--pieces-per-chunk", type=int, default=0)
    ap.add_argument("--seed", type=int, default=0x675A0001)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--autotune", type=int, default=1, help="pick the tuning by measurement before timing")
    ap.add_argument("--batch-members", type=int, default=10000, help="batched-members leg (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--plain-members", type=int, default=500,
                    help="members of the batch timed in the plain one-launch-per-member form")
    ap.add_argument("--mixed-mib", type=int, default=2048, help="mixed-entropy member size (0: skip the leg)")
    ap.add_argument("--mixed-steps", type=int, default=3)
    ap.add_argument("--serve-seconds", type=float, default=3.0, help="serving leg, seconds per rate (0: skip)")
    ap.add_argument("--serve-pool", type=int, default=2000)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="after the timed steps, run one plain device-to-device copy of this many MiB "
                         "(the reference kernel for scaling FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for --gpus {args.gpus}); sharding members or pieces "
              "across GPUs is not implemented", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import gzippy_amd
    from gzippy_amd import workloads as W

    n = args.size_mib << 20
    t0 = time.time()
    raw = W.generate("text", n, args.seed)
    gz = W.gzip_compress(raw, args.level)
    log(f"workload: {n} B -> {gz.size} B ({W.encoder_name()} L{args.level}) in {time.time() - t0:.1f}s")

    dev = torch.device("cuda:0")
    d_in = torch.from_numpy(gz).to(dev)
    d_raw = torch.from_numpy(raw).to(dev)
    d_out = torch.empty(n, dtype=torch.uint8, device=dev)
    ctx = gzippy_amd.context(0)
    ctx.set_tuning(chunk_bytes=args.chunk_kib << 10, pieces_per_chunk=args.pieces_per_chunk)
    torch.cuda.synchronize()
    tuning = None
    if args.autotune:
        t, tms = gzippy_amd.autotune(d_in, [0], [gz.size], d_out, [0], [n], budget_ms=1500, ctx=ctx)
        tuning = {"chunk_bytes": t.chunk_bytes, "pieces_per_chunk": t.pieces_per_chunk, "waves_per_cu": t.waves_per_cu,
                  "lut_bits": t.lut_bits, "autotune_ms": round(tms, 3)}
        log(f"autotune: {tuning}")

    stats = None
    for i in range(args.warmup):
        d_out.zero_()
        got, crc, stats = gzippy_amd.decompress_device(d_in, d_out, ctx=ctx)
        torch.cuda.synchronize()
        if not (got == n and torch.equal(d_out, d_raw)):
            log("FAILED: output differs from the generator bytes")
            sys.exit(1)
        log(f"warmup {i}: ok, {stats['ms_total']:.2f} ms device, pieces {stats['pieces']}, "
            f"redecoded {stats['redecoded']}, inflate launches {stats['inflate_launches']}")
    del d_raw

    ms_infl, launches, ms_dev = 0.0, 0, 0.0
    ms_clean, ms_marker, clean_bytes, marker_bytes = 0.0, 0.0, 0, 0
    per_step = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        got, crc, st = gzippy_amd.decompress_device(d_in, d_out, ctx=ctx)
        per_step.append(time.perf_counter() - ts)
        ms_infl += st["ms_inflate_kernel"]
        ms_clean += st["ms_inflate_clean"]
        ms_marker += st["ms_inflate_marker"]
        clean_bytes += st["clean_in_bytes"] + st["clean_out_bytes"]
        marker_bytes += st["marker_in_bytes"] + st["marker_out_bytes"]
        launches += st["inflate_launches"]
        ms_dev += st["ms_total"]
        stats = st
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    ms_per_step = t / args.steps * 1e3
    gbs = n * args.steps / t / 1e9
    C, D = gz.size, n
    infl_ms_step = ms_infl / args.steps
    # dominant kernel: the step's inflate launch (the u16-ring kernel when the
    # member is split into speculative pieces); its algorithmic bytes are the
    # compressed bytes it consumes plus the bytes it writes (u16 entries
    # count 2)
    u16 = ms_marker >= ms_clean
    kname = "k_inflate<unsigned short>" if u16 else "k_inflate<unsigned char>"
    k_launches = launches / args.steps
    k_ms_launch = (ms_marker if u16 else ms_clean) / args.steps / max(1.0, k_launches)
    # algorithmic bytes of the launch: the member's compressed bytes read once
    # plus its decompressed bytes written once (C + D); the u16 marker
    # entries it writes on the way are overhead, reported beside it
    k_alg = (C + D) / max(1.0, k_launches)
    k_written = (marker_bytes if u16 else clean_bytes) // args.steps / max(1.0, k_launches)
    achieved = k_alg / (k_ms_launch / 1e3) / 1e9
    stage_gbs = (C + D) / ((ms_clean + ms_marker) / args.steps / 1e3) / 1e9
    lat = sorted(per_step)
    p50 = lat[len(lat) // 2] * 1e3
    p99 = lat[min(len(lat) - 1, int(round(0.99 * (len(lat) - 1))))] * 1e3

    line = {
        "metric": METRIC, "value": round(gbs, 3), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {
            "workload": f"single gzip member: {args.size_mib} MiB text generator (order-1 Markov, Zipf 1.2, "
                        f"35% copies), {W.encoder_name()} level {args.level}, speculative parallel decode",
            "compressed_bytes": int(C), "decompressed_bytes": int(D), "seed": args.seed,
            "pieces": stats["pieces"], "redecoded": stats["redecoded"],
            "inflate_launches_per_step": launches / args.steps,
            "phase_ms": {k: round(stats[k], 3) for k in ("ms_find", "ms_inflate", "ms_resolve", "ms_crc", "ms_total")},
            "latency_ms": {"p50": round(p50, 3), "p99": round(p99, 3)},
            "tuning": tuning,
        },
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "algorithmic_bytes_per_launch": int(k_alg), "algorithmic": "C + D (compressed read once, "
                     "decompressed written once)", "consumed_plus_written_bytes_per_launch": int(k_written),
                     "kernel_ms_per_launch": round(k_ms_launch, 3),
                     "launches_per_step": k_launches,
                     "inflate_stage": {"ms_per_step": round(infl_ms_step, 3),
                                       "achieved_c_plus_d_gbs": round(stage_gbs, 2)}},
    }
    line["host_buffers"] = host_leg(gz, n, ctx)
    log(f"host buffers: {line['host_buffers']}")
    del d_in, d_out
    torch.cuda.empty_cache()
    if args.mixed_mib:
        line["mixed"] = mixed_leg(args, dev, ctx)
    if args.batch_members:
        torch.cuda.empty_cache()
        line["batch"] = batch_leg(args, dev, ctx)
        torch.cuda.empty_cache()
        line["batch_graph"] = [graph_leg(args, dev, ctx, 512, 64), graph_leg(args, dev, ctx, 10000, 256)]
    if args.serve_seconds > 0:
        torch.cuda.empty_cache()
        line["serving"] = serve_leg(args, (200, 1000, 5000))
    traffic = measured_traffic(args, kname)
    if traffic is not None:
        line["roofline"]["traffic"] = traffic["bytes_per_launch"]
        line["roofline"]["traffic_source"] = traffic["source"]
    if args.calibrate_copy_mib:
        nb = args.calibrate_copy_mib << 20
        a = torch.empty(nb, dtype=torch.uint8, device=dev).fill_(1)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"calibration copy: {nb} B")
    if not args.no_cpu:
        log("cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(gz, raw)
        log(f"cpu baseline: {line['cpu_baseline']}")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
