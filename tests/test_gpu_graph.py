"""The hipGraph batch form with members split on the device: the block
finder, the piece plan, the link check and the re-decode of a miss, the
window chain and the marker resolution all run inside the graph, so one
replay decodes members of any size with no host decision in between.

Every replay is checked byte for byte against the generator's bytes and
against gzgpu_decompress_batch (status, length, CRC32), with eager torch work
and synchronisations between the replays."""
import struct
import zlib

import numpy as np
import pytest

from gzippy_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _layout(ins, raws, caps=None):
    in_len = np.array([x.size for x in ins], dtype=np.uint64)
    in_off = np.concatenate([[0], np.cumsum(in_len)[:-1]]).astype(np.uint64)
    out_cap = np.array([max(r.size, 1) for r in raws] if caps is None else caps, dtype=np.uint64)
    out_off = np.concatenate([[0], np.cumsum(out_cap)[:-1]]).astype(np.uint64)
    return in_off, in_len, out_off, out_cap


def _large_members(scale=1):
    spec = [("text", 24, 6), ("u32", 8, 1), ("runs", 32, 9), ("random", 8, 6), ("text", 6, 1), ("u32", 5, 9)]
    out = []
    for i, (kind, mib, level) in enumerate(spec):
        raw = W.generate(kind, (mib << 20) // scale + 777 * i, 4100 + i)
        out.append((raw, W.gzip_compress(raw, level)))
    return out


def _check_replays(gpu, g, d_out, raws, out_off, reps, want_status=None):
    import torch
    stats = []
    for rep in range(reps):
        d_out.zero_()
        x = torch.arange(1 << 20, device="cuda", dtype=torch.int32).sum()  # eager work between replays
        torch.cuda.synchronize()
        assert int(x) > 0
        status, olen, crc, st = g.launch(raise_on_error=False)
        say("replay", rep, "ms", round(st["ms_total"], 3), "pieces", st["pieces"], "redecoded", st["redecoded"],
            "bad", np.nonzero(status)[0][:8])
        host = d_out.cpu().numpy()
        for i, r in enumerate(raws):
            if want_status is not None and want_status[i] != 0:
                assert status[i] == want_status[i], (rep, i, status[i])
                continue
            assert status[i] == 0, (rep, i, status[i])
            assert olen[i] == r.size, (rep, i)
            assert np.array_equal(host[out_off[i]: out_off[i] + r.size], r), (rep, i)
        stats.append((status.copy(), olen.copy(), crc.copy(), st))
    return stats


def test_graph_splits_large_members(gpu):
    """Small members and members of 5-32 MiB (text, u32 deltas, byte runs,
    stored random data) in one graph: the large ones are split on the device
    (pieces > members) and every replay is exact and equal to the batch path."""
    import torch
    small = W.batch_members(120, seed=808, max_size=1 << 20)
    mem = [(m[0], m[1]) for m in small] + _large_members()
    raws = [m[0] for m in mem]
    ins = [m[1] for m in mem]
    in_off, in_len, out_off, out_cap = _layout(ins, raws)
    d_in = torch.from_numpy(np.concatenate(ins)).cuda()
    d_out = torch.zeros(int(out_cap.sum()), dtype=torch.uint8, device="cuda")
    ref_out = torch.zeros_like(d_out)
    ref = gpu.decompress_batch(d_in, in_off, in_len, ref_out, out_off, out_cap)
    with gpu.BatchGraph(d_in, in_off, in_len, d_out, out_off, out_cap) as g:
        stats = _check_replays(gpu, g, d_out, raws, out_off, 3)
    for status, olen, crc, st in stats:
        assert np.array_equal(status, ref[0]) and np.array_equal(olen, ref[1]) and np.array_equal(crc, ref[2])
        assert st["pieces"] > len(raws) + 100, st["pieces"]  # the large members were split
        assert st["candidates"] > 0


def test_graph_forced_misses_redecode(gpu):
    """Every second planned start one bit late (the miss_every hook, taken at
    creation): the link checks on the device re-decode each missed piece in
    place from the end of the piece before it (and, for what is still
    unlinked, the member's remainder from the last confirmed boundary); the
    bytes stay exact."""
    import torch
    ctx = gpu.context()
    old = ctx.tuning()
    mem = _large_members(scale=2)
    raws = [m[0] for m in mem]
    ins = [m[1] for m in mem]
    in_off, in_len, out_off, out_cap = _layout(ins, raws)
    d_in = torch.from_numpy(np.concatenate(ins)).cuda()
    d_out = torch.zeros(int(out_cap.sum()), dtype=torch.uint8, device="cuda")
    try:
        ctx.set_tuning(miss_every=2)
        with gpu.BatchGraph(d_in, in_off, in_len, d_out, out_off, out_cap) as g:
            stats = _check_replays(gpu, g, d_out, raws, out_off, 2)
    finally:
        ctx.set_tuning(old.chunk_bytes, old.pieces_per_chunk, old.waves_per_cu, old.batch_split_bytes, 0)
    for _, _, _, st in stats:
        assert st["redecoded"] >= 4, st  # every split member with two or more starts missed


def test_graph_split_member_errors(gpu):
    """Split members that must fail do so alone: output capacity half the
    member's size (OUTPUT_TOO_SMALL), a trailer CRC changed (CRC_MISMATCH),
    a payload byte flipped (an error, never silent bytes), a bad header; the
    other members stay exact, equal to the batch path."""
    import torch
    mem = _large_members(scale=2)
    raws = [m[0] for m in mem]
    ins = [m[1].copy() for m in mem]
    caps = [r.size for r in raws]
    caps[0] = raws[0].size // 2
    ins[2][-8] ^= 0x5A  # trailer CRC
    ins[3][ins[3].size // 2] ^= 0x10  # stored payload byte
    ins[4][0] = 0x1e  # magic
    in_off, in_len, out_off, out_cap = _layout(ins, raws, caps)
    d_in = torch.from_numpy(np.concatenate(ins)).cuda()
    d_out = torch.zeros(int(out_cap.sum()), dtype=torch.uint8, device="cuda")
    ref_out = torch.zeros_like(d_out)
    ref = gpu.decompress_batch(d_in, in_off, in_len, ref_out, out_off, out_cap, raise_on_error=False)
    say("batch statuses", ref[0])
    assert ref[0][0] == 9 and ref[0][2] == 5 and ref[0][3] != 0 and ref[0][4] == 1
    want = ref[0].copy()
    with gpu.BatchGraph(d_in, in_off, in_len, d_out, out_off, out_cap) as g:
        stats = _check_replays(gpu, g, d_out, raws, out_off, 2, want_status=want)
    for status, olen, crc, _ in stats:
        assert np.array_equal(status, ref[0]), (status, ref[0])
        good = status == 0
        assert np.array_equal(olen[good], ref[1][good]) and np.array_equal(crc[good], ref[2][good])


def _graph_check(gpu, ins, raws, label, caps=None, reps=2, expect_ok=None):
    """Decode `ins` as one graph (twice, with eager work between); members
    whose expect_ok is False must fail alone, every other one must be exact."""
    import torch
    in_off, in_len, out_off, out_cap = _layout(ins, raws, caps)
    d_in = torch.from_numpy(np.concatenate(ins)).cuda()
    d_out = torch.zeros(int(out_cap.sum()), dtype=torch.uint8, device="cuda")
    stats = []
    with gpu.BatchGraph(d_in, in_off, in_len, d_out, out_off, out_cap) as g:
        for rep in range(reps):
            d_out.zero_()
            torch.cuda.synchronize()
            status, olen, crc, st = g.launch(raise_on_error=False)
            say(label, "replay", rep, "ms", round(st["ms_total"], 3), "pieces", st["pieces"], "redecoded",
                st["redecoded"], "bad", np.nonzero(status)[0][:8])
            host = d_out.cpu().numpy()
            for i, r in enumerate(raws):
                if expect_ok is not None and not expect_ok[i]:
                    continue
                assert status[i] == 0, (label, rep, i, status[i])
                assert olen[i] == r.size and np.array_equal(host[out_off[i]: out_off[i] + r.size], r), (label, rep, i)
            stats.append((status.copy(), st))
    return stats


def test_graph_hazard_contents(gpu):
    """The speculation hazards of the host path's tests, in one graph:
    fixed-Huffman-only members (no dynamic header for the finder), gzip of
    deflate data (real dynamic headers inside stored payloads), the phantom
    end-of-stream vectors inside stored payloads, copies at distance 32767
    across piece seams, empty members and every header field."""
    import test_gpu_hazards as H
    ins, raws = [], []

    def add(raw, gz):
        raws.append(np.frombuffer(bytes(raw), np.uint8).copy() if not isinstance(raw, np.ndarray) else raw)
        ins.append(np.frombuffer(bytes(gz), np.uint8).copy() if not isinstance(gz, np.ndarray) else gz)

    for kind in ("text", "u32", "runs"):
        raw = W.generate(kind, 6 << 20, seed=55)
        add(raw, H.gz_fixed(raw))
    inner = W.gzip_compress(W.generate("u32", 24 << 20, seed=8), 1)
    for level in (1, 6):
        add(inner, W.gzip_compress(inner, level))
    fixed_empty_bfinal, footer = bytes([0x03, 0x00]), bytes([0xAA, 0xBB, 0xCC, 0xDD, 0x01, 0, 0, 0])
    payload = (fixed_empty_bfinal + footer + bytes(100)) * 20000 + bytes(W.generate("random", 1 << 20, 3))
    for level in (0, 1, 6):
        co = zlib.compressobj(level, zlib.DEFLATED, 31)
        add(payload, co.compress(payload) + co.flush())
    per = H.periodic(8 << 20, 32767)
    add(per, W.gzip_compress(per, 6, encoder="libdeflate"))
    add(np.zeros(0, np.uint8), W.gzip_compress(np.zeros(0, np.uint8), 6))
    body = W.generate("text", 3 << 20, 77)
    for kw in (dict(fname=b"data.txt"), dict(fcomment=b"a comment " * 10), dict(fhcrc=True),
               dict(fextra=b"XY" + struct.pack("<H", 65531) + bytes(65531), fname=b"x" * 5000)):
        add(body, H.gz_with_fields(body, **kw))
    stats = _graph_check(gpu, ins, raws, "hazards")
    assert stats[0][1]["pieces"] > len(raws) + 50


def test_graph_corruption_sweep(gpu):
    """Seeded truncations and bit flips of a split 1 MiB member, each variant
    a member of the same graph beside intact copies: a corrupted member
    either fails alone or decodes to the original bytes, and the intact ones
    stay exact (the reference's fuzz contract, per member)."""
    raw = W.generate("text", 1 << 20, 99)
    gz = W.gzip_compress(raw, 6)
    rng = np.random.default_rng(4321)
    ins, raws, bad = [], [], []
    for case in range(60):
        v = gz.copy()
        if case % 3 == 0:
            v = v[: int(rng.integers(1, gz.size))]
        else:
            for p, b in zip(rng.integers(0, gz.size, 1 + case % 4), rng.integers(0, 8, 1 + case % 4)):
                v[p] ^= np.uint8(1 << int(b))
        ins.append(v)
        raws.append(raw)
        bad.append(True)
        if case % 4 == 0:  # an intact copy between the corrupted ones
            ins.append(gz.copy())
            raws.append(raw)
            bad.append(False)
    import torch
    in_off, in_len, out_off, out_cap = _layout(ins, raws)
    d_in = torch.from_numpy(np.concatenate(ins)).cuda()
    d_out = torch.zeros(int(out_cap.sum()), dtype=torch.uint8, device="cuda")
    with gpu.BatchGraph(d_in, in_off, in_len, d_out, out_off, out_cap) as g:
        status, olen, crc, st = g.launch(raise_on_error=False)
    say("graph corruption", "pieces", st["pieces"], "redecoded", st["redecoded"], "errors", int((status != 0).sum()))
    host = d_out.cpu().numpy()
    errors = 0
    for i, r in enumerate(raws):
        exact = status[i] == 0 and olen[i] == r.size and np.array_equal(host[out_off[i]: out_off[i] + r.size], r)
        if not bad[i]:
            assert exact, i
        elif status[i] != 0:
            errors += 1
        else:
            assert exact, (i, "a corrupted member returned wrong bytes without an error")
    assert st["pieces"] > len(raws) * 4  # the members were split
    assert errors > 30
