"""End-to-end GPU decode through the C ABI: byte-exact output, CRC32 and
ISIZE, error behaviour, multi-member, batch == per-member."""
import io
import zlib

import numpy as np
import pytest

from gzippy_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def dev(arr):
    import torch
    return torch.from_numpy(np.ascontiguousarray(arr)).cuda()


@pytest.fixture
def fine_grid(gpu):
    """Small chunks so even MiB-sized members split into many speculative pieces."""
    ctx = gpu.context()
    old = ctx.tuning()
    ctx.set_tuning(chunk_bytes=64 << 10, pieces_per_chunk=4)
    yield ctx
    ctx.set_tuning(old.chunk_bytes, old.pieces_per_chunk, old.waves_per_cu, old.batch_split_bytes)


def test_text_member_16mib(gpu):
    raw, gz = W.text_member(16 << 20, seed=0x675A0000)
    say("text member", raw.size, gz.size)
    st = {}
    out = gpu.decompress(gz, stats=st)
    assert out == raw.tobytes()
    say("stats", st)


@pytest.mark.parametrize("kind,level", [("text", 6), ("text", 1), ("u32", 1), ("runs", 9)])
def test_midblock_starts(gpu, kind, level):
    """Pieces that start inside a block (at the sync point of a grid bit) and
    predecessors that stop there exactly: the output stays byte-exact and
    nearly every such start links on the first launch."""
    ctx = gpu.context()
    old = ctx.tuning()
    try:
        # 2 KiB compressed per piece: most grid points fall inside blocks
        ctx.set_tuning(chunk_bytes=64 << 10, pieces_per_chunk=32)
        raw = W.generate(kind, 6 << 20, seed=11 * level + len(kind))
        gz = W.gzip_compress(raw, level)
        st = {}
        out = gpu.decompress(gz, stats=st)
        say("midblock", kind, level, gz.size, {k: st[k] for k in ("pieces", "midblock_pieces", "midblock_linked",
                                                                  "redecoded")})
        assert out == raw.tobytes()
        assert st["midblock_pieces"] > 0
        assert st["midblock_linked"] >= 0.9 * st["midblock_pieces"]
    finally:
        ctx.set_tuning(old.chunk_bytes, old.pieces_per_chunk, old.waves_per_cu, old.batch_split_bytes)


@pytest.mark.parametrize("enc", ["libdeflate", "zlib"])
def test_final_block_split(gpu, enc):
    """The member's final block (BFINAL = 1, never a hit of the reference's
    finder predicate) is found by the planner's scan of the member's tail and
    takes starts inside it like any other block: the output stays exact and
    those starts link."""
    ctx = gpu.context()
    old = ctx.tuning()
    try:
        ctx.set_tuning(chunk_bytes=64 << 10, pieces_per_chunk=32)
        raw = W.generate("text", 6 << 20, seed=91)
        gz = W.gzip_compress(raw, 6, encoder=enc)
        st = {}
        out = gpu.decompress(gz, stats=st)
        say("final block", enc, gz.size, {k: st[k] for k in ("pieces", "midblock_pieces", "midblock_linked",
                                                            "final_block_pieces", "redecoded")})
        assert out == raw.tobytes()
        assert st["final_block_pieces"] > 0
        assert st["midblock_linked"] >= 0.9 * st["midblock_pieces"]
    finally:
        ctx.set_tuning(old.chunk_bytes, old.pieces_per_chunk, old.waves_per_cu, old.batch_split_bytes)


@pytest.mark.parametrize("kind,level", [("text", 6), ("u32", 1), ("runs", 9)])
def test_lut_root_width(gpu, kind, level):
    """The 10-bit litlen LUT root (tuning lut_bits, an autotuner candidate):
    a split member (marker pieces, u16-ring kernel) and a batch of whole
    members (u8-ring kernel) decode to the same bytes as with the default
    9-bit root; other widths are refused."""
    import torch
    ctx = gpu.context()
    old = ctx.tuning()
    try:
        raw = W.generate(kind, 6 << 20, seed=17 * level + len(kind))
        gz = W.gzip_compress(raw, level)
        small = [W.generate(kind, (64 << 10) + 77 * i, seed=i) for i in range(8)]
        sgz = [W.gzip_compress(x, level) for x in small]
        for bits in (10, 9, 0):
            ctx.set_tuning(lut_bits=bits)
            st = {}
            assert gpu.decompress(gz, stats=st) == raw.tobytes(), bits
            say("lut root", kind, level, bits, st["pieces"])
            blob = np.concatenate(sgz)
            off = np.cumsum([0] + [g.size for g in sgz[:-1]])
            d_in = torch.from_numpy(blob).cuda()
            cap = [x.size + 64 for x in small]
            ooff = np.cumsum([0] + cap[:-1])
            d_out = torch.zeros(int(sum(cap)), dtype=torch.uint8, device="cuda")
            status, olen, crc, _ = gpu.decompress_batch(d_in, off, [g.size for g in sgz], d_out, ooff, cap)
            out = d_out.cpu().numpy()
            for i, x in enumerate(small):
                assert status[i] == 0 and olen[i] == x.size
                assert np.array_equal(out[ooff[i]: ooff[i] + x.size], x), (bits, i)
        with pytest.raises(ValueError):
            ctx.set_tuning(lut_bits=11)
    finally:
        ctx.set_tuning(old.chunk_bytes, old.pieces_per_chunk, old.waves_per_cu, old.batch_split_bytes)


@pytest.mark.parametrize("kind", ["text", "random", "runs", "u32"])
@pytest.mark.parametrize("level", [1, 6, 9])
def test_kinds_levels_split(gpu, fine_grid, kind, level):
    say("split", kind, level)
    raw = W.generate(kind, (3 << 20) + 1234, seed=7 * level + len(kind))
    for enc in ("libdeflate", "zlib"):
        gz = W.gzip_compress(raw, level, encoder=enc)
        st = {}
        out = gpu.decompress(gz, stats=st)
        assert out == raw.tobytes(), (kind, level, enc, st)
        if kind == "text":
            assert st["pieces"] > 1


def test_mixed_segments(gpu, fine_grid):
    raw = W.generate_mixed(12 << 20, seed=3, seg=1 << 20)
    for level in (1, 9):
        gz = W.gzip_compress(raw, level)
        say("mixed", level, gz.size)
        assert gpu.decompress(gz) == raw.tobytes()


def test_stored_runs_skipped_by_the_finder(gpu):
    """A member of 4 MiB segments cycling text / random / runs / u32: the
    random segments are runs of stored blocks, whose payloads the block
    finder skips (stat stored_runs_skipped); the output stays exact, also
    with the payload of one stored block corrupted (a CRC error, not silent
    bytes)."""
    import torch
    raw = W.generate_mixed(24 << 20, seed=31, seg=4 << 20)
    for level in (1, 6):
        gz = W.gzip_compress(raw, level)
        out = torch.zeros(raw.size, dtype=torch.uint8, device="cuda")
        n, crc, st = gpu.decompress_device(dev(gz), out)
        say("stored runs", level, gz.size, {k: st[k] for k in ("pieces", "redecoded", "stored_runs_skipped",
                                                              "candidates", "ms_find")})
        assert n == raw.size and np.array_equal(out.cpu().numpy(), raw)
        assert st["stored_runs_skipped"] > 20
    # a flipped byte inside a skipped payload: the stored copy carries it,
    # the trailer CRC catches it
    gz = W.gzip_compress(raw, 1)
    bad = gz.copy()
    bad[gz.size * 5 // 24] ^= 0x40  # inside the second segment (random: stored blocks)
    with pytest.raises(gpu.GzippyError):
        out = torch.zeros(raw.size, dtype=torch.uint8, device="cuda")
        gpu.decompress_device(dev(bad), out)


def test_device_api(gpu):
    import torch
    raw, gz = W.text_member(8 << 20, seed=9)
    out = torch.empty(raw.size, dtype=torch.uint8, device="cuda")
    n, crc, st = gpu.decompress_device(dev(gz), out)
    assert n == raw.size and crc == zlib.crc32(raw.tobytes())
    assert np.array_equal(out.cpu().numpy(), raw)


@pytest.mark.parametrize("kind,level,kib", [("text", 6, 24576), ("u32", 1, 24576), ("runs", 9, 24576),
                                            ("text", 6, 1024), ("u32", 1, 3072), ("text", 6, 160), ("u32", 1, 96)])
def test_window_chain_prefix_equals_walk(gpu, kind, level, kib):
    """Pieces too short to flip form one long window chain: resolved as a
    parallel prefix (every chain) and piece by piece (no chain), the output
    is the same, byte-exact.  The small members give chains of a few pieces,
    whose prefix is a single step (slab reads and output writes in one
    launch)."""
    import os
    raw = W.generate(kind, kib << 10, seed=17 + level)
    gz = W.gzip_compress(raw, level)
    old = os.environ.get("GZGPU_CHAIN_SCAN_MIN")
    try:
        for v in ("0", "1000000000"):
            os.environ["GZGPU_CHAIN_SCAN_MIN"] = v
            st = {}
            out = gpu.decompress(gz, stats=st)
            say("chain", kind, level, v, {k: st[k] for k in ("pieces", "launches", "ms_resolve")})
            assert out == raw.tobytes(), v
    finally:
        if old is None:
            os.environ.pop("GZGPU_CHAIN_SCAN_MIN", None)
        else:
            os.environ["GZGPU_CHAIN_SCAN_MIN"] = old


def test_issue_priority_balance_off_same_output(gpu):
    """GZGPU_PROG=0 turns off the inflate launch's issue-priority balance
    (the launch-wide progress counter of a single member, the remaining-work
    weights of a batch whose pieces outnumber the resident wavefronts): the
    bytes are the same.  The switch is read once per process, so a child
    process decodes with it off."""
    import os
    import subprocess
    import sys
    code = r"""
import numpy as np, torch
import gzippy_amd
from gzippy_amd import workloads as W
raw = W.generate("text", 8 << 20, 23)
gz = W.gzip_compress(raw, 6)
assert gzippy_amd.decompress(gz) == raw.tobytes()
mem = W.batch_members(5000, seed=29, median=8 << 10, max_size=64 << 10)
raws = [m[0] for m in mem]
ins = [m[1] for m in mem]
in_len = np.array([x.size for x in ins], dtype=np.uint64)
in_off = np.concatenate([[0], np.cumsum(in_len)[:-1]]).astype(np.uint64)
out_cap = np.array([r.size for r in raws], dtype=np.uint64)
out_off = np.concatenate([[0], np.cumsum(out_cap)[:-1]]).astype(np.uint64)
d_in = torch.from_numpy(np.concatenate(ins)).cuda()
D = int(out_cap.sum())
d_out = torch.empty(D + 1, dtype=torch.uint8, device="cuda")
status, olen, crc, st = gzippy_amd.decompress_batch(d_in, in_off, in_len, d_out, out_off, out_cap)
assert (status == 0).all()
assert torch.equal(d_out[:D].cpu(), torch.from_numpy(np.concatenate(raws)))
print("ok", st["pieces"])
"""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], env=dict(os.environ, GZGPU_PROG="0"), cwd=root,
                       capture_output=True, text=True, timeout=300)
    say("balance off:", r.stdout.strip()[-200:], r.stderr.strip()[-300:] if r.returncode else "")
    assert r.returncode == 0 and "ok" in r.stdout


def test_writer_api(gpu):
    raw, gz = W.text_member(2 << 20, seed=10)
    sink = io.BytesIO()
    n = gpu.decompress_to_writer(gz, sink)
    assert n == raw.size and sink.getvalue() == raw.tobytes()


def test_writer_gets_output_in_order_chunks(gpu):
    """A large member reaches the writer as consecutive chunks, in order."""
    raw, gz = W.text_member(21 << 20, seed=13)

    class Rec:
        def __init__(self):
            self.parts = []

        def write(self, b):
            self.parts.append(bytes(b))
            return len(b)

    rec = Rec()
    n = gpu.decompress_to_writer(gz, rec)
    say("writer calls", [len(p) for p in rec.parts])
    assert n == raw.size and len(rec.parts) >= 3
    assert b"".join(rec.parts) == raw.tobytes()


def test_small_and_edge_members(gpu):
    for raw in [b"", b"a", b"ab" * 3, bytes(range(256)) * 300]:
        for level in (0, 1, 6, 9):
            c = zlib.compressobj(level, zlib.DEFLATED, 31)
            gz = c.compress(raw) + c.flush()
            assert gpu.decompress(gz) == raw, (len(raw), level)


def test_non_gzip_returns_empty(gpu):
    assert gpu.decompress(b"") == b""
    assert gpu.decompress(b"plain text, not gzip") == b""


def test_multi_member_and_trailing_garbage(gpu):
    a, b = W.generate("text", 300_000, 1), W.generate("u32", 500_001, 2)
    gz = np.concatenate([W.gzip_compress(a, 6), W.gzip_compress(b, 1)]).tobytes()
    assert gpu.decompress(gz) == a.tobytes() + b.tobytes()
    assert gpu.decompress(gz + b"garbage!") == a.tobytes() + b.tobytes()


@pytest.mark.parametrize("form,block", [("gz", 1 << 20), ("bgzf", 65280)])
def test_sized_member_streams(gpu, form, block):
    """Members that carry their size (gzippy's "GZ" subfield, BGZF's "BC")
    decode together in one pipeline; a bad member is still reported."""
    raw = W.generate_mixed(6 << 20, 21, seg=768 << 10)
    gz = W.sized_member_stream(raw, block, 6, form)
    st = {}
    out = gpu.decompress(gz, stats=st)
    say("sized", form, len(gz), {k: st[k] for k in ("pieces", "inflate_launches", "launches")})
    assert out == raw.tobytes()
    assert st["inflate_launches"] <= 4  # one pipeline, not one per member
    sink = io.BytesIO()
    assert gpu.decompress_to_writer(gz, sink) == raw.size and sink.getvalue() == raw.tobytes()
    # a CRC byte of a middle member: the member-by-member walk reports it
    bad = bytearray(gz)
    members = len(range(0, raw.size, block))
    mid = W.sized_member_stream(raw[: (members // 2) * block], block, 6, form)
    mid_len = len(mid) - (len(W.BGZF_EOF) if form == "bgzf" else 0)
    bad[mid_len - 8] ^= 0x10  # CRC32 of the member ending at mid_len
    with pytest.raises(gpu.GzippyError) as e:
        gpu.decompress(bytes(bad))
    assert e.value.kind == "CrcMismatch"


def test_corruption_is_an_error(gpu):
    raw, gz = W.text_member(1 << 20, seed=11)
    bad = gz.copy()
    bad[-6] ^= 0x10
    with pytest.raises(gpu.GzippyError) as e:
        gpu.decompress(bad)
    assert e.value.kind == "CrcMismatch"
    bad = gz.copy()
    bad[-3] ^= 0x01
    with pytest.raises(gpu.GzippyError) as e:
        gpu.decompress(bad)
    assert e.value.kind == "SizeMismatch"
    with pytest.raises(gpu.GzippyError):
        gpu.decompress(gz[: gz.size // 2])
    bad = gz.copy()
    bad[gz.size // 2: gz.size // 2 + 64] ^= 0x5A
    with pytest.raises(gpu.GzippyError):
        gpu.decompress(bad)


def test_batch_matches_per_member(gpu):
    import torch
    mem = W.batch_members(300, seed=12, max_size=4 << 20)
    ins = [m[1] for m in mem]
    raws = [m[0] for m in mem]
    in_off = np.cumsum([0] + [x.size for x in ins[:-1]])
    in_len = np.array([x.size for x in ins])
    out_cap = np.array([r.size for r in raws])
    out_off = np.cumsum([0] + list(out_cap[:-1]))
    d_in = dev(np.concatenate(ins))
    d_out = torch.zeros(int(out_cap.sum()) + 1, dtype=torch.uint8, device="cuda")
    say("batch", len(mem), int(out_cap.sum()))
    status, olen, crc, st = gpu.decompress_batch(d_in, in_off, in_len, d_out, out_off, out_cap)
    assert (status == 0).all()
    host = d_out.cpu().numpy()
    for i, r in enumerate(raws):
        assert olen[i] == r.size
        assert np.array_equal(host[out_off[i]: out_off[i] + r.size], r), i
        assert crc[i] == zlib.crc32(r.tobytes())
    # each member alone through the plain single-member path
    for i in range(0, 300, 37):
        out = torch.zeros(max(raws[i].size, 1), dtype=torch.uint8, device="cuda")
        n, c, _ = gpu.decompress_device(dev(ins[i]), out)
        assert n == raws[i].size and np.array_equal(out[:n].cpu().numpy(), raws[i])


def test_batch_split_members_unaligned_outputs(gpu):
    """Members large enough to be split into speculative pieces, written at
    odd output offsets: the marker resolution writes 16-byte blocks aligned
    in the address space, whatever the member's output base."""
    import torch
    ctx = gpu.context()
    old = ctx.tuning()
    ctx.set_tuning(chunk_bytes=64 << 10, pieces_per_chunk=4, batch_split_bytes=32 << 10)
    try:
        kinds = ["text", "runs", "u32", "text", "random", "text"]
        raws, ins = [], []
        for i, k in enumerate(kinds):
            raw = W.generate(k, (1 << 20) + 4099 * i + 13, 700 + i)
            raws.append(raw)
            ins.append(W.gzip_compress(raw, 6 if i % 2 else 1))
        in_off = np.cumsum([0] + [x.size for x in ins[:-1]])
        in_len = np.array([x.size for x in ins])
        out_cap = np.array([r.size for r in raws])
        out_off = np.cumsum([3] + [c + 7 for c in out_cap[:-1]])
        d_in = dev(np.concatenate(ins))
        d_out = torch.zeros(int(out_off[-1] + out_cap[-1]) + 5, dtype=torch.uint8, device="cuda")
        say("split batch", len(raws), [x.size for x in ins])
        st = {}
        status, olen, crc, st = gpu.decompress_batch(d_in, in_off, in_len, d_out, out_off, out_cap)
        assert (status == 0).all(), status
        say("pieces", st["pieces"])
        assert st["pieces"] > len(raws)
        host = d_out.cpu().numpy()
        for i, r in enumerate(raws):
            assert olen[i] == r.size
            assert np.array_equal(host[out_off[i]: out_off[i] + r.size], r), i
            assert crc[i] == zlib.crc32(r.tobytes())
            # the gaps between members stay untouched
            assert not host[out_off[i] + r.size: out_off[i] + r.size + 7].any()
    finally:
        ctx.set_tuning(old.chunk_bytes, old.pieces_per_chunk, old.waves_per_cu, old.batch_split_bytes)


def test_unaligned_input_pointers(gpu, fine_grid):
    """Members whose compressed bytes start at every offset mod 16: decoded
    (single piece and split), and block-finder hits shifted exactly."""
    import torch
    cases = [W.text_member(300_000, seed=41), W.text_member(3 << 20, seed=42)]
    for raw, gz in cases:
        ref_hits = None
        for shift in (0, 1, 2, 3, 5, 8, 13, 15):
            buf = np.zeros(gz.size + 32, dtype=np.uint8)
            buf[shift: shift + gz.size] = gz
            d_buf = dev(buf)
            d_in = d_buf[shift: shift + gz.size]
            assert d_in.data_ptr() % 16 == (d_buf.data_ptr() + shift) % 16
            out = torch.zeros(raw.size, dtype=torch.uint8, device="cuda")
            say("unaligned", gz.size, "shift", shift)
            n, crc, st = gpu.decompress_device(d_in, out)
            assert n == raw.size and crc == zlib.crc32(raw.tobytes())
            assert np.array_equal(out.cpu().numpy(), raw)
            hits = gpu.find_blocks(d_in, 0, gz.size * 8, in_len=gz.size)
            if ref_hits is None:
                ref_hits = hits
            assert hits == ref_hits, shift


def test_autotune_keeps_output_exact(gpu):
    """The autotuner's pick is one of its candidates, stays in the context,
    and every member still decodes bit-exact under it."""
    import torch
    ctx = gpu.context()
    old = ctx.tuning()
    try:
        mem = W.batch_members(40, seed=77, max_size=2 << 20)
        ins = [m[1] for m in mem]
        raws = [m[0] for m in mem]
        in_off = np.cumsum([0] + [x.size for x in ins[:-1]])
        in_len = np.array([x.size for x in ins])
        out_cap = np.array([r.size for r in raws])
        out_off = np.cumsum([0] + list(out_cap[:-1]))
        d_in = dev(np.concatenate(ins))
        d_out = torch.zeros(int(out_cap.sum()) + 1, dtype=torch.uint8, device="cuda")
        best, ms = gpu.autotune(d_in, in_off, in_len, d_out, out_off, out_cap, budget_ms=500)
        say("autotune", best.chunk_bytes, best.pieces_per_chunk, best.waves_per_cu, best.batch_split_bytes, ms)
        assert ms > 0
        now = ctx.tuning()
        assert (now.chunk_bytes, now.pieces_per_chunk, now.waves_per_cu, now.batch_split_bytes) == \
            (best.chunk_bytes, best.pieces_per_chunk, best.waves_per_cu, best.batch_split_bytes)
        d_out.zero_()
        status, olen, crc, _ = gpu.decompress_batch(d_in, in_off, in_len, d_out, out_off, out_cap)
        assert (status == 0).all()
        host = d_out.cpu().numpy()
        for i, r in enumerate(raws):
            assert np.array_equal(host[out_off[i]: out_off[i] + r.size], r), i
    finally:
        ctx.set_tuning(old.chunk_bytes, old.pieces_per_chunk, old.waves_per_cu, old.batch_split_bytes)


def test_batch_block_finder_covers_every_member(gpu):
    """The batch runs one block-finder pass over all members to be split: its
    candidate count equals the sum of the counts found member by member, also
    when a member appears twice at the same input offset."""
    import torch
    ctx = gpu.context()
    old = ctx.tuning()
    ctx.set_tuning(chunk_bytes=64 << 10, pieces_per_chunk=4, batch_split_bytes=32 << 10)
    try:
        raws, ins = [], []
        for i, k in enumerate(["text", "u32", "text", "runs"]):
            raw = W.generate(k, (1 << 20) + 777 * i, 900 + i)
            raws.append(raw)
            ins.append(W.gzip_compress(raw, (1, 6, 9, 6)[i]))
        d_in = dev(np.concatenate(ins))
        in_off = np.cumsum([0] + [x.size for x in ins[:-1]])
        in_len = np.array([x.size for x in ins])
        cand_one = []
        for i in range(len(ins)):
            d_out = torch.zeros(raws[i].size, dtype=torch.uint8, device="cuda")
            status, olen, _, st = gpu.decompress_batch(d_in, [in_off[i]], [in_len[i]], d_out, [0], [raws[i].size])
            assert status[0] == 0
            cand_one.append(st["candidates"])
        say("per-member candidates", cand_one)
        assert sum(c > 0 for c in cand_one) >= 3  # the highly compressible member is not split
        # all members, member 1 twice (same input range)
        order = [0, 1, 2, 3, 1]
        out_cap = np.array([raws[i].size for i in order])
        out_off = np.cumsum([0] + list(out_cap[:-1]))
        d_out = torch.zeros(int(out_cap.sum()), dtype=torch.uint8, device="cuda")
        status, olen, _, st = gpu.decompress_batch(d_in, in_off[order], in_len[order], d_out, out_off, out_cap)
        assert (status == 0).all(), status
        assert st["candidates"] == sum(cand_one), (st["candidates"], cand_one)
        host = d_out.cpu().numpy()
        for j, i in enumerate(order):
            assert np.array_equal(host[out_off[j]: out_off[j] + raws[i].size], raws[i]), j
    finally:
        ctx.set_tuning(old.chunk_bytes, old.pieces_per_chunk, old.waves_per_cu, old.batch_split_bytes)


def test_batch_graph_replays_match_batch(gpu):
    """The hipGraph form: every replay equals gzgpu_decompress_batch on the
    same members (status, length, CRC, bytes), with torch work and syncs
    between replays; a byte changed in d_in between launches shows up in the
    next launch (the graph reads the current contents) and goes away when
    restored.  Bad members fail alone."""
    import torch
    mem = W.batch_members(300, seed=515, max_size=1 << 20)
    ins = [m[1].copy() for m in mem]
    raws = [m[0] for m in mem]
    ins[5] = np.frombuffer(b"not a gzip member at all....", np.uint8).copy()  # bad header
    raws[5] = np.zeros(0, np.uint8)
    bad_crc = ins[7].copy()
    bad_crc[-8] ^= 0x5A
    ins[7] = bad_crc
    in_off = np.cumsum([0] + [x.size for x in ins[:-1]])
    in_len = np.array([x.size for x in ins])
    out_cap = np.array([max(r.size, 1) for r in raws])
    out_cap[9] = max(1, raws[9].size // 2)  # too small
    out_off = np.cumsum([0] + list(out_cap[:-1]))
    host_in = np.concatenate(ins)
    d_in = dev(host_in)
    ref_out = torch.zeros(int(out_cap.sum()), dtype=torch.uint8, device="cuda")
    ref = gpu.decompress_batch(d_in, in_off, in_len, ref_out, out_off, out_cap, raise_on_error=False)
    d_out = torch.zeros_like(ref_out)
    with gpu.BatchGraph(d_in, in_off, in_len, d_out, out_off, out_cap) as g:
        for rep in range(3):
            d_out.zero_()
            torch.cuda.synchronize()
            status, olen, crc, st = g.launch(raise_on_error=False)
            say("graph replay", rep, "ms", round(st["ms_total"], 3), "bad", np.nonzero(status)[0][:5])
            assert np.array_equal(status, ref[0]), rep
            good = status == 0
            assert np.array_equal(olen[good], ref[1][good]) and np.array_equal(crc[good], ref[2][good])
            for i in np.nonzero(good)[0][:50]:
                assert np.array_equal(d_out[out_off[i]: out_off[i] + raws[i].size].cpu().numpy(), raws[i]), (rep, i)
        # change one member's trailer CRC in place, then restore it
        j = int(np.nonzero(ref[0] == 0)[0][3])
        pos = int(in_off[j] + in_len[j] - 8)
        d_in[pos] ^= 0xFF
        status, _, _, _ = g.launch(raise_on_error=False)
        assert status[j] != 0 and (status[np.arange(len(status)) != j] == ref[0][np.arange(len(status)) != j]).all()
        d_in[pos] ^= 0xFF
        status, _, _, _ = g.launch(raise_on_error=False)
        assert np.array_equal(status, ref[0])
    assert ref[0][5] != 0 and ref[0][7] != 0 and ref[0][9] != 0


def test_unsized_multi_member_stream_one_pipeline(gpu):
    """A pigz-style stream of 64 members without size fields: the members are
    found at their gzip magics and decode together in one pipeline (not one
    pipeline per member, each over the rest of the input)."""
    raws = [W.generate(("text", "u32", "runs", "random")[i % 4], 200_000 + 3331 * i, 500 + i) for i in range(64)]
    gz = b"".join(W.gzip_compress(r, (1, 6, 9)[i % 3]).tobytes() for i, r in enumerate(raws))
    st = {}
    out = gpu.decompress(gz, stats=st)
    say("unsized members", len(gz), {k: st[k] for k in ("pieces", "inflate_launches", "launches")})
    assert out == b"".join(r.tobytes() for r in raws)
    assert st["inflate_launches"] <= 4
    sink = io.BytesIO()
    assert gpu.decompress_to_writer(gz + b"trailing", sink) == len(out) and sink.getvalue() == out


def test_false_member_magic_inside_a_member(gpu):
    """Gzip magic bytes inside a member's payload (stored data that holds a
    gzip member) make a false candidate boundary: the span decode does not
    end at its trailer, and the member walk decodes the stream exactly."""
    inner = W.gzip_compress(W.generate("text", 300_000, 9), 6).tobytes()
    payload = bytes(W.generate("random", 100_000, 8)) + inner + bytes(W.generate("random", 100_000, 7))
    a = W.gzip_compress(np.frombuffer(payload, np.uint8), 0, encoder="zlib").tobytes()
    b = W.gzip_compress(W.generate("u32", 400_000, 6), 6).tobytes()
    for gz, want in ((a + b, payload + W.generate("u32", 400_000, 6).tobytes()), (a, payload)):
        st = {}
        assert gpu.decompress(gz, stats=st) == want
        say("false magic", len(gz), st["inflate_launches"])


@pytest.mark.parametrize("kind,level", [("random", 6), ("text", 0), ("random", 1)])
def test_stored_members_split_without_speculation(gpu, kind, level):
    """A member of stored blocks only (incompressible data, or level 0) is
    copied block by block from its LEN chain: no block finder, no
    speculation (StoredParallel, stored_split.rs:265)."""
    import torch
    raw = W.generate(kind, (40 << 20) + 4097, seed=60 + level)
    gz = W.gzip_compress(raw, level, encoder="zlib" if level == 0 else None)
    d_in = dev(gz)
    out = torch.zeros(raw.size, dtype=torch.uint8, device="cuda")
    n, crc, st = gpu.decompress_device(d_in, out)
    say("stored member", kind, level, gz.size, {k: st[k] for k in ("stored_members", "stored_blocks", "pieces",
                                                                    "candidates", "ms_total")})
    assert n == raw.size and crc == zlib.crc32(raw.tobytes())
    assert np.array_equal(out.cpu().numpy(), raw)
    assert st["stored_members"] == 1 and st["pieces"] == 0 and st["candidates"] == 0
    # at an odd output offset and through the host path too
    out2 = torch.zeros(raw.size + 3, dtype=torch.uint8, device="cuda")
    n, crc, st = gpu.decompress_device(d_in, out2[3:])
    assert n == raw.size and np.array_equal(out2[3:].cpu().numpy(), raw) and not out2[:3].any()
    assert gpu.decompress(gz) == raw.tobytes()
    # a corrupt payload byte: the CRC says so
    bad = gz.copy()
    bad[gz.size // 2] ^= 0x40
    with pytest.raises(gpu.GzippyError) as e:
        gpu.decompress(bad)
    assert e.value.kind == "CrcMismatch"


def test_stored_prefix_with_huffman_tail(gpu):
    """Stored blocks, then Huffman blocks in the same member (ratio still
    under 1.15): the LEN chain breaks at the first Huffman block, so the
    member takes the speculative path, byte-exact."""
    import struct
    a = W.generate("random", 24 << 20, 71)
    b = W.generate("text", 1 << 20, 72)
    c0 = zlib.compressobj(0, zlib.DEFLATED, -15)
    body = c0.compress(a.tobytes()) + c0.flush(zlib.Z_SYNC_FLUSH)
    c1 = zlib.compressobj(6, zlib.DEFLATED, -15, 9, zlib.Z_DEFAULT_STRATEGY, a[-32768:].tobytes())
    body += c1.compress(b.tobytes()) + c1.flush()
    raw = np.concatenate([a, b]).tobytes()
    gz = b"\x1f\x8b\x08\x00\x00\x00\x00\x00\x00\xff" + body + struct.pack("<II", zlib.crc32(raw), len(raw))
    st = {}
    assert gpu.decompress(gz, stats=st) == raw
    say("stored prefix", len(gz), {k: st[k] for k in ("stored_members", "pieces", "redecoded")})
    assert st["stored_members"] == 0


def test_streamed_output_and_late_crc_error(gpu):
    """A single large member reaches the writer range by range while the rest
    of it resolves; a CRC error is still reported, after output that equals
    the true bytes' prefix (partial output before a terminal error,
    single_member.rs:17-22)."""
    raw, gz = W.text_member(40 << 20, seed=17)

    class Rec:
        def __init__(self):
            self.parts = []

        def write(self, b):
            self.parts.append(bytes(b))
            return len(b)

    rec = Rec()
    assert gpu.decompress_to_writer(gz, rec) == raw.size
    assert b"".join(rec.parts) == raw.tobytes() and len(rec.parts) >= 5
    bad = gz.copy()
    bad[-6] ^= 0x01  # CRC
    rec = Rec()
    with pytest.raises(gpu.GzippyError) as e:
        gpu.decompress_to_writer(bad, rec)
    assert e.value.kind == "CrcMismatch"
    got = b"".join(rec.parts)
    assert raw.tobytes().startswith(got)
    say("streamed before the CRC error", len(got), "of", raw.size)


def test_crc32_entry_point_offsets_and_lengths(gpu):
    """gzgpu_crc32 (crc32_fold, crc32.rs:155) over a device buffer at every
    16-byte phase of its start, with lengths across the 16-byte block, the
    4096-byte fold stride and the 64 KiB segment boundaries, equals zlib's
    CRC32: the masked edge blocks and the divided-out pad bytes."""
    import torch
    rng = np.random.default_rng(5)
    host = rng.integers(0, 256, 300_000, dtype=np.uint8)
    d = torch.from_numpy(host).cuda()
    for off in (0, 1, 3, 8, 15, 16, 17, 4095):
        for n in (0, 1, 15, 16, 17, 255, 4096, 4097, 65535, 65536, 65537, 131077, 200_000):
            got = gpu.crc32(d[off:], length=n)
            assert got == zlib.crc32(host[off:off + n].tobytes()), (off, n)
