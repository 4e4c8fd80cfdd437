"""Serving loop (continuous batching) through the C ABI: every result equals
decoding that request alone and the generator's bytes, for single members,
multi-member streams, a false member boundary inside a member, non-gzip input
and trailing garbage."""
import zlib

import numpy as np
import pytest

from gzippy_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def stored_member_with_magic():
    """A level-0 member whose payload holds a gzip magic: the member-boundary
    scan splits it wrongly and the request must be re-decoded as a whole."""
    raw = np.frombuffer(b"A" * 5000 + bytes([0x1F, 0x8B, 0x08, 0x00]) + b"B" * 30 + bytes(20) + b"C" * 4000,
                        dtype=np.uint8).copy()
    co = zlib.compressobj(0, zlib.DEFLATED, 31)
    gz = np.frombuffer(co.compress(raw.tobytes()) + co.flush(), dtype=np.uint8).copy()
    return raw, gz


def build_requests(seed):
    rng = np.random.default_rng(seed)
    reqs = []  # (gz bytes, expected raw bytes)
    kinds = ["text", "random", "runs", "u32"]
    for i in range(90):
        k = kinds[i % 4]
        n = int(np.clip(rng.lognormal(np.log(32 << 10), 1.2), 100, 2 << 20))
        raw = W.generate(k, n, seed * 1000 + i)
        reqs.append((W.gzip_compress(raw, (1, 6, 9)[i % 3]).tobytes(), raw.tobytes()))
    for i in range(20):
        parts, raws = [], []
        for j in range(int(rng.integers(4, 13))):
            n = int(rng.integers(50, 200 << 10))
            raw = W.generate(kinds[(i + j) % 4], n, seed * 7777 + 31 * i + j)
            raws.append(raw.tobytes())
            parts.append(W.gzip_compress(raw, 6).tobytes())
        reqs.append((b"".join(parts), b"".join(raws)))
    raw, gz = stored_member_with_magic()
    reqs.append((gz.tobytes(), raw.tobytes()))
    reqs.append((b"this is not gzip at all", b""))
    raw = W.generate("text", 70_000, seed + 5)
    reqs.append((W.gzip_compress(raw, 6).tobytes() + b"\x00trailing garbage", raw.tobytes()))
    raw = W.generate("text", 3 << 20, seed + 6)  # large enough to be split into pieces
    reqs.append((W.gzip_compress(raw, 6).tobytes(), raw.tobytes()))
    return reqs


def test_served_results_equal_alone_and_reference(gpu):
    import gzippy_amd
    reqs = build_requests(91)
    say("requests", len(reqs), "bytes", sum(len(g) for g, _ in reqs))
    with gzippy_amd.Server(workers=2) as srv:
        tickets = [srv.submit(g) for g, _ in reqs]
        infos = [srv.poll(t, -1) for t in tickets]
        outs = [srv.fetch(t, inf) for t, inf in zip(tickets, infos)]
        st = srv.stats()
    say("server stats", st)
    assert st["requests_done"] == len(reqs)
    assert st["batches"] < len(reqs)  # requests shared batches
    assert st["fallbacks"] >= 1  # the false boundary was caught and re-decoded
    for i, ((g, raw), out, inf) in enumerate(zip(reqs, outs, infos)):
        assert inf["status"] == 0, (i, inf)
        assert out == raw, i
        assert out == gzippy_amd.decompress(g), i  # the same request alone
    # multi-member requests report their member count
    assert max(inf["members"] for inf in infos) >= 4


def test_served_corrupt_request_fails_alone(gpu):
    import gzippy_amd
    raw = W.generate("text", 200_000, 5)
    good = W.gzip_compress(raw, 6).tobytes()
    bad = bytearray(good)
    bad[-8] ^= 0xFF  # CRC32 of the trailer
    with gzippy_amd.Server() as srv:
        tg, tb = srv.submit(good), srv.submit(bytes(bad))
        assert srv.result(tg) == raw.tobytes()
        with pytest.raises(gzippy_amd.GzippyError):
            srv.result(tb)
