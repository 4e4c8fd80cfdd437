"""Writes tests/golden/silesia_blocks.npz from the reference's frozen GZBLK01
corpus (corpus/silesia_blocks/*.bin; format at
tools/extract_blocks/src/main.rs:7-31).

Kept per block: the 32 KiB predecessor window, the compressed bytes, the
block's bit offset and bit length, and the expected output as (length,
CRC32, SHA-256) — enough to pin a byte-exact decode without storing the
decoded bytes themselves.  Run once in a tree where the reference exists:

    python tests/golden/make_silesia_fixtures.py
"""
import glob
import hashlib
import os
import struct
import sys

import numpy as np

SRC = "/root/reference/corpus/silesia_blocks"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "silesia_blocks.npz")


def parse(path):
    b = open(path, "rb").read()
    (magic, idx, btype, bitoff, bitlen, bytelen, declen, predlen, nlit, ndist, mxl, mxd, mc, crc,
     _res) = struct.unpack("<8sIIQQQIIHHBBHII", b[:64])
    assert magic == b"GZBLK01\0", path
    o = 64
    pred = b[o:o + predlen]; o += predlen
    comp = b[o:o + bytelen]; o += bytelen
    dec = b[o:o + declen]; o += declen
    lcl = b[o:o + nlit]; o += nlit
    dcl = b[o:o + ndist]; o += ndist
    assert o == len(b), path
    return dict(idx=idx, btype=btype, bitoff=bitoff, bitlen=bitlen, pred=pred, comp=comp, dec=dec,
                crc=crc, lcl=lcl, dcl=dcl)


def main():
    files = sorted(glob.glob(os.path.join(SRC, "*.bin")))
    if not files:
        sys.exit("reference corpus not found at " + SRC)
    blocks = [parse(f) for f in files]
    arrays = {}
    meta = []
    for i, blk in enumerate(blocks):
        arrays[f"win{i}"] = np.frombuffer(blk["pred"], dtype=np.uint8)
        arrays[f"comp{i}"] = np.frombuffer(blk["comp"], dtype=np.uint8)
        arrays[f"lcl{i}"] = np.frombuffer(blk["lcl"], dtype=np.uint8)
        arrays[f"dcl{i}"] = np.frombuffer(blk["dcl"], dtype=np.uint8)
        sha = np.frombuffer(hashlib.sha256(blk["dec"]).digest(), dtype=np.uint8)
        arrays[f"sha{i}"] = sha
        meta.append([blk["idx"], blk["btype"], blk["bitoff"], blk["bitlen"], len(blk["dec"]), blk["crc"]])
    arrays["meta"] = np.array(meta, dtype=np.uint64)
    np.savez_compressed(DST, **arrays)
    print("wrote", DST, os.path.getsize(DST), "bytes,", len(blocks), "blocks")


if __name__ == "__main__":
    main()
