"""A small PNG encoder for the texture-loader tests (ISO/IEC 15948): any
colour type and bit depth, a chosen filter per scanline, optional PLTE/tRNS."""
import struct
import zlib

import numpy as np


def _chunk(kind, data):
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def encode(samples, ctype, depth, palette=None, trns=None, filters=(0, 1, 2, 3, 4)):
    """samples: (h, w, channels) integer array of raw sample values."""
    h, w, ch = samples.shape
    rows = []
    for y in range(h):
        if depth >= 8:
            vals = samples[y].reshape(-1).astype(np.uint64)
            row = b"".join(int(v).to_bytes(depth // 8, "big") for v in vals)
        else:
            bits = "".join(format(int(v), f"0{depth}b") for v in samples[y].reshape(-1))
            bits += "0" * (-len(bits) % 8)
            row = bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8))
        rows.append(np.frombuffer(row, np.uint8).astype(int))
    bpp = max(1, ch * depth // 8)
    raw = bytearray()
    prev = np.zeros_like(rows[0])
    for y, cur in enumerate(rows):
        ft = filters[y % len(filters)]
        out = []
        for x in range(len(cur)):
            a = cur[x - bpp] if x >= bpp else 0
            b = prev[x]
            c = prev[x - bpp] if x >= bpp else 0
            pred = [0, a, b, (a + b) // 2, _paeth(a, b, c)][ft]
            out.append((cur[x] - pred) & 0xFF)
        raw += bytes([ft]) + bytes(out)
        prev = cur
    png = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 0))
    if palette is not None:
        png += _chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    if trns is not None:
        png += _chunk(b"tRNS", bytes(np.asarray(trns, np.uint8)))
    return png + _chunk(b"IDAT", zlib.compress(bytes(raw))) + _chunk(b"IEND", b"")
