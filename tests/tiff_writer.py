"""A small TIFF writer for the volume-ingest tests: one-channel integer pages,
classic or BigTIFF, either byte order, strips or tiles, compression none, LZW,
Deflate or PackBits, predictor none or horizontal.  Written from the TIFF 6.0
specification so that the reader is checked against an independent encoder."""
from __future__ import annotations

import struct
import zlib

import numpy as np

NONE, LZW, DEFLATE, PACKBITS = 1, 5, 8, 32773


def lzw_encode(data: bytes) -> bytes:
    """TIFF LZW: Clear first, MSB-first codes, width grows when the next free
    code reaches 2**width (the decoder, one entry behind, switches one code
    early), Clear again before the table fills, EndOfInformation last."""
    out = bytearray()
    buf = nbits = 0
    width = 9

    def put(code):
        nonlocal buf, nbits
        buf = (buf << width) | code
        nbits += width
        while nbits >= 8:
            nbits -= 8
            out.append((buf >> nbits) & 0xFF)
        buf &= (1 << nbits) - 1

    def reset():
        nonlocal width
        width = 9
        return {bytes([i]): i for i in range(256)}, 258

    put(256)
    table, nxt = reset()
    w = b""
    for c in data:
        wc = w + bytes([c])
        if wc in table:
            w = wc
            continue
        put(table[w])
        table[wc] = nxt
        nxt += 1
        if nxt == (1 << width) and width < 12:
            width += 1
        if nxt >= 4093:
            put(256)
            table, nxt = reset()
        w = bytes([c])
    if w:
        put(table[w])
        nxt += 1
        if nxt == (1 << width) and width < 12:
            width += 1
    put(257)
    if nbits:
        out.append((buf << (8 - nbits)) & 0xFF)
    return bytes(out)


def packbits_encode(data: bytes) -> bytes:
    out = bytearray()
    i, n = 0, len(data)
    while i < n:
        j = i
        while j + 1 < n and data[j + 1] == data[i] and j - i < 127:
            j += 1
        if j > i:
            out += bytes([(257 - (j - i + 1)) & 0xFF, data[i]])
            i = j + 1
            continue
        j = i
        while j < n and j - i < 128 and not (j + 1 < n and data[j + 1] == data[j]):
            j += 1
        j = max(j, i + 1)
        out.append(j - i - 1)
        out += data[i:j]
        i = j
    return bytes(out)


def _diff(block: np.ndarray) -> np.ndarray:
    """Horizontal differencing per row, modulo the sample width."""
    nat = block.astype(block.dtype.newbyteorder("="))
    u = nat.view(f"u{nat.dtype.itemsize}")
    d = u.copy()
    d[:, 1:] = u[:, 1:] - u[:, :-1]
    return d.view(nat.dtype)


def _encode(block: np.ndarray, order: str, compression: int, predictor: int) -> bytes:
    if predictor == 2:
        block = _diff(block)
    raw = np.ascontiguousarray(block).astype(block.dtype.newbyteorder(order)).tobytes()
    if compression == NONE:
        return raw
    if compression == LZW:
        return lzw_encode(raw)
    if compression == DEFLATE:
        return zlib.compress(raw)
    return packbits_encode(raw)


def write_tiff(path, pages, order="<", bigtiff=False, compression=NONE, predictor=1, rows_per_strip=None,
               tile=None):
    """pages: list of 2-D integer arrays of one dtype (height, width)."""
    e = order
    blob = bytearray(b"II" if e == "<" else b"MM")
    if bigtiff:
        blob += struct.pack(e + "HHHQ", 43, 8, 0, 0)
    else:
        blob += struct.pack(e + "HI", 42, 0)
    first_ptr = 8 if bigtiff else 4
    prev_ptr = first_ptr
    for page in pages:
        page = np.asarray(page)
        h, w = page.shape
        bits = page.dtype.itemsize * 8
        fmt = 2 if page.dtype.kind == "i" else 1
        blocks = []
        if tile:
            tw, th = tile
            for ty in range(0, h, th):
                for tx in range(0, w, tw):
                    b = np.zeros((th, tw), page.dtype)
                    part = page[ty:ty + th, tx:tx + tw]
                    b[:part.shape[0], :part.shape[1]] = part
                    blocks.append(_encode(b, e, compression, predictor))
        else:
            rps = rows_per_strip or h
            for r0 in range(0, h, rps):
                blocks.append(_encode(page[r0:r0 + rps], e, compression, predictor))
        offsets = []
        for b in blocks:
            offsets.append(len(blob))
            blob += b
            if len(blob) % 2:
                blob += b"\0"
        counts = [len(b) for b in blocks]
        tags = [(256, 4, [w]), (257, 4, [h]), (258, 3, [bits]), (259, 3, [compression]), (262, 3, [1]),
                (277, 3, [1]), (339, 3, [fmt])]
        if predictor != 1:
            tags.append((317, 3, [predictor]))
        off_type = 16 if bigtiff else 4
        if tile:
            tags += [(322, 4, [tile[0]]), (323, 4, [tile[1]]), (324, off_type, offsets), (325, 4, counts)]
        else:
            tags += [(273, off_type, offsets), (278, 4, [rows_per_strip or h]), (279, 4, counts)]
        tags.sort()
        sizes = {3: ("H", 2), 4: ("I", 4), 16: ("Q", 8)}
        inline = 8 if bigtiff else 4
        extra = {}
        for tag, typ, vals in tags:  # out-of-line values first
            f, sz = sizes[typ]
            if sz * len(vals) > inline:
                extra[tag] = len(blob)
                blob += struct.pack(e + f * len(vals), *vals)
                if len(blob) % 2:
                    blob += b"\0"
        ifd = len(blob)
        blob += struct.pack(e + ("Q" if bigtiff else "H"), len(tags))
        for tag, typ, vals in tags:
            f, sz = sizes[typ]
            blob += struct.pack(e + "HH", tag, typ) + struct.pack(e + ("Q" if bigtiff else "I"), len(vals))
            if tag in extra:
                blob += struct.pack(e + ("Q" if bigtiff else "I"), extra[tag])
            else:
                v = struct.pack(e + f * len(vals), *vals)
                blob += v + b"\0" * (inline - len(v))
        next_ptr = len(blob)
        blob += b"\0" * (8 if bigtiff else 4)
        blob[prev_ptr:prev_ptr + (8 if bigtiff else 4)] = struct.pack(e + ("Q" if bigtiff else "I"), ifd)
        prev_ptr = next_ptr
    with open(path, "wb") as fh:
        fh.write(bytes(blob))
