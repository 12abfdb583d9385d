"""A small TIFF writer for tests: byte order, BigTIFF, strips or tiles, no
compression, Deflate or PackBits, the horizontal predictor and arbitrary
extra tags.  It writes what the TIFF 6.0 and BigTIFF specifications define,
independently of the library's reader."""
import struct
import zlib

import numpy as np

TYPES = {"BYTE": (1, "B"), "ASCII": (2, "s"), "SHORT": (3, "H"), "LONG": (4, "I"),
         "RATIONAL": (5, "II"), "LONG8": (16, "Q")}


def packbits(data):
    out = bytearray()
    i, n = 0, len(data)
    while i < n:
        j = i + 1
        while j < n and j - i < 128 and data[j] == data[i]:
            j += 1
        if j - i >= 2:
            out += struct.pack("b", -(j - i - 1)) + bytes([data[i]])
            i = j
            continue
        j = i + 1
        while j < n and j - i < 128 and not (j + 1 < n and data[j] == data[j + 1]):
            j += 1
        out += struct.pack("b", j - i - 1) + bytes(data[i:j])
        i = j
    return bytes(out)


def encode(block, compression, predictor):
    block = np.ascontiguousarray(block, np.uint8)
    if predictor == 2:
        d = block.astype(np.int16)
        d[:, 1:] = d[:, 1:] - d[:, :-1]
        block = (d & 0xFF).astype(np.uint8)
    raw = block.tobytes()
    if compression == 1:
        return raw
    if compression in (8, 32946):
        return zlib.compress(raw)
    if compression == 32773:
        return packbits(raw)
    raise ValueError(compression)


def write(path, pages, order="<", big=False, compression=1, predictor=1, rows_per_strip=None,
          tile=None, photometric=1, bits=8, samples=1, extra=None, first_page_extra=None):
    """pages: list of 2-D uint8 arrays (height, width).  extra / first_page_extra:
    {tag: (type name, values)} for every page / the first page only."""
    body = bytearray()
    base = 16 if big else 8
    ifd_offsets = []
    ifd_blobs = []

    def put(data):
        while (base + len(body)) % 2:
            body.append(0)
        off = base + len(body)
        body.extend(data)
        return off

    for pi, page in enumerate(pages):
        h, w = page.shape[:2]
        chunks = []
        if tile:
            tw, th = tile
            for ty in range(0, h, th):
                for tx in range(0, w, tw):
                    t = np.zeros((th, tw), np.uint8)
                    part = page[ty:ty + th, tx:tx + tw]
                    t[:part.shape[0], :part.shape[1]] = part
                    chunks.append(encode(t, compression, predictor))
        else:
            rps = rows_per_strip or h
            for y0 in range(0, h, rps):
                chunks.append(encode(page[y0:y0 + rps].reshape(-1, w * samples), compression,
                                     predictor))
        offs = [put(c) for c in chunks]
        counts = [len(c) for c in chunks]
        tags = {256: ("LONG", [w]), 257: ("LONG", [h]), 258: ("SHORT", [bits] * samples),
                259: ("SHORT", [compression]), 262: ("SHORT", [photometric]),
                277: ("SHORT", [samples])}
        if predictor != 1:
            tags[317] = ("SHORT", [predictor])
        offtype = "LONG8" if big else "LONG"
        if tile:
            tags.update({322: ("LONG", [tile[0]]), 323: ("LONG", [tile[1]]),
                         324: (offtype, offs), 325: (offtype, counts)})
        else:
            tags.update({273: (offtype, offs), 278: ("LONG", [rows_per_strip or h]),
                         279: (offtype, counts)})
        for src in (extra or {}), (first_page_extra or {}) if pi == 0 else {}:
            tags.update(src)
        ifd_blobs.append(tags)
    # IFDs after the data
    for tags in ifd_blobs:
        entries = []
        for tag in sorted(tags):
            tname, vals = tags[tag]
            code, f = TYPES[tname]
            if tname == "ASCII":
                data = vals.encode() + b"\0" if isinstance(vals, str) else bytes(vals)
                count = len(data)
            elif tname == "RATIONAL":
                data = b"".join(struct.pack(order + "II", a, b) for a, b in vals)
                count = len(vals)
            else:
                data = b"".join(struct.pack(order + f, v) for v in vals)
                count = len(vals)
            entries.append((tag, code, count, data))
        inline = 8 if big else 4
        resolved = []
        for tag, code, count, data in entries:
            if len(data) <= inline:
                resolved.append((tag, code, count, data.ljust(inline, b"\0"), None))
            else:
                resolved.append((tag, code, count, None, put(data)))
        while (base + len(body)) % 2:
            body.append(0)
        ifd_offsets.append(base + len(body))
        ifd = bytearray(struct.pack(order + ("Q" if big else "H"), len(resolved)))
        for tag, code, count, data, off in resolved:
            ifd += struct.pack(order + "HH", tag, code)
            ifd += struct.pack(order + ("Q" if big else "I"), count)
            ifd += data if data is not None else struct.pack(order + ("Q" if big else "I"), off)
        ifd += b"\0" * (8 if big else 4)  # next offset, patched below
        body.extend(ifd)
    # chain the IFDs
    for i, off in enumerate(ifd_offsets):
        nxt = ifd_offsets[i + 1] if i + 1 < len(ifd_offsets) else 0
        n = struct.unpack_from(order + ("Q" if big else "H"), body, off - base)[0]
        at = off - base + (8 if big else 2) + n * (20 if big else 12)
        struct.pack_into(order + ("Q" if big else "I"), body, at, nxt)
    bom = b"II" if order == "<" else b"MM"
    if big:
        head = bom + struct.pack(order + "HHHQ", 43, 8, 0, ifd_offsets[0])
    else:
        head = bom + struct.pack(order + "HI", 42, ifd_offsets[0])
    with open(path, "wb") as fh:
        fh.write(head + bytes(body))
