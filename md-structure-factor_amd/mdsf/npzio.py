"""Parallel ``.npz`` writer (what ``np.savez_compressed`` writes, dens.py:346).

``np.savez_compressed`` deflates every member on one core; at 256^3 the S(q)
outputs are ~1 GB and the write dominates the wall time of ``compute_sf``.
Here each member's bytes (``.npy`` header + C-order data) are cut into
chunks that are deflated on a thread pool (zlib releases the GIL).  Every
chunk but the last ends with a full flush, so the chunks concatenate into one
valid raw deflate stream; the member's CRC-32 is computed on the calling
thread while the pool compresses.
The archive is a plain ZIP (ZIP64 records where sizes need them) that
``np.load`` / ``zipfile`` read like any other.  A member may be given as a
``concurrent.futures.Future`` of its array; it is resolved when the writer
reaches it, so the caller can produce later members while earlier ones are
deflated.
"""
import concurrent.futures
import io
import os
import struct
import time
import zlib

import numpy as np

CHUNK = 8 << 20
_ZIP64_LIMIT = 0xFFFFFFFF


def _npy_header(arr):
    d = np.lib.format.header_data_from_array_1_0(arr)
    bio = io.BytesIO()
    try:
        np.lib.format.write_array_header_1_0(bio, d)
    except ValueError:  # header too large for format 1.0
        bio = io.BytesIO()
        np.lib.format.write_array_header_2_0(bio, d)
    return bio.getvalue()


def _deflate(view, level, last):
    c = zlib.compressobj(level, zlib.DEFLATED, -15)
    out = c.compress(view)
    return out + c.flush(zlib.Z_FINISH if last else zlib.Z_FULL_FLUSH)


def _dos_time(t):
    lt = time.localtime(t)
    dt = ((lt.tm_year - 1980) << 9) | (lt.tm_mon << 5) | lt.tm_mday
    tm = (lt.tm_hour << 11) | (lt.tm_min << 5) | (lt.tm_sec // 2)
    return tm, dt


def savez_compressed(path, threads=None, level=6, force_zip64=False, **arrays):
    """Write ``arrays`` to ``path`` (".npz" appended if missing), deflated in parallel."""
    path = str(path)
    if not path.endswith(".npz"):
        path += ".npz"
    threads = threads or min(16, os.cpu_count() or 1)
    tm, dt = _dos_time(time.time())
    entries = []
    with open(path, "wb") as f, concurrent.futures.ThreadPoolExecutor(threads) as pool:
        for key, val in arrays.items():
            if isinstance(val, concurrent.futures.Future):
                val = val.result()   # produced while earlier members were deflated
            arr = np.asanyarray(val)
            if arr.dtype.hasobject:
                raise TypeError(f"{key}: object arrays are not written")
            if not arr.flags.c_contiguous:
                arr = arr.copy(order="C")   # (np.ascontiguousarray would make 0-d arrays 1-d)
            name = (key + ".npy").encode()
            head = _npy_header(arr)
            data = memoryview(arr.reshape(-1).view(np.uint8)) if arr.size else memoryview(b"")
            parts = [memoryview(head)] + [data[i:i + CHUNK] for i in range(0, len(data), CHUNK)]
            futs = [pool.submit(_deflate, p, level, i == len(parts) - 1)
                    for i, p in enumerate(parts)]
            crc, usize = 0, 0
            for p in parts:
                crc = zlib.crc32(p, crc)
                usize += len(p)
            res = [fu.result() for fu in futs]
            csize = sum(len(b) for b in res)
            offset = f.tell()
            z64 = force_zip64 or usize >= _ZIP64_LIMIT or csize >= _ZIP64_LIMIT
            extra = struct.pack("<HHQQ", 1, 16, usize, csize) if z64 else b""
            f.write(struct.pack("<IHHHHHIIIHH", 0x04034B50, 45 if z64 else 20, 0, 8, tm, dt,
                                crc & 0xFFFFFFFF, _ZIP64_LIMIT if z64 else csize,
                                _ZIP64_LIMIT if z64 else usize, len(name), len(extra)))
            f.write(name)
            f.write(extra)
            for b in res:
                f.write(b)
            entries.append((name, crc, csize, usize, offset, z64))
        cd_start = f.tell()
        for name, crc, csize, usize, offset, z64 in entries:
            big_off = offset >= _ZIP64_LIMIT
            fields = []
            if z64:
                fields += [usize, csize]
            if big_off or (z64 and force_zip64):
                fields.append(offset)
            extra = struct.pack("<HH" + "Q" * len(fields), 1, 8 * len(fields), *fields) \
                if fields else b""
            f.write(struct.pack("<IHHHHHHIIIHHHHHII", 0x02014B50, 45, 45 if z64 else 20, 0, 8,
                                tm, dt, crc & 0xFFFFFFFF, _ZIP64_LIMIT if z64 else csize,
                                _ZIP64_LIMIT if z64 else usize, len(name), len(extra), 0, 0, 0,
                                0, _ZIP64_LIMIT if (big_off or (z64 and force_zip64)) else offset))
            f.write(name)
            f.write(extra)
        cd_end = f.tell()
        cd_size = cd_end - cd_start
        n = len(entries)
        if force_zip64 or cd_start >= _ZIP64_LIMIT or n >= 0xFFFF:
            f.write(struct.pack("<IQHHIIQQQQ", 0x06064B50, 44, 45, 45, 0, 0, n, n, cd_size,
                                cd_start))
            f.write(struct.pack("<IIQI", 0x07064B50, 0, cd_end, 1))
            f.write(struct.pack("<IHHHHIIH", 0x06054B50, 0, 0, 0xFFFF, 0xFFFF, 0xFFFFFFFF,
                                0xFFFFFFFF, 0))
        else:
            f.write(struct.pack("<IHHHHIIH", 0x06054B50, 0, 0, n, n, cd_size, cd_start, 0))
    return path
