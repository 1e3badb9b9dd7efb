"""The FLWA animation format restated in Python (TEST INFRASTRUCTURE ONLY).

Only the tests import this module, as the checker of the native recorder and player; the
product package never does.  It follows src/animation/format.rs of the reference:
AnimationHeader::write_to / read_from (:101-176, 52 bytes little-endian), FrameIndex
(:188-218, 16 bytes), encode_frame / decode_frame (:221-248, f32 little-endian), and the file
layout AnimationRecorder writes (recorder.rs:63-205: header, frames, index table) and
AnimationPlayer reads (player.rs:36-64).  The LZ4 decoder follows the published LZ4 block
format with the 4-byte size prefix lz4_flex::compress_prepend_size writes.
"""
from __future__ import annotations

import struct
from typing import List, Tuple

import numpy as np

MAGIC = b"FLWA"
VERSION = 1
HEADER_SIZE = 52
INDEX_SIZE = 16


def encode_header(width, height, depth, channels, frame_count, dt, compression=0, delta=False) -> bytes:
    flags = (compression & 0x0F) | (1 << 4 if delta else 0)
    return (MAGIC + struct.pack("<HHIIIIQf", VERSION, flags, width, height, depth, channels, frame_count, dt)
            + bytes(16))


def decode_header(b: bytes) -> dict:
    if b[:4] != MAGIC:
        raise ValueError("Invalid FLWA magic bytes")
    version, flags, w, h, d, c, n, dt = struct.unpack("<HHIIIIQf", b[4:36])
    if version != VERSION:
        raise ValueError(f"Unsupported FLWA version: {version}")
    return {"width": w, "height": h, "depth": d, "channels": c, "frame_count": n, "dt": dt,
            "compression": flags & 0x0F, "delta_encoding": bool(flags & 0x10)}


def encode_index(offset: int, size: int) -> bytes:
    return struct.pack("<QQ", offset, size)


def encode_frame(channels) -> bytes:
    return b"".join(np.asarray(c, dtype="<f4").tobytes() for c in channels)


def write_file(width, height, depth, channels, dt, frames: List[List[np.ndarray]], packed=None) -> bytes:
    """Header, frames (raw, or the given packed bytes per frame with compression 1), index table."""
    body, index = b"", []
    off = HEADER_SIZE
    for i, fr in enumerate(frames):
        data = packed[i] if packed is not None else encode_frame(fr)
        index.append((off, len(data)))
        body += data
        off += len(data)
    head = encode_header(width, height, depth, channels, len(frames), dt, 1 if packed is not None else 0)
    return head + body + b"".join(encode_index(o, s) for o, s in index)


def read_file(b: bytes) -> Tuple[dict, List[bytes]]:
    h = decode_header(b[:HEADER_SIZE])
    n = h["frame_count"]
    start = len(b) - n * INDEX_SIZE
    frames = []
    for i in range(n):
        off, size = struct.unpack("<QQ", b[start + i * INDEX_SIZE:start + (i + 1) * INDEX_SIZE])
        data = b[off:off + size]
        frames.append(lz4_decompress(data) if h["compression"] == 1 else data)
    return h, frames


def decode_frame(data: bytes, w, h, d, c) -> List[np.ndarray]:
    a = np.frombuffer(data, dtype="<f4").astype(np.float32)
    n = w * h * d
    assert a.size == n * c, "Frame size mismatch"
    return [a[i * n:(i + 1) * n].copy() for i in range(c)]


def lz4_decompress(src: bytes) -> bytes:
    """LZ4 block with a u32 little-endian size prefix."""
    size = struct.unpack("<I", src[:4])[0]
    out = bytearray()
    ip = 4
    while ip < len(src):
        tok = src[ip]
        ip += 1
        lit = tok >> 4
        if lit == 15:
            while True:
                b = src[ip]
                ip += 1
                lit += b
                if b != 255:
                    break
        out += src[ip:ip + lit]
        ip += lit
        if ip >= len(src):
            break
        off = src[ip] | (src[ip + 1] << 8)
        ip += 2
        ml = (tok & 15) + 4
        if (tok & 15) == 15:
            while True:
                b = src[ip]
                ip += 1
                ml += b
                if b != 255:
                    break
        if off == 0 or off > len(out):
            raise ValueError("corrupt LZ4 block")
        for _ in range(ml):
            out.append(out[-off])
    if len(out) != size:
        raise ValueError("corrupt LZ4 block")
    return bytes(out)


def lz4_literal_block(data: bytes) -> bytes:
    """A valid LZ4 block that stores `data` as one run of literals."""
    n = len(data)
    out = bytearray(struct.pack("<I", n))
    out.append((15 if n >= 15 else n) << 4)
    if n >= 15:
        r = n - 15
        while r >= 255:
            out.append(255)
            r -= 255
        out.append(r)
    return bytes(out + data)
