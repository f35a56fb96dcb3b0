"""An independent encoder of the snapshot format documented in
include/slimemold.h, for tests: builds a snapshot from arrays."""
import struct

import numpy as np

MAGIC = b"SLMSNAP1"


def checksum(data: bytes) -> int:
    """FNV-1a over little-endian 64-bit words, the last one zero-padded."""
    pad = (-len(data)) % 8
    words = np.frombuffer(data + b"\0" * pad, dtype="<u8")
    h = 0xcbf29ce484222325
    for w in words.tolist():
        h = ((h ^ w) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def build(width, height, agents, trail_raw, storage=0, flags=0, steps=0, reassigns=0, seed=0,
          params=(10.0, 0.0, 30.0, 50.0, 0.43, 0.3, 20.0, 1.0, 1.0, 2.0, 1.0), xy64=None,
          version=1):
    agents = np.ascontiguousarray(agents, dtype="<f4")
    n = agents.shape[0]
    body = agents.tobytes()
    if flags & 2:
        body += np.ascontiguousarray(xy64, dtype="<f8").tobytes()
    trail = np.ascontiguousarray(trail_raw).tobytes()
    head = MAGIC + struct.pack("<IIIIQQiIQQ", version, 128, width, height, n, seed, storage, flags,
                               steps, reassigns)
    head += struct.pack("<11f", *params) + struct.pack("<I", 0)
    head += struct.pack("<QQ", len(body), len(trail))
    assert len(head) == 128
    data = head + body + trail
    return data + struct.pack("<Q", checksum(data))
