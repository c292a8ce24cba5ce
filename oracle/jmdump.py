"""Reader/writer of the record files exchanged with the reference harness
(oracle/ref_harness.cpp): 8-byte magic ``JMDUMP01`` then records of
``char name[32]; int32 kind; int32 pad; int64 count; payload`` with kind
0 = float64, 1 = int32, 2 = bytes. Test infrastructure only."""
from __future__ import annotations

import struct

import numpy as np

MAGIC = b"JMDUMP01"
_DT = {0: np.float64, 1: np.int32, 2: np.uint8}


def read(path: str) -> dict:
    out = {}
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            raise ValueError(f"{path}: not a JMDUMP01 file")
        while True:
            hdr = f.read(48)
            if len(hdr) < 48:
                break
            name = hdr[:32].split(b"\0", 1)[0].decode()
            kind, _pad, count = struct.unpack("<iiq", hdr[32:])
            dt = _DT[kind]
            data = np.frombuffer(f.read(count * np.dtype(dt).itemsize), dtype=dt).copy()
            out[name] = data
    return out


def write(path: str, records: dict) -> None:
    with open(path, "wb") as f:
        f.write(MAGIC)
        for name, arr in records.items():
            a = np.asarray(arr)
            if a.dtype == np.float64:
                kind = 0
            elif a.dtype == np.int32:
                kind = 1
            else:
                raise TypeError(f"{name}: unsupported dtype {a.dtype}")
            nm = name.encode()[:31].ljust(32, b"\0")
            f.write(nm + struct.pack("<iiq", kind, 0, a.size))
            f.write(np.ascontiguousarray(a).tobytes())


def zones(d: dict) -> list:
    """[(name, kind, ifa_f, ifa_l)] in the reference's faZoneList order."""
    names = bytes(d["ZONE_NAMES"]).decode().split("\n")[:-1]
    z = d["ZONES"].reshape(-1, 3)
    return [(n, int(k), int(f), int(l)) for n, (k, f, l) in zip(names, z)]
