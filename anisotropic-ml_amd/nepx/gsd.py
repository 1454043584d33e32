"""Minimal reader/writer of GSD 2.x trajectory files (HOOMD schema), no `gsd` package needed.

File layout (GSD 2.0 format): a 256-byte header (magic 0x65DF65DF65DF65DF,
index location and capacity, name-list location and capacity, schema and
file versions, application and schema strings), a sorted index of 32-byte
entries (frame u64, N u64, location i64, M u32, name id u16, type u8, flags
u8) and a name list of NUL-terminated chunk names; chunk data is raw
little-endian N x M arrays.

Replaces gsd.hoomd as used by nep_simulator/GsdHandler.py:9-69 (read the last
frame) and nep_simulator/Sim.py:117-144 (append a frame every dump period).
Frames written here also carry the thermostat state under
log/nepx/thermostat so a run can resume exactly (the reference drops it and
re-hard-codes it, Sim.py:99-111).
"""
import os
import struct

import numpy as np

MAGIC = 0x65DF65DF65DF65DF
HEADER = struct.Struct("<QQQQQII64s64s80s")
ENTRY = struct.Struct("<QQqIHBB")
NAME_SIZE = 64
TYPES = {1: np.uint8, 2: np.uint16, 3: np.uint32, 4: np.uint64, 5: np.int8, 6: np.int16, 7: np.int32,
         8: np.int64, 9: np.float32, 10: np.float64, 11: np.uint8}
TYPE_ID = {np.dtype(v).str: k for k, v in TYPES.items() if k != 11}

# HOOMD schema defaults for chunks absent from a frame (and from frame 0)
DEFAULTS = {
    "configuration/step": lambda n: np.array([0], dtype=np.uint64),
    "configuration/box": lambda n: np.array([1, 1, 1, 0, 0, 0], dtype=np.float32),
    "particles/mass": lambda n: np.ones(n, dtype=np.float32),
    "particles/typeid": lambda n: np.zeros(n, dtype=np.uint32),
    "particles/orientation": lambda n: np.tile(np.array([1, 0, 0, 0], dtype=np.float32), (n, 1)),
    "particles/velocity": lambda n: np.zeros((n, 3), dtype=np.float32),
    "particles/angmom": lambda n: np.zeros((n, 4), dtype=np.float32),
    "particles/moment_inertia": lambda n: np.zeros((n, 3), dtype=np.float32),
    "particles/position": lambda n: np.zeros((n, 3), dtype=np.float32),
    "particles/body": lambda n: np.full(n, -1, dtype=np.int32),
}


class GsdError(ValueError):
    pass


def _read_index(f):
    f.seek(0)
    raw = f.read(HEADER.size)
    if len(raw) < HEADER.size:
        raise GsdError("file too short for a GSD header")
    (magic, idx_loc, idx_n, nl_loc, nl_n, schema_v, gsd_v, app, schema, _) = HEADER.unpack(raw)
    if magic != MAGIC:
        raise GsdError("not a GSD file (bad magic)")
    if (gsd_v >> 16) < 2:
        raise GsdError(f"GSD file version {gsd_v >> 16}.{gsd_v & 0xffff} not supported (need 2.x)")
    f.seek(nl_loc)
    names = [n.decode() for n in f.read(nl_n * NAME_SIZE).split(b"\0")]
    f.seek(idx_loc)
    buf = f.read(idx_n * ENTRY.size)
    entries = []
    for k in range(idx_n):
        e = ENTRY.unpack_from(buf, k * ENTRY.size)
        if e[2] == 0:
            break
        entries.append(e)
    return dict(names=names, entries=entries, application=app.split(b"\0")[0].decode(),
                schema=schema.split(b"\0")[0].decode())


def n_frames(path):
    with open(path, "rb") as f:
        idx = _read_index(f)
    return 1 + max((e[0] for e in idx["entries"]), default=-1)


def read_frame(path, frame=-1):
    """All chunks of one frame as numpy arrays (frame-0 and schema defaults filled in)."""
    with open(path, "rb") as f:
        idx = _read_index(f)
        nf = 1 + max((e[0] for e in idx["entries"]), default=-1)
        if nf == 0:
            raise GsdError("GSD file has no frames")
        fr = nf - 1 if frame < 0 else int(frame)
        if fr >= nf:
            raise GsdError(f"frame {fr} out of range ({nf} frames)")
        chunks = {}
        for want in (0, fr):
            for (efr, N, loc, M, nid, typ, _flags) in idx["entries"]:
                if efr != want:
                    continue
                name = idx["names"][nid]
                if want == 0 and fr != 0 and name.startswith("log/"):
                    continue
                f.seek(loc)
                dt = np.dtype(TYPES[typ]).newbyteorder("<")
                a = np.frombuffer(f.read(N * M * dt.itemsize), dtype=dt).copy()
                if typ == 11:
                    a = a.tobytes()
                elif M > 1:
                    a = a.reshape(N, M)
                chunks[name] = a
    n = int(chunks["particles/N"][0]) if "particles/N" in chunks else 0
    for k, fn in DEFAULTS.items():
        if k not in chunks:
            chunks[k] = fn(n)
    chunks["_frame"] = fr
    chunks["_frames"] = nf
    return chunks


class GsdWriter:
    """Append frames to a GSD 2.0 file.

    Each append writes the frame's chunks at the end of the file, then a fresh
    index and name list after them, then rewrites the header to point there.
    mode 'w' truncates, 'a' appends to an existing file.
    """

    def __init__(self, path, mode="w", application="nepx", schema="hoomd", schema_version=(1, 4)):
        self.path = path
        self.names = []
        self.entries = []
        self.app = application
        self.schema = schema
        self.schema_version = schema_version
        if mode == "a" and os.path.exists(path):
            with open(path, "rb") as f:
                idx = _read_index(f)
            self.names = [n for n in idx["names"] if n]
            self.entries = list(idx["entries"])
        elif mode not in ("w", "a"):
            raise ValueError(mode)
        else:
            with open(path, "wb") as f:
                f.write(b"\0" * HEADER.size)
            self._finish()

    @property
    def frames(self):
        return 1 + max((e[0] for e in self.entries), default=-1)

    def _name_id(self, name):
        if name not in self.names:
            self.names.append(name)
        return self.names.index(name)

    def append(self, chunks):
        """chunks: dict name -> numpy array (1-D or 2-D) or bytes (character chunk)."""
        frame = self.frames
        with open(self.path, "r+b") as f:
            f.seek(0, os.SEEK_END)
            for name in sorted(chunks):
                a = chunks[name]
                if isinstance(a, (bytes, bytearray)):
                    data, N, M, typ = bytes(a), len(a), 1, 11
                else:
                    a = np.ascontiguousarray(a)
                    if a.ndim == 0:
                        a = a.reshape(1)
                    dt = a.dtype.newbyteorder("<")
                    a = a.astype(dt, copy=False)
                    if dt.str not in TYPE_ID:
                        raise GsdError(f"unsupported dtype {a.dtype} for chunk {name}")
                    typ = TYPE_ID[dt.str]
                    N = a.shape[0]
                    M = 1 if a.ndim == 1 else int(np.prod(a.shape[1:]))
                    data = a.tobytes()
                loc = f.tell()
                f.write(data)
                self.entries.append((frame, N, loc, M, self._name_id(name), typ, 0))
        self._finish()

    def _finish(self):
        with open(self.path, "r+b") as f:
            f.seek(0, os.SEEK_END)
            idx_loc = f.tell()
            cap = max(16, len(self.entries))
            ents = sorted(self.entries, key=lambda e: (e[0], e[4]))
            buf = bytearray(cap * ENTRY.size)
            for k, e in enumerate(ents):
                ENTRY.pack_into(buf, k * ENTRY.size, *e)
            f.write(buf)
            nl_loc = f.tell()
            blob = b"".join(n.encode() + b"\0" for n in self.names)
            nl_n = max(1, (len(blob) + NAME_SIZE - 1) // NAME_SIZE)
            f.write(blob + b"\0" * (nl_n * NAME_SIZE - len(blob)))
            sv = (self.schema_version[0] << 16) | self.schema_version[1]
            hdr = HEADER.pack(MAGIC, idx_loc, cap, nl_loc, nl_n, sv, (2 << 16) | 0,
                              self.app.encode()[:63], self.schema.encode()[:63], b"\0" * 80)
            f.seek(0)
            f.write(hdr)
