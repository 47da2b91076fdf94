"""Wire format between the solver and the browser viewer.

Server -> client frame (server/app.py:23-41, decoded by web/js/net.js:18-29):
    [u32 little-endian header length L][L bytes of JSON header, space-padded
    so that 4 + L is a multiple of 8][payload]
The header carries `offsets`: byte offset of each named part relative to the
start of the payload; each part starts 8-byte aligned. Parts the viewer reads:
parts16 (u16, 4 per particle), parts8 (u8, 1 per particle), slice (u8,
2 per pixel), cp (i8, 1 per vertex).

Client -> server binary command (web/js/net.js:9-16, parsed at
server/app.py:220-224): [u32 LE header length][JSON header][payload], the
payload gzip-compressed by the client (net.js:4-7).
"""
import gzip
import json
import struct

import numpy as np

PART_DTYPES = {'parts16': np.uint16, 'parts8': np.uint8, 'slice': np.uint8, 'cp': np.int8}


def pack(header, parts):
    """Frame bytes for (name, ndarray) parts, offsets added to the header."""
    offsets, pos = {}, 0
    for name, arr in parts:
        offsets[name] = pos
        pos += arr.nbytes
        pos += (-pos) % 8
    hb = json.dumps(dict(header, offsets=offsets)).encode()
    hb += b' ' * ((-(4 + len(hb))) % 8)
    out = bytearray(4 + len(hb) + pos)
    struct.pack_into('<I', out, 0, len(hb))
    out[4:4 + len(hb)] = hb
    base = 4 + len(hb)
    for name, arr in parts:
        b = np.ascontiguousarray(arr).tobytes()
        o = base + offsets[name]
        out[o:o + len(b)] = b
    return bytes(out)


def frame_header(solver, out, steps_since, mlups, running):
    """The frame header of server/app.py:182-184, 189-190."""
    h = {'type': 'frame', 'dims': [solver.nx, solver.ny, solver.nz], 'numP': solver.num_p,
         'numVerts': solver.num_verts, 'forces': out['forces'], 'steps': steps_since,
         'mlups': mlups if running else 0, 'running': running}
    if 'slice' in out:
        h.update(sliceW=out['sliceW'], sliceH=out['sliceH'], sliceAxis=out['sliceAxis'])
    return h


def frame(solver, out, steps_since, mlups, running):
    """Full frame for one Solver.frame() result (server/app.py:182-194)."""
    parts = [(k, out[k]) for k in ('parts16', 'parts8', 'slice', 'cp') if k in out]
    return pack(frame_header(solver, out, steps_since, mlups, running), parts)


def decode_frame(buf):
    """What the viewer reads from a frame (web/js/net.js:18-29)."""
    buf = bytes(buf)
    hl = struct.unpack_from('<I', buf, 0)[0]
    h = json.loads(buf[4:4 + hl].decode())
    base = 4 + hl
    o = h['offsets']
    d = {'header': h}
    counts = {'parts16': h.get('numP', 0) * 4, 'parts8': h.get('numP', 0),
              'slice': h.get('sliceW', 0) * h.get('sliceH', 0) * 2, 'cp': h.get('numVerts', 0)}
    for name, dt in PART_DTYPES.items():
        if name in o:
            d[name] = np.frombuffer(buf, dtype=dt, count=counts[name], offset=base + o[name])
    return d


def command_bytes(header, payload, compress=True):
    """Client binary command (web/js/net.js:9-16, gzip as net.js:4-7)."""
    hb = json.dumps(header, separators=(',', ':')).encode()   # JSON.stringify's compact form
    body = gzip.compress(bytes(payload)) if compress else bytes(payload)
    return struct.pack('<I', len(hb)) + hb + body


def parse_command(buf):
    """(header, raw payload) of a binary command (server/app.py:221-224)."""
    buf = bytes(buf)
    hl = struct.unpack_from('<I', buf, 0)[0]
    return json.loads(buf[4:4 + hl]), buf[4 + hl:]


def solid_from_command(header, payload, dims):
    """The mask of a 'solid' command, or None when dims or size do not match
    (server/app.py:138-143 drops such uploads silently)."""
    if tuple(header.get('dims', ())) != tuple(dims):
        return None
    arr = np.frombuffer(gzip.decompress(payload), dtype=np.uint8)
    return arr if arr.size == dims[0] * dims[1] * dims[2] else None


def surface_from_command(header, payload):
    """(verts, norms) of a 'surface' command, or None on a size mismatch (app.py:144-148)."""
    raw = np.frombuffer(gzip.decompress(payload), dtype=np.float32)
    nv = int(header['numVerts'])
    if raw.size != nv * 6:
        return None
    return raw[:nv * 3].reshape(-1, 3), raw[nv * 3:].reshape(-1, 3)
