"""Wire format (server/app.py:23-41, 182-190, 220-224; web/js/net.js:9-29):
frames packed here decode with the reference viewer's own decoder (fixture
made by tests/golden/make_wire_golden.js under Node), and commands encoded by
the reference client parse here."""
import gzip
import json
import os
import struct

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def _golden():
    return json.load(open(os.path.join(GOLDEN, 'wire_golden.json')))


def test_frames_decode_with_the_reference_decoder():
    import sys
    sys.path.insert(0, GOLDEN)
    from make_wire_frames import cases
    from windtunnel import wire
    g = {f['name']: f for f in _golden()['frames']}
    for name, header, parts in cases():
        fr = wire.pack(header, parts)
        ref = g[name]
        mine = wire.decode_frame(fr)
        assert mine['header'] == ref['header']
        for k, arr in parts:
            assert ref[k] == arr.tolist()              # what the viewer reads
            assert np.array_equal(mine[k], arr)
        assert set(k for k in ('parts16', 'parts8', 'slice', 'cp') if k in ref) == {k for k, _ in parts}


def test_frame_layout_rules():
    from windtunnel import wire
    parts = [('parts16', np.arange(12, dtype=np.uint16)), ('parts8', np.arange(3, dtype=np.uint8)),
             ('cp', np.arange(-5, 6, dtype=np.int8))]
    fr = wire.pack({'type': 'frame', 'numP': 3, 'numVerts': 11}, parts)
    hl = struct.unpack_from('<I', fr, 0)[0]
    assert (4 + hl) % 8 == 0                            # payload 8-byte aligned (app.py:32)
    h = json.loads(fr[4:4 + hl])
    assert all(v % 8 == 0 for v in h['offsets'].values())
    assert h['offsets'] == {'parts16': 0, 'parts8': 24, 'cp': 32}
    assert len(fr) == 4 + hl + 32 + 16


def test_commands_from_the_reference_client_parse():
    from windtunnel import wire
    for c in _golden()['commands']:
        raw = bytes.fromhex(c['bytes'])
        h, payload = wire.parse_command(raw)
        assert h == c['header'] and list(payload) == c['payload']
        assert wire.command_bytes(c['header'], bytes(c['payload']), compress=False) == raw


def test_solid_and_surface_uploads():
    from windtunnel import wire
    dims = (8, 4, 6)
    mask = (np.arange(8 * 4 * 6) % 3 == 0).astype(np.uint8)
    h, p = wire.parse_command(wire.command_bytes({'type': 'solid', 'dims': list(dims)}, mask.tobytes()))
    assert np.array_equal(wire.solid_from_command(h, p, dims), mask)
    assert wire.solid_from_command({'dims': [8, 4, 5]}, p, dims) is None          # wrong dims: dropped
    assert wire.solid_from_command(h, gzip.compress(b'12'), dims) is None         # wrong size: dropped
    v = np.arange(6, dtype=np.float32).reshape(2, 3)
    n = np.eye(3, dtype=np.float32)[:2]
    h, p = wire.parse_command(wire.command_bytes({'type': 'surface', 'numVerts': 2},
                                                 np.concatenate([v.ravel(), n.ravel()]).tobytes()))
    vv, nn = wire.surface_from_command(h, p)
    assert np.array_equal(vv, v) and np.array_equal(nn, n)
    assert wire.surface_from_command({'numVerts': 3}, p) is None
    with pytest.raises(Exception):
        wire.solid_from_command({'dims': [2, 1, 1]}, b'not gzip', (2, 1, 1))
