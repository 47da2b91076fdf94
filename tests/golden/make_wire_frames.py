"""Step 1 of the wire-format fixture: frames packed by windtunnel.wire.pack.
Writes tests/golden/wire_frames.bin + wire_frames.json (index). Step 2 is
make_wire_golden.js, which decodes them with the reference viewer's decoder."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, '..', '..', 'wind-tunnel_amd'))
from windtunnel import wire  # noqa: E402


def cases():
    rng = np.random.default_rng(5)
    yield 'forces_only', {'type': 'frame', 'dims': [64, 28, 36], 'numP': 0, 'numVerts': 0,
                          'forces': [0.125, -1.5e-3, 2.0], 'steps': 10, 'mlups': 1234.5, 'running': True}, []
    p = 7
    yield 'particles', {'type': 'frame', 'dims': [64, 28, 36], 'numP': p, 'numVerts': 0, 'forces': [1.0, 2.0, 3.0],
                        'steps': 3, 'mlups': 0, 'running': False}, \
        [('parts16', rng.integers(0, 65536, 4 * p).astype(np.uint16)), ('parts8', rng.integers(0, 256, p).astype(np.uint8))]
    w, h, nv = 5, 3, 11
    yield 'all_parts', {'type': 'frame', 'dims': [5, 3, 9], 'numP': 3, 'numVerts': nv, 'forces': [0.5, 0.25, -0.75],
                        'steps': 20, 'mlups': 99.0, 'running': True, 'sliceW': w, 'sliceH': h, 'sliceAxis': 2}, \
        [('parts16', rng.integers(0, 65536, 12).astype(np.uint16)), ('parts8', rng.integers(0, 256, 3).astype(np.uint8)),
         ('slice', rng.integers(0, 256, 2 * w * h).astype(np.uint8)), ('cp', rng.integers(-127, 128, nv).astype(np.int8))]


if __name__ == '__main__':
    index, blob = [], b''
    for name, header, parts in cases():
        fr = wire.pack(header, parts)
        index.append({'name': name, 'offset': len(blob), 'bytes': len(fr)})
        blob += fr + b'\0' * ((-len(fr)) % 8)
    open(os.path.join(HERE, 'wire_frames.bin'), 'wb').write(blob)
    json.dump(index, open(os.path.join(HERE, 'wire_frames.json'), 'w'), indent=1)
