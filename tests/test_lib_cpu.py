"""Host-side checks of libwt.so that need no GPU: the C ABI loads and exports
every entry point include/wt.h declares, the host geometry matches the
reference voxelizer bit for bit, and the size/admission arithmetic holds."""
import ctypes
import json
import os
import re
import subprocess
import zlib

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, 'include', 'wt.h')
GOLDEN = os.path.join(REPO, 'tests', 'golden')


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(wt_[a-z_0-9]+)\s*\(', text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    from windtunnel import _lib
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 20, names
    out = subprocess.run(['nm', '-D', '--defined-only', _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if ' T ' in l)
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert hasattr(lib, n)
    # every declared function has a ctypes signature in the binding, and vice versa
    assert sorted(_lib.SIGNATURES) == names


def test_error_reporting_without_gpu_calls():
    from windtunnel import _lib
    lib = _lib.load()
    s = (ctypes.c_int64 * 6)()
    assert lib.wt_buffer_sizes(2, 8, 8, 0, s) < 0
    assert 'at least 3' in _lib.last_error()
    assert lib.wt_buffer_sizes(8, 8, 8, 9, s) < 0
    assert 'storage form' in _lib.last_error()
    assert lib.wt_version() == 1


def test_launch_shape_is_checked_per_grid_before_any_launch():
    """The launch shape travels in the grid descriptor of every call; a bad one is
    rejected before anything reaches the GPU (the fake buffer addresses are never
    dereferenced)."""
    from windtunnel import _lib
    lib = _lib.load()
    fake = ctypes.c_void_p(0x1000)
    for cpt, order, what in ((3, 0, 'cells per thread'), (1, 8, 'tile order')):
        g = _lib.Grid(16, 8, 8, 0, cpt, order, (ctypes.c_void_p * 2)(0x1000, 0x2000), 0x3000, 0x4000,
                      (ctypes.c_void_p * 2)(0x5000, 0x6000), 0x7000, 0x8000)
        p = _lib.Params(0.05, 0.52, 0.16)
        assert lib.wt_step(ctypes.byref(g), 0, 1, ctypes.byref(p), 0, None) < 0
        assert what in _lib.last_error()
    del fake


@pytest.mark.parametrize('form,field,addr', [(0, 'mac', 0x4008), (1, 'cls', 0x3004), (0, 'f1', 0x2002),
                                             (3, 'shell0', 0x5010), (0, 'partials', 0x7004)])
def test_misaligned_grid_buffers_are_rejected_before_any_launch(form, field, addr):
    """Grid buffers are 16-byte aligned (the FP64 face snapshot 32, partials and
    sums 8): a misaligned one is an error, not a kernel reading across it (the fake
    buffer addresses are never dereferenced)."""
    from windtunnel import _lib
    lib = _lib.load()
    f = [0x1000, 0x2000]
    sh = [0x5000, 0x6000]
    cls, mac, part = 0x3000, 0x4000, 0x7000
    if field == 'mac':
        mac = addr
    elif field == 'cls':
        cls = addr
    elif field == 'f1':
        f[1] = addr
    elif field == 'shell0':
        sh[0] = addr
    else:
        part = addr
    g = _lib.Grid(16, 8, 8, form, 1, 0, (ctypes.c_void_p * 2)(*f), cls, mac, (ctypes.c_void_p * 2)(*sh), part,
                  0x8000)
    p = _lib.Params(0.05, 0.52, 0.16)
    assert lib.wt_step(ctypes.byref(g), 0, 1, ctypes.byref(p), 0, None) < 0
    assert 'aligned' in _lib.last_error()


def test_noalloc_step_rejects_tunnels_beyond_its_scratch():
    """wt_step_noalloc serves tunnels of at most 2^26 cells and says so, before any
    device call (the fake buffer addresses are never dereferenced)."""
    from windtunnel import _lib
    lib = _lib.load()
    g = _lib.Grid(512, 512, 257, 0, 1, 0, (ctypes.c_void_p * 2)(0x1000, 0x2000), 0x3000, 0x4000,
                  (ctypes.c_void_p * 2)(0x5000, 0x6000), 0, 0x8000)
    p = _lib.Params(0.05, 0.52, 0.16)
    assert lib.wt_step_noalloc(ctypes.byref(g), 0, 1, ctypes.byref(p), 1, None) < 0
    assert 'exceed' in _lib.last_error() and str(1 << 26) in _lib.last_error()


def test_flags_and_batch_gates_without_gpu_calls():
    """Unknown flag bits, and WT_FLAG_NO_MAC on the fused advection step (whose
    field the next advection reads), are rejected before any device call; the
    encoded batch table opens every tunnel's gate. The fake buffer addresses are
    never dereferenced."""
    from windtunnel import _lib
    lib = _lib.load()
    g = _lib.Grid(16, 8, 8, 0, 1, 0, (ctypes.c_void_p * 2)(0x1000, 0x2000), 0x3000, 0x4000,
                  (ctypes.c_void_p * 2)(0x5000, 0x6000), 0x7000, 0x8000)
    p = _lib.Params(0.05, 0.52, 0.16)
    assert lib.wt_step(ctypes.byref(g), 0, 1, ctypes.byref(p), 8, None) < 0
    assert 'flag' in _lib.last_error()
    box = _lib.SeedBox(2.0, 4.0, 4.0, 2.0, 2.0)
    v = ctypes.c_void_p
    assert lib.wt_step_advect(ctypes.byref(g), 0, ctypes.byref(p), _lib.WT_FLAG_NO_MAC, v(0x9000), v(0xa000),
                              v(0xb000), None, 16, 1.0, 0.05, 100.0, 1, ctypes.byref(box), None) < 0
    assert 'WT_FLAG_NO_MAC' in _lib.last_error()
    nt = 3
    grids = (_lib.Grid * nt)(g, g, g)
    prm = (_lib.Params * nt)(p, p, p)
    cur = (ctypes.c_int32 * nt)(0, 1, 0)
    nbytes, off = lib.wt_batch_workspace_bytes(nt), lib.wt_batch_gate_offset(nt)
    # descriptors, gate bytes padded to 8, then the persistent kernel's 8 claim
    # counters on 128-byte lines from the next multiple of 128
    assert off > 0 and nbytes == -(-(off + 8) // 128) * 128 + 8 * 128
    img = (ctypes.c_uint8 * nbytes)(*([0xa5] * nbytes))
    info = _lib.BatchInfo()
    assert lib.wt_batch_encode(grids, prm, cur, nt, None, img, ctypes.byref(info)) == 0
    assert list(img[off:off + 8]) == [1, 1, 1, 0, 0, 0, 0, 0]
    assert not any(img[off + 8:nbytes]), 'the claim counters must start at zero'
    assert info.tile_order == _lib.WT_BATCH_BACKWARDS and info.persist_blocks == 0
    assert info.ntunnels == nt and info.total_tiles == nt * (16 * 8 * 8 // 256)
    # the batch's largest cells-per-thread picks the kernel variant; its tiles follow each shape
    assert info.max_cpt == 1
    g2, g4 = _lib.Grid.from_buffer_copy(g), _lib.Grid.from_buffer_copy(g)
    g2.cpt, g4.cpt = 2, 4
    for mixed, want in (((g, g2, g), 2), ((g4, g, g2), 4)):
        grids = (_lib.Grid * nt)(*mixed)
        assert lib.wt_batch_encode(grids, prm, cur, nt, None, img, ctypes.byref(info)) == 0
        assert info.max_cpt == want
        assert info.total_tiles == sum(-(-16 * 8 * 8 // (256 * m.cpt)) for m in mixed)
    assert lib.wt_batch_step(ctypes.c_void_p(0x9004), ctypes.byref(info), 0, 0, None) < 0
    assert 'aligned' in _lib.last_error()
    info.max_cpt = 3
    assert lib.wt_batch_step(ctypes.c_void_p(0x9000), ctypes.byref(info), 0, 0, None) < 0
    assert 'cells per thread' in _lib.last_error()
    info.max_cpt = 1
    info.tile_order = 8
    assert lib.wt_batch_step(ctypes.c_void_p(0x9000), ctypes.byref(info), 0, 0, None) < 0
    assert 'tile order' in _lib.last_error()
    info.tile_order = _lib.WT_BATCH_PERSISTENT
    info.persist_blocks = -1
    assert lib.wt_batch_step(ctypes.c_void_p(0x9000), ctypes.byref(info), 0, 0, None) < 0
    assert 'persist_blocks' in _lib.last_error()


@pytest.mark.parametrize('dims,form,policy', [
    ((128, 64, 64), 0, 2),      # 2 x 19 x 4 B x 0.52M = 80 MB: all cached
    ((160, 96, 96), 0, 2),      # 213.75 MiB, within the 224 MiB bound
    ((200, 96, 96), 0, 0),      # 280 MB
    ((256, 112, 136), 1, 0),    # FP16, 296 MB
    ((256, 128, 128), 1, 0),    # FP16, 319 MB
    ((384, 192, 192), 0, 1),    # 2.15 GB >= 2 GiB: streaming
    ((512, 256, 256), 1, 1),    # FP16 33.5M cells: 2.55 GB, above 2 GiB
    ((512, 256, 256), 0, 1),
])
def test_cache_policy_boundaries(dims, form, policy):
    """wt_cache_policy: 2 within 224 MiB of populations (both buffers), 1 from
    2 GiB, else 0; computed from the element size of the storage form."""
    from windtunnel import _lib
    lib = _lib.load()
    n = dims[0] * dims[1] * dims[2]
    b = 2 * 19 * n * {0: 4, 1: 2, 2: 2, 3: 8}[form]
    expect = 1 if b >= 1 << 31 else (2 if b <= 224 << 20 else 0)
    assert lib.wt_cache_policy(*dims, form) == expect == policy


@pytest.mark.parametrize('form,elem', [(0, 4), (1, 2), (2, 2), (3, 8)])
def test_buffer_sizes_and_footprint(form, elem):
    from windtunnel import _lib
    lib = _lib.load()
    nx, ny, nz = 512, 256, 256
    n = nx * ny * nz
    s = (ctypes.c_int64 * 6)()
    assert lib.wt_buffer_sizes(nx, ny, nz, form, s) == 0
    mac = 8 if form in (1, 2) else 16          # binary16 x 4 for FP16/BF16, float4 otherwise
    assert s[0] == 19 * n * elem and s[1] == n and s[2] == mac * n
    shell = 2 * (ny * nz + nx * nz + nx * ny) * 4 * (8 if form == 3 else 4)
    assert s[3] == shell
    total = lib.wt_state_bytes(nx, ny, nz, form)
    assert total == 2 * s[0] + s[1] + s[2] + 2 * s[3] + s[4] + s[5]
    # per cell: 2 x 19 populations + the mac entry + 1 B class, plus O(n^(2/3)) extras
    per_cell = 2 * 19 * elem + mac + 1
    assert per_cell * n <= total <= per_cell * n * 1.02
    if form == 0:
        assert per_cell == 169      # lbm.py:30 BYTES_PER_CELL


def _golden_cases():
    meta = json.load(open(os.path.join(GOLDEN, 'voxel_golden.json')))
    blob = open(os.path.join(GOLDEN, 'voxel_golden.bin'), 'rb').read()
    for c in meta['cases']:
        tris = np.frombuffer(blob[c['tris_offset']:c['tris_offset'] + c['tris_bytes']], np.float32)
        mask = np.frombuffer(zlib.decompress(blob[c['mask_offset']:c['mask_offset'] + c['mask_zbytes']]), np.uint8)
        yield c, tris, mask


@pytest.mark.parametrize('case', [c['name'] for c, _, _ in _golden_cases()])
def test_voxelize_matches_reference_bit_for_bit(case):
    """Masks and counts of web/js/voxelize.js, produced by the reference itself under
    Node (tests/golden/make_voxel_golden.js)."""
    from windtunnel import geometry
    for c, tris, mask in _golden_cases():
        if c['name'] != case:
            continue
        got, counts = geometry.voxelize(tris, c['dims'], c['ground'])
        assert got.shape == mask.shape
        bad = np.flatnonzero(got != mask)
        assert bad.size == 0, (case, bad[:10], got[bad[:10]], mask[bad[:10]])
        assert counts == {'count': c['count'], 'frontal': c['frontal'], 'planform': c['planform']}


def test_sphere_counts_closed_form():
    """Sphere D=16 at (40,32,32), cell centres with the <= r^2 rule: 2,109 cells and a
    frontal area of 197 = the Gauss circle count N(8) (lattice points with y^2+z^2 <= 64)."""
    from windtunnel import geometry
    dims = (128, 64, 64)
    mask, counts = geometry.sphere(dims, (40, 32, 32), 16)
    assert counts['count'] == 2109 and mask.sum() == 2109
    assert counts['frontal'] == sum(1 for y in range(-8, 9) for z in range(-8, 9) if y * y + z * z <= 64) == 197
    assert counts['planform'] == 197
    z, y, x = np.meshgrid(np.arange(64), np.arange(64), np.arange(128), indexing='ij')
    brute = ((x - 40) ** 2 + (y - 32) ** 2 + (z - 32) ** 2 <= 64).astype(np.uint8).reshape(-1)
    assert np.array_equal(mask, brute)


def test_ellipsoid_union_brute_force_and_closure():
    from windtunnel import geometry
    dims = (60, 30, 34)
    nx, ny, nz = dims
    rng = np.random.default_rng(9)
    ells = np.column_stack([rng.uniform(15, 45, 6), rng.uniform(5, 25, 6), rng.uniform(8, 26, 6),
                            rng.uniform(2, 9, (6, 3))])
    mask, counts = geometry.ellipsoids(ells, dims, ground=True)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing='ij')
    inside = np.zeros((nz, ny, nx), bool)
    for cx, cy, cz, ax, ay, az in ells:
        dx, dy, dz = (x - cx) / ax, (y - cy) / ay, (z - cz) / az
        inside |= dx * dx + dy * dy + dz * dz <= 1.0
    inside[[0, -1], :, :] = False
    inside[:, [0, -1], :] = False
    inside[:, :, [0, -1]] = False
    m3 = mask.reshape(nz, ny, nx)
    body = m3 == 1
    # every union cell is solid; extra solid cells are only enclosed pockets
    assert np.all(body[inside])
    from scipy import ndimage
    fluid = ~inside
    lab, _ = ndimage.label(fluid)
    border = set(np.unique(np.concatenate([lab[0].ravel(), lab[-1].ravel(), lab[:, 0].ravel(), lab[:, -1].ravel(),
                                            lab[:, :, 0].ravel(), lab[:, :, -1].ravel()])))
    enclosed = fluid & ~np.isin(lab, list(border))
    expect = inside | enclosed
    expect[:, 0, :] = False
    assert np.array_equal(body, expect)
    assert np.all(m3[:, 0, :] == 2)
    assert counts['count'] == int((inside | enclosed).sum())


def test_closure_fills_enclosed_pocket():
    """A hollow ellipsoid shell (union of 6 slabs) encloses fluid that the flood fill closes."""
    from windtunnel import geometry
    dims = (30, 30, 30)
    ells = []
    for ax in range(3):
        for s in (-1, 1):
            c = [15.0, 15.0, 15.0]
            c[ax] += s * 6
            a = [8.0, 8.0, 8.0]
            a[ax] = 1.2
            ells.append(c + a)
    mask, counts = geometry.ellipsoids(ells, dims)
    assert mask[15 + 30 * (15 + 30 * 15)] == 1     # centre was fluid, enclosed -> solid


@pytest.mark.parametrize('form', ['fp32', 'fp16', 'bf16', 'fp64'])
def test_autotune_classes_follow_the_library_cache_policy(form):
    """The tuner's size classes are cut where the library's cache policy changes:
    its Python rule equals wt_cache_policy on both sides of both boundaries."""
    from windtunnel import _lib, autotune
    lib = _lib.load()
    fid = _lib.FORMS[form]
    eb = autotune.ELEM_BYTES[form]
    for bound in (224 << 20, 1 << 31):
        n0 = bound // (2 * 19 * eb)              # the largest cell count at or below the bound
        a0 = n0 // 9                             # grids a x 3 x 3 on both sides of it
        pols = []
        for a in (a0 - 1, a0, a0 + 1, a0 + 2):
            pols.append(autotune.cache_policy(form, 9 * a))
            assert lib.wt_cache_policy(a, 3, 3, fid) == pols[-1], (form, a)
        assert pols[0] == pols[1] != pols[2] == pols[3], (form, bound, pols)
    for name, dims in autotune.CLASSES:
        n = dims[0] * dims[1] * dims[2]
        assert autotune.size_class(n, form) == name
        assert lib.wt_cache_policy(*dims, fid) == autotune.cache_policy(form, n)


def test_autotune_batch_table():
    """The committed tuning table holds, for batches, one cells-per-thread value
    per storage form, chosen from the timed uniform candidates, and one per size
    class found from it by coordinate descent (a class keeps a value only where
    its timing beat the uniform one); a tunnel in a batch takes its class's value.
    The tuning batch spans the serving mix's sizes and the 48-tunnel batch's."""
    import json
    from windtunnel import autotune
    with open(autotune.TUNING_PATH) as fh:
        table = json.load(fh)
    for form in ('fp32', 'fp16', 'bf16', 'fp64'):
        c = autotune.batch_cpt(form)
        assert c in autotune.BATCH_CPTS and c == table['batch_cpt'][form]
        ms = table['batch_timings'][form]
        uniform = {k: v for k, v in ms.items() if k in {str(x) for x in autotune.BATCH_CPTS}}
        assert min(uniform, key=uniform.get) == str(c)
        per_class = table['batch_cpt_class'][form]
        assert set(per_class) <= {name for name, _ in autotune.CLASSES}
        for name, dims in autotune.CLASSES:
            n = dims[0] * dims[1] * dims[2]
            assert autotune.size_class(n, form) == name
            assert autotune.batch_cpt(form, n) == per_class.get(name, c)
        if per_class != {k: c for k in per_class}:
            key = ','.join(f'{k}{v}' for k, v in sorted(per_class.items()))
            assert ms[key] < ms[str(c)]
    assert autotune.batch_cpt('no-such-form') == 1
    assert autotune.batch_cpt('no-such-form', 10 ** 6) == 1
    dims = autotune.batch_dims()
    assert len(dims) == autotune.BATCH_SERVING + autotune.BATCH_SMALL
    big = [d[0] * d[1] * d[2] for d in dims[:autotune.BATCH_SERVING]]
    assert all(d[0] == 2 * d[1] == 2 * d[2] and d[1] % 8 == 0 for d in dims[:autotune.BATCH_SERVING])
    assert min(big) > 0.3e6 and max(big) < 40e6
    assert all(d in autotune.BATCH_SMALL_DIMS for d in dims[autotune.BATCH_SERVING:])
