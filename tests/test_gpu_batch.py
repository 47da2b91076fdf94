"""Batched tunnels: every tunnel advanced inside one batched launch must equal
the same tunnel advanced alone, bit for bit (populations, macroscopic field and
the FP64 force/mass sums), also when replayed from a captured graph; for both
batched kernels: one workgroup per item, and the persistent one (PERSIST: run at
the device's full occupancy, -1, or with a few workgroups striding over many
items each)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SIZES = [(32, 16, 16), (48, 24, 24), (40, 20, 22), (56, 28, 24), (64, 32, 32)]
PARAMS = [(0.05, 0.52, 0.16), (0.08, 0.60, 0.17), (0.03, 0.51, 0.16), (0.06, 0.55, 0.10), (0.04, 0.53, 0.2)]


def _mask(dims, seed):
    from windtunnel import geometry
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    if seed % 2:
        c = (nx * 0.3, ny * 0.5, nz * 0.5)
        return geometry.sphere(dims, c, min(ny, nz) * 0.4)[0]
    ells = [[nx * 0.35, ny * rng.uniform(0.3, 0.6), nz * 0.5, nx * 0.12, ny * 0.2, nz * 0.2]]
    return geometry.ellipsoids(ells, dims, ground=bool(seed % 4 == 0))[0]


def _make(form, pre_steps=()):
    from windtunnel.tunnel import Tunnel
    ts = []
    for k, (dims, p) in enumerate(zip(SIZES, PARAMS)):
        t = Tunnel(dims, form)
        t.init(p[0])
        t.set_solid(_mask(dims, k))
        if k in pre_steps:
            t.step(1, *p)
        ts.append(t)
    return ts


def _state(t):
    return (t.raw_populations().cpu().numpy().copy(), t.mac.cpu().numpy().copy(), t.sums.cpu().numpy().copy())


# batched kernel of a test case: 0 one workgroup per item, -1 persistent at full
# occupancy, k > 0 persistent with k workgroups
def _batch(tunnels, params, persist=0, **kw):
    from windtunnel.batch import Batch
    return Batch(tunnels, params, persistent=persist != 0, persist_blocks=max(persist, 0), **kw)


def _same(a, b):
    return all(np.array_equal(np.ascontiguousarray(x).view(np.uint8), np.ascontiguousarray(y).view(np.uint8))
               for x, y in zip(a, b))


@pytest.mark.parametrize('persist', [0, -1, 7])
@pytest.mark.parametrize('form', ['fp32', 'fp16', 'bf16', 'fp64'])
def test_batch_equals_alone(gpu, form, persist):
    from windtunnel._lib import WT_FLAG_FORCES, WT_FLAG_MASS
    alone = _make(form, pre_steps=(2,))
    mixed = _make(form, pre_steps=(2,))
    b = _batch(mixed, PARAMS, persist)
    flags = WT_FLAG_FORCES | WT_FLAG_MASS
    for k in range(9):
        b.step(flags)
        for t, p in zip(alone, PARAMS):
            t.step(1, *p, flags=flags)
        torch.cuda.synchronize()
        bs = b.last_sums().cpu().numpy()
        for j, (ta, tm) in enumerate(zip(alone, mixed)):
            assert ta.cur == tm.cur
            assert _same(_state(ta)[:2], _state(tm)[:2]), (form, k, ta.nx)
            assert _same([ta.sums.cpu().numpy()], [bs[j]]), (form, k, ta.nx)


@pytest.mark.parametrize('persist', [0, 5])
def test_batch_graph_replay_equals_eager(gpu, persist):
    from windtunnel._lib import WT_FLAG_FORCES
    eager = _make('fp32')
    graphed = _make('fp32')
    be = _batch(eager, PARAMS)
    bg = _batch(graphed, PARAMS, persist)
    bg.capture(WT_FLAG_FORCES)
    for rep in range(5):
        be.step(WT_FLAG_FORCES)
        s0 = be.last_sums().clone()
        be.step(WT_FLAG_FORCES)
        bg.replay()
        torch.cuda.synchronize()
        # eager host work between replays
        _ = torch.arange(1000, device='cuda').sum().item()
        for ta, tb in zip(eager, graphed):
            assert _same(_state(ta)[:2], _state(tb)[:2]), rep
        assert _same([s0.cpu().numpy(), be.last_sums().cpu().numpy()], [bg.sums[0].cpu().numpy(), bg.sums[1].cpu().numpy()])


def test_batch_rejects_mixed_forms(gpu):
    from windtunnel.batch import Batch
    from windtunnel.tunnel import Tunnel
    with pytest.raises(ValueError):
        Batch([Tunnel((8, 8, 8), 'fp32'), Tunnel((8, 8, 8), 'fp16')], [(0.05, 0.52, 0.16)] * 2)


@pytest.mark.parametrize('persist', [0, -1])
def test_batch_with_a_large_tunnel_equals_alone(gpu, persist):
    """A 2.4M-cell tunnel (10 first-level reduction chunks) batched with three
    small ones (1 chunk each): every tunnel's force/mass sums equal the same
    tunnel stepped alone, bit for bit. Its 9.4k items outnumber the persistent
    kernel's workgroups at full occupancy."""
    from windtunnel import geometry
    from windtunnel._lib import WT_FLAG_FORCES, WT_FLAG_MASS
    from windtunnel.tunnel import Tunnel
    big_dims, big_p = (192, 112, 112), (0.05, 0.52, 0.16)
    assert big_dims[0] * big_dims[1] * big_dims[2] > 8 * 1024 * 256
    big_mask = geometry.sphere(big_dims, (60, 56, 56), 20)[0]

    def big():
        t = Tunnel(big_dims, 'fp32')
        t.init(big_p[0])
        t.set_solid(big_mask)
        return t
    alone = _make('fp32')[:3] + [big()]
    mixed = _make('fp32')[:3] + [big()]
    params = PARAMS[:3] + [big_p]
    b = _batch(mixed, params, persist)
    flags = WT_FLAG_FORCES | WT_FLAG_MASS
    for k in range(4):
        b.step(flags)
        for t, p in zip(alone, params):
            t.step(1, *p, flags=flags)
        torch.cuda.synchronize()
        bs = b.last_sums().cpu().numpy()
        for j, t in enumerate(alone):
            assert _same([t.sums.cpu().numpy()], [bs[j]]), (k, t.nx)


@pytest.mark.parametrize('cpt', [None, 2, 4])
def test_batch_runs_every_tunnel_at_the_batch_cells_per_thread(gpu, cpt):
    """By default a batch runs every tunnel at the cells per thread tuned for
    tunnels of its size class in batches of its storage form (or at an explicit
    one), whatever the tunnels' own launch shapes, which it leaves alone; results
    equal the tunnels alone."""
    from windtunnel import autotune
    from windtunnel._lib import WT_FLAG_FORCES
    from windtunnel.batch import Batch
    alone, mixed = _make('fp16'), _make('fp16')
    for t in mixed:
        t.shape = (2, 0)
    b = Batch(mixed, PARAMS, cpt=cpt)
    want = [cpt or autotune.batch_cpt('fp16', t.n) for t in mixed]
    assert b.info.max_cpt == max(want)
    assert b.info.total_tiles == sum(-(-t.n // (256 * w)) for t, w in zip(mixed, want))
    assert all(t.shape == (2, 0) for t in mixed)
    for _ in range(2):
        b.step(WT_FLAG_FORCES)
        for t, p in zip(alone, PARAMS):
            t.step(1, *p, flags=WT_FLAG_FORCES)
    torch.cuda.synchronize()
    for ta, tm in zip(alone, mixed):
        assert _same(_state(ta)[:2], _state(tm)[:2])


@pytest.mark.parametrize('persist', [0, 3])
def test_batch_mixed_launch_shapes_and_sweep_orders(gpu, persist):
    """One launch per step with each tunnel's own cells per thread (and 64-bit
    addressing for one of them), sweeping forwards or alternating: every tunnel
    gets the bits it gets alone."""
    from windtunnel._lib import WT_BATCH_PERSISTENT, WT_FLAG_FORCES, WT_FLAG_MASS
    shapes = [(1, 0), (2, 0), (4, 2), (2, 4), (1, 2)]
    alone = _make('fp32', pre_steps=(1,))
    mixed = _make('fp32', pre_steps=(1,))
    for a, m, sh in zip(alone, mixed, shapes):
        a.shape = (1, 0)
        m.shape = sh
    b = _batch(mixed, PARAMS, persist, own_shapes=True)
    assert b.info.total_tiles == sum(-(-t.n // (256 * t.shape[0])) for t in mixed)
    assert b.info.max_cpt == 4
    for order in (2, 0):
        b.info.tile_order = order | (WT_BATCH_PERSISTENT if persist else 0)
        flags = WT_FLAG_FORCES | WT_FLAG_MASS
        for k in range(3):
            b.step(flags)
            for t, p in zip(alone, PARAMS):
                t.step(1, *p, flags=flags)
            torch.cuda.synchronize()
            bs = b.last_sums().cpu().numpy()
            for j, (ta, tm) in enumerate(zip(alone, mixed)):
                assert _same(_state(ta)[:2], _state(tm)[:2]), (order, k, j)
                assert _same([ta.sums.cpu().numpy()], [bs[j]]), (order, k, j)


@pytest.mark.parametrize('persist', [0, 11])
@pytest.mark.parametrize('form', ['fp16', 'fp32'])
def test_batch_of_more_than_64_tunnels_equals_alone(gpu, form, persist):
    """70 tunnels of four sizes in one batch: the FP16/BF16 kernel finds a work
    item's tunnel by wave-wide votes over 64 descriptors at a time, so this batch
    needs two of them (FP32 searches); every tunnel ends bit-identical to itself
    run alone, and its FP64 sums equal the alone sums."""
    from windtunnel._lib import WT_FLAG_FORCES, WT_FLAG_MASS
    from windtunnel.batch import Batch
    from windtunnel.tunnel import Tunnel
    sizes = [(16, 8, 8), (24, 12, 10), (40, 8, 12), (8, 8, 8)]
    rng = np.random.default_rng(70)
    alone, mixed, params = [], [], []
    for k in range(70):
        dims = sizes[k % len(sizes)]
        p = (float(rng.uniform(0.03, 0.08)), float(rng.uniform(0.51, 0.6)), 0.16)
        mask = _mask(dims, k) if dims[1] >= 12 else np.zeros(dims[0] * dims[1] * dims[2], np.uint8)
        pair = []
        for _ in range(2):
            t = Tunnel(dims, form)
            t.init(p[0])
            t.set_solid(mask)
            pair.append(t)
        alone.append(pair[0])
        mixed.append(pair[1])
        params.append(p)
    b = _batch(mixed, params, persist)
    assert b.info.ntunnels == 70
    flags = WT_FLAG_FORCES | WT_FLAG_MASS
    for _ in range(3):
        b.step(flags)
        for t, p in zip(alone, params):
            t.step(1, *p, flags=flags)
    torch.cuda.synchronize()
    bs = b.last_sums().cpu().numpy()
    for j, (ta, tm) in enumerate(zip(alone, mixed)):
        assert _same(_state(ta)[:2], _state(tm)[:2]), j
        assert _same([ta.sums.cpu().numpy()], [bs[j]]), j
