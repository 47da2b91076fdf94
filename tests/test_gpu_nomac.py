"""WT_FLAG_NO_MAC: steps that leave the macroscopic field unwritten.

The field is an output of the step only (face cells read the shell snapshot,
forces come from the populations), so with the flag every population, shell
and FP64 sum is the same bit for bit, the field keeps what the last storing
step wrote, and the next storing step writes exactly the field a run that
stored it every step has. In a batch, tunnels with their gate byte set still
store it. Reference anchor: Solver.step reads the field once, after its last
step (lbm.py:443-453)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

FORMS = ['fp32', 'fp16', 'bf16', 'fp64']
UIN, TAU, SMAG = 0.05, 0.52, 0.16


def _bytes(t):
    return np.ascontiguousarray(t.detach().cpu().numpy()).view(np.uint8).copy()


def _tunnel(dims, form):
    from windtunnel import geometry
    from windtunnel.tunnel import Tunnel
    nx, ny, nz = dims
    ells = [[nx * 0.35, ny * 0.45, nz * 0.5, nx * 0.12, ny * 0.25, nz * 0.2]]
    t = Tunnel(dims, form)
    t.init(UIN)
    t.set_solid(geometry.ellipsoids(ells, dims, ground=True)[0])
    return t


def _state(t):
    return [_bytes(t.raw_populations()), _bytes(t.shell[t.cur]), _bytes(t.sums)]


@pytest.mark.parametrize('form', FORMS)
def test_no_mac_steps_change_nothing_but_the_field(gpu, form):
    from windtunnel._lib import WT_FLAG_FORCES, WT_FLAG_MASS, WT_FLAG_NO_MAC
    dims = (58, 22, 20)
    a, b = _tunnel(dims, form), _tunnel(dims, form)
    red = WT_FLAG_FORCES | WT_FLAG_MASS
    a.step(1, UIN, TAU, SMAG)
    b.step(1, UIN, TAU, SMAG)
    held = _bytes(b.mac)
    a.step(3, UIN, TAU, SMAG, flags=red)
    b.step(3, UIN, TAU, SMAG, flags=red | WT_FLAG_NO_MAC)
    torch.cuda.synchronize()
    assert a.cur == b.cur
    for x, y in zip(_state(a), _state(b)):
        assert np.array_equal(x, y), form
    assert np.array_equal(_bytes(b.mac), held), form            # untouched
    assert not np.array_equal(_bytes(a.mac), held), form        # the stored field moved on
    a.step(1, UIN, TAU, SMAG, flags=red)
    b.step(1, UIN, TAU, SMAG, flags=red)
    for x, y in zip(_state(a) + [_bytes(a.mac)], _state(b) + [_bytes(b.mac)]):
        assert np.array_equal(x, y), form


@pytest.mark.parametrize('form', ['fp32', 'fp16'])
def test_batch_gates_choose_who_stores_the_field(gpu, form):
    from windtunnel._lib import WT_FLAG_FORCES, WT_FLAG_NO_MAC
    from windtunnel.batch import Batch
    sizes = [(32, 16, 16), (48, 24, 24), (40, 20, 22), (56, 28, 24), (64, 32, 32)]
    params = [(UIN, TAU, SMAG)] * len(sizes)
    alone = [_tunnel(d, form) for d in sizes]
    mixed = [_tunnel(d, form) for d in sizes]
    held = [_bytes(t.mac) for t in mixed]
    b = Batch(mixed, params)
    gates = [True, False, True, False, False]
    b.set_gates(gates)
    for _ in range(3):
        b.step(WT_FLAG_FORCES | WT_FLAG_NO_MAC)
        for t in alone:
            t.step(1, UIN, TAU, SMAG, flags=WT_FLAG_FORCES)
    torch.cuda.synchronize()
    sums = b.last_sums().cpu().numpy()
    for k, (ta, tm) in enumerate(zip(alone, mixed)):
        assert np.array_equal(_bytes(ta.raw_populations()), _bytes(tm.raw_populations())), k
        assert np.array_equal(ta.sums.cpu().numpy().view(np.uint64), sums[k].view(np.uint64)), k
        want = _bytes(ta.mac) if gates[k] else held[k]
        assert np.array_equal(_bytes(tm.mac), want), k
    # without the flag every tunnel stores it again, gates or not
    b.step(WT_FLAG_FORCES)
    for t in alone:
        t.step(1, UIN, TAU, SMAG, flags=WT_FLAG_FORCES)
    torch.cuda.synchronize()
    for k, (ta, tm) in enumerate(zip(alone, mixed)):
        assert np.array_equal(_bytes(ta.mac), _bytes(tm.mac)), k


def test_fused_advection_step_rejects_no_mac(gpu):
    from windtunnel._lib import WT_FLAG_NO_MAC
    t = _tunnel((40, 16, 16), 'fp32')
    n = 64
    pos = torch.zeros((n, 4), dtype=torch.float32, device='cuda')
    pos[:, 0] = 5.0
    pos[:, 1] = 8.0
    pos[:, 2] = 8.0
    spd = torch.ones(n, dtype=torch.float32, device='cuda')
    ids = torch.arange(n, dtype=torch.int32, device='cuda')
    prev = t.mac.clone()
    with pytest.raises(RuntimeError, match='WT_FLAG_NO_MAC'):
        t.step_advect(UIN, TAU, SMAG, WT_FLAG_NO_MAC, prev, pos, spd, ids, 1.0, 100.0, 1, (2.0, 8.0, 8.0, 4.0, 4.0))


def test_particle_free_session_slices_equal_a_storing_run(gpu):
    """A session without particles stores the field only on its frame steps; its
    slice images equal those of a tunnel that stores it every step."""
    from windtunnel.session import FIELD_RANGES, Session
    from windtunnel import geometry
    dims = (64, 24, 24)
    nx, ny, nz = dims
    ells = np.array([[nx * 0.35, ny * 0.45, nz * 0.5, nx * 0.12, ny * 0.25, nz * 0.2]])
    mask, counts = geometry.ellipsoids(ells, dims, ground=True)
    s = Session(dims, mask, 'fp32', UIN, TAU, SMAG, particles=0, slices=[(2, nz // 2, 2)], frame_every=5,
                lref=20.0, seed=7, frontal=counts['frontal'])
    t = _tunnel(dims, 'fp32')               # the same body
    lo, hi = FIELD_RANGES[2]
    want = torch.zeros(2 * nx * ny, dtype=torch.uint8, device='cuda')
    for k in range(10):
        s.step()
        t.step(1, UIN, TAU, SMAG)
    t.slice(want, 2, nz // 2, 2, UIN, 20.0, lo, hi)
    s.flush()
    torch.cuda.synchronize()
    assert np.array_equal(s.slice_host[0].numpy(), want.cpu().numpy())
