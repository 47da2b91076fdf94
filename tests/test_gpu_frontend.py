"""Multi-session WebSocket front end on the GPU: three browser sessions in one
solver thread (two FP32 'test' tunnels stepped by one batched launch per
lattice step, one FP16 'low' tunnel alone) driven over real WebSockets with
the viewer's commands. After pausing, each session's populations, macroscopic
field and FP64 force equal those of the same tunnel run alone through
lbm.Solver for the same number of steps, bit for bit."""
import asyncio
import json
import time

import numpy as np
import pytest

aiohttp = pytest.importorskip('aiohttp')

pytestmark = pytest.mark.gpu

SESSIONS = [
    # res, storage, sphere centre, diameter, lattice parameters
    ('test', 'fp32', (20, 14, 18), 8, dict(uIn=0.05, tau=0.56, smag=0.16)),
    ('test', 'fp32', (24, 12, 17), 10, dict(uIn=0.06, tau=0.6, smag=0.12)),
    ('low', 'fp16', (40, 28, 36), 16, dict(uIn=0.05, tau=0.53, smag=0.17)),
]


async def recv(ws, timeout=20):
    msg = await ws.receive(timeout=timeout)
    if msg.type == aiohttp.WSMsgType.TEXT:
        return json.loads(msg.data)
    if msg.type == aiohttp.WSMsgType.BINARY:
        from windtunnel import wire
        return wire.decode_frame(msg.data)
    raise AssertionError(f'socket ended: {msg.type}')


async def until(ws, pred, timeout=20):
    t0 = time.monotonic()
    while time.monotonic() - t0 < timeout:
        m = await recv(ws, timeout)
        if 'header' in m:
            await ws.send_json({'type': 'ready'})
        if pred(m):
            return m
    raise AssertionError('timed out waiting for a message')


def test_sessions_in_mix_equal_alone(gpu):
    import lbm
    from aiohttp import web
    from windtunnel import frontend, geometry, wire

    masks = []
    for res, _, c, d, _ in SESSIONS:
        masks.append(geometry.sphere(lbm.RESOLUTIONS[res], c, d)[0])
    snaps = {}

    async def body(url, engine):
        async with aiohttp.ClientSession() as cs:
            socks = [await cs.ws_connect(url) for _ in SESSIONS]
            for ws in socks:
                h = await recv(ws)
                assert h['type'] == 'hello' and h['cuda'] is True
            for ws, (res, storage, _, _, _) in zip(socks, SESSIONS):
                await ws.send_json({'type': 'grid', 'res': res, 'storage': storage})
                g = await until(ws, lambda m: m.get('type') in ('grid', 'error'))
                assert g['type'] == 'grid' and g['storage'] == storage, g
            for ws, (res, _, _, _, prm), mask in zip(socks, SESSIONS, masks):
                dims = list(lbm.RESOLUTIONS[res])
                await ws.send_json(dict({'type': 'params', 'running': False, 'steps': 10, 'particles': 2000,
                                         'maxAge': 500.0, 'lref': 8.0,
                                         'slice': {'axis': 2, 'index': dims[2] // 2, 'field': 0,
                                                   'vmin': 0.0, 'vmax': 1.5}}, **prm))
                await ws.send_bytes(wire.command_bytes({'type': 'solid', 'dims': dims}, mask.tobytes()))
                await ws.send_json({'type': 'reset'})
                await ws.send_json({'type': 'respawn'})
                await ws.send_json({'type': 'params', 'running': True})
            # run about a second of frames on every session
            t0 = time.monotonic()
            nframes = [0] * len(socks)
            while time.monotonic() - t0 < 1.5:
                for k, ws in enumerate(socks):
                    m = await until(ws, lambda m: 'header' in m)
                    h = m['header']
                    dims = list(lbm.RESOLUTIONS[SESSIONS[k][0]])
                    assert h['dims'] == dims
                    # frames still in flight from before the set-up commands have no particles
                    if h['running'] and h['numP'] == 2000:
                        nframes[k] += 1
                        assert m['parts16'].size == 8000 and m['parts8'].size == 2000
                        assert m['slice'].size == 2 * dims[0] * dims[1]
                        assert all(np.isfinite(h['forces']))
            assert min(nframes) >= 5, nframes
            for ws in socks:
                await ws.send_json({'type': 'params', 'running': False})
            for ws in socks:
                await until(ws, lambda m: 'header' in m and not m['header']['running'])
                await until(ws, lambda m: 'header' in m)
            sess = sorted(engine.sessions.values(), key=lambda s: s.sid)
            assert len(sess) == len(SESSIONS)
            for k, s in enumerate(sess):
                T = s.solver.T
                snaps[k] = dict(steps=s.lattice_steps, f=T.raw_populations().cpu().clone(),
                                mac=T.mac.cpu().clone(), sums=T.sums.cpu().clone(), cur=T.cur)
            assert engine.batched_launches > 0
            assert not engine.errors, engine.errors
            for ws in socks:
                await ws.close()

    async def main():
        app = frontend.make_app()
        runner = web.AppRunner(app)
        await runner.setup()
        site = web.TCPSite(runner, '127.0.0.1', 0)
        await site.start()
        port = site._server.sockets[0].getsockname()[1]
        try:
            await asyncio.wait_for(body(f'http://127.0.0.1:{port}/ws', app[frontend.ENGINE]), 90)
        finally:
            await runner.cleanup()

    asyncio.run(main())

    for k, (res, storage, _, _, prm) in enumerate(SESSIONS):
        snap = snaps[k]
        assert snap['steps'] >= 50, snap['steps']
        alone = lbm.Solver(lbm.RESOLUTIONS[res], storage=storage)
        alone.set_solid(masks[k])
        alone.reset(prm['uIn'])
        alone.step(snap['steps'], dict(prm, maxAge=500.0, seedX=2.0, seedCy=20.0, seedCz=20.0,
                                       seedHy=10.0, seedHz=10.0))
        out = alone.frame({}, False)
        T = alone.T
        assert T.cur == snap['cur']
        assert torch_equal(T.raw_populations().cpu(), snap['f']), f'session {k}: populations differ'
        assert torch_equal(T.mac.cpu(), snap['mac']), f'session {k}: mac differs'
        assert out['forces'] == snap['sums'][:3].tolist(), f'session {k}: forces differ'


def torch_equal(a, b):
    import torch
    return bool(torch.equal(a.view(torch.uint8) if a.dtype != torch.uint8 else a,
                            b.view(torch.uint8) if b.dtype != torch.uint8 else b))
