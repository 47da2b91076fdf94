"""Multi-session WebSocket front end, protocol and admission, with a stand-in
solver (no GPU): hello/grid/params/reset/respawn/ready handling of
server/app.py:78-148, 205-227 per connection, binary solid/surface commands,
frame layout as web/js/net.js:18-29 decodes it, and memory admission."""
import asyncio
import json

import numpy as np
import pytest

aiohttp = pytest.importorskip('aiohttp')

from windtunnel import frontend, wire  # noqa: E402

RES = {'test': (16, 8, 8), 'low': (32, 16, 16), 'high': (64, 32, 32)}


def fake_footprint(dims, storage='fp32'):
    per = {'fp32': 169, 'fp16': 93, 'bf16': 93, 'fp64': 321}[storage]
    return dims[0] * dims[1] * dims[2] * per


class FakeSolver:
    """Records what the engine asks of it; force = (total steps, uIn, number of solid cells)."""

    def __init__(self, dims, device=None, storage='fp32'):
        self.nx, self.ny, self.nz = dims
        self.n = self.nx * self.ny * self.nz
        self.storage = storage
        self.num_p = 0
        self.num_verts = 0
        self.total = 0
        self.uin = None
        self.solid = np.zeros(self.n, np.uint8)
        self.resets = 0
        self.verts = None

    def reset(self, uin):
        self.uin = uin
        self.total = 0
        self.resets += 1

    def set_solid(self, arr):
        assert arr.size == self.n
        self.solid = np.array(arr, dtype=np.uint8)

    def set_surface(self, verts, norms):
        self.verts, self.norms = np.array(verts), np.array(norms)
        self.num_verts = len(self.verts)

    def set_particles(self, count, p):
        self.num_p = int(count)

    def step(self, steps, p):
        self.total += int(steps)

    def slice_dims(self, axis):
        return {2: (self.nx, self.ny), 1: (self.nx, self.nz), 0: (self.nz, self.ny)}[axis]

    def frame(self, p, want_cp):
        out = {}
        if self.num_p:
            out['parts16'] = np.full(self.num_p * 4, 7, np.uint16)
            out['parts8'] = np.full(self.num_p, 3, np.uint8)
        sl = p.get('slice')
        if sl and sl.get('field', -1) >= 0:
            w, h = self.slice_dims(int(sl['axis']))
            out['slice'] = np.full(2 * w * h, 9, np.uint8)
            out['sliceW'], out['sliceH'], out['sliceAxis'] = w, h, int(sl['axis'])
        if want_cp and self.num_verts:
            out['cp'] = np.full(self.num_verts, -5, np.int8)
        out['forces'] = [float(self.total), float(self.uin), float(int((self.solid != 0).sum()))]
        return out


def run_with_server(coro_fn, memory=10 ** 7, frame_hz=200, paused_hz=50):
    """Start the app on 127.0.0.1 with the fake solver, run coro_fn(url, engine)."""
    from aiohttp import web

    async def main():
        app = frontend.make_app(solver_factory=FakeSolver, footprint=fake_footprint, resolutions=RES,
                                memory=memory, device_name='fake', frame_hz=frame_hz, paused_hz=paused_hz)
        runner = web.AppRunner(app)
        await runner.setup()
        site = web.TCPSite(runner, '127.0.0.1', 0)
        await site.start()
        port = site._server.sockets[0].getsockname()[1]
        try:
            return await asyncio.wait_for(coro_fn(f'http://127.0.0.1:{port}/ws', app[frontend.ENGINE]), 60)
        finally:
            await runner.cleanup()
    return asyncio.run(main())


async def recv_json(ws, kind=None, timeout=10):
    while True:
        msg = await ws.receive(timeout=timeout)
        if msg.type == aiohttp.WSMsgType.TEXT:
            d = json.loads(msg.data)
            if kind is None or d['type'] == kind:
                return d
        elif msg.type != aiohttp.WSMsgType.BINARY:
            raise AssertionError(f'socket ended: {msg.type}')


async def recv_frame(ws, timeout=10, ack=True):
    while True:
        msg = await ws.receive(timeout=timeout)
        if msg.type == aiohttp.WSMsgType.BINARY:
            d = wire.decode_frame(msg.data)
            if ack:
                await ws.send_json({'type': 'ready'})
            return d
        if msg.type != aiohttp.WSMsgType.TEXT:
            raise AssertionError(f'socket ended: {msg.type}')


def test_hello_grid_and_frames():
    async def body(url, engine):
        async with aiohttp.ClientSession() as cs:
            async with cs.ws_connect(url) as ws:
                h = await recv_json(ws, 'hello')
                assert h['gpu'] == 'fake' and h['memory'] == 10 ** 7 and h['res'] is None
                assert [r['name'] for r in h['resolutions']] == list(RES)
                # 85 % of 1e7 B: 'high' (65,536 cells x 169 B = 11.1 MB) does not fit
                assert [r['fits'] for r in h['resolutions']] == [True, True, False]
                assert h['default'] == 'test'
                await ws.send_json({'type': 'grid', 'res': 'high'})
                e = await recv_json(ws, 'error')
                assert e['message'] == 'Resolution high does not fit on this device'
                await ws.send_json({'type': 'grid', 'res': 'low'})
                g = await recv_json(ws, 'grid')
                assert g == {'type': 'grid', 'res': 'low', 'dims': [32, 16, 16], 'storage': 'fp32'}
                await ws.send_json({'type': 'params', 'uIn': 0.05, 'steps': 4,
                                    'slice': {'axis': 2, 'index': 99, 'field': 0, 'vmin': 0, 'vmax': 1}})
                seen = 0
                for _ in range(6):
                    f = await recv_frame(ws)
                    hd = f['header']
                    assert hd['type'] == 'frame' and hd['dims'] == [32, 16, 16] and hd['running'] is True
                    seen += hd['steps']
                    if 'slice' in f:
                        assert (hd['sliceW'], hd['sliceH'], hd['sliceAxis']) == (32, 16, 2)
                        assert f['slice'].size == 2 * 32 * 16 and (f['slice'] == 9).all()
                # forces[0] counts the solver's steps: frames report every one of them
                await ws.send_json({'type': 'params', 'running': False})
                while True:
                    f = await recv_frame(ws)
                    seen += f['header']['steps']
                    if not f['header']['running']:
                        break
                assert f['header']['mlups'] == 0
                assert f['header']['forces'][0] == seen and seen > 0
                assert f['header']['forces'][1] == 0.1      # reset ran with the default uIn
    run_with_server(body)


def test_ready_flow_control():
    async def body(url, engine):
        async with aiohttp.ClientSession() as cs:
            async with cs.ws_connect(url) as ws:
                await recv_json(ws, 'hello')
                await ws.send_json({'type': 'grid', 'res': 'test'})
                await recv_json(ws, 'grid')
                await recv_frame(ws, ack=False)
                # no ready: nothing more may arrive
                with pytest.raises(asyncio.TimeoutError):
                    await ws.receive(timeout=0.4)
                await ws.send_json({'type': 'ready'})
                f = await recv_frame(ws, ack=False)
                assert f['header']['steps'] > 10      # the unsent steps accumulate
    run_with_server(body)


def test_sessions_are_independent():
    async def body(url, engine):
        async with aiohttp.ClientSession() as cs:
            a = await cs.ws_connect(url)
            b = await cs.ws_connect(url)
            await recv_json(a, 'hello')
            await recv_json(b, 'hello')
            await a.send_json({'type': 'grid', 'res': 'test'})
            await b.send_json({'type': 'grid', 'res': 'low'})
            await a.send_json({'type': 'params', 'uIn': 0.07})
            await recv_json(a, 'grid')
            await recv_json(b, 'grid')
            fa, fb = await recv_frame(a), await recv_frame(b)
            assert fa['header']['dims'] == [16, 8, 8] and fb['header']['dims'] == [32, 16, 16]
            # pause only b: a keeps stepping, b stops
            await b.send_json({'type': 'params', 'running': False})
            while (await recv_frame(b))['header']['running']:
                pass
            fb1 = (await recv_frame(b))['header']['forces'][0]
            fa1 = (await recv_frame(a))['header']['forces'][0]
            for _ in range(5):
                await recv_frame(a)
            fb2 = (await recv_frame(b))['header']['forces'][0]
            fa2 = (await recv_frame(a))['header']['forces'][0]
            assert fb2 == fb1 and fa2 > fa1
            assert len(engine.sessions) == 2
            await a.close()
            await b.close()
    run_with_server(body)


def test_binary_commands_and_particles():
    async def body(url, engine):
        async with aiohttp.ClientSession() as cs:
            async with cs.ws_connect(url) as ws:
                await recv_json(ws, 'hello')
                await ws.send_json({'type': 'grid', 'res': 'test'})
                await recv_json(ws, 'grid')
                n = 16 * 8 * 8
                mask = (np.arange(n) % 5 == 0).astype(np.uint8)
                # wrong dims and wrong size are dropped silently (app.py:139-143)
                await ws.send_bytes(wire.command_bytes({'type': 'solid', 'dims': [8, 8, 8]}, mask.tobytes()))
                await ws.send_bytes(wire.command_bytes({'type': 'solid', 'dims': [16, 8, 8]}, mask[:-1].tobytes()))
                await ws.send_bytes(wire.command_bytes({'type': 'solid', 'dims': [16, 8, 8]}, mask.tobytes()))
                verts = np.arange(30, dtype=np.float32).reshape(10, 3)
                norms = -verts
                await ws.send_bytes(wire.command_bytes({'type': 'surface', 'numVerts': 10},
                                                       np.concatenate([verts.ravel(), norms.ravel()]).tobytes()))
                await ws.send_json({'type': 'params', 'particles': 33, 'cp': True})
                await ws.send_json({'type': 'respawn'})
                got_cp = False
                for _ in range(12):
                    f = await recv_frame(ws)
                    if f['header']['numP'] == 33:
                        assert f['parts16'].size == 132 and (f['parts16'] == 7).all()
                        assert f['parts8'].size == 33
                        assert f['header']['forces'][2] == int(mask.sum())
                    if 'cp' in f:
                        assert f['header']['numVerts'] == 10 and (f['cp'] == -5).all()
                        got_cp = True
                assert got_cp
                sv = next(iter(engine.sessions.values())).solver
                assert np.array_equal(sv.solid, mask)
                assert np.array_equal(sv.verts, verts) and np.array_equal(sv.norms, norms)
                # particle count change respawns only once particles exist (app.py:129-132)
                await ws.send_json({'type': 'params', 'particles': 40})
                for _ in range(3):
                    f = await recv_frame(ws)
                assert f['header']['numP'] == 40
    run_with_server(body)


def test_admission_frees_on_disconnect():
    async def body(url, engine):
        async with aiohttp.ClientSession() as cs:
            a = await cs.ws_connect(url)
            await recv_json(a, 'hello')
            await a.send_json({'type': 'grid', 'res': 'low'})        # 8192 cells x 169 B
            await recv_json(a, 'grid')
            b = await cs.ws_connect(url)
            h = await recv_json(b, 'hello')
            fits = {r['name']: r['fits'] for r in h['resolutions']}
            assert fits == {'test': True, 'low': False, 'high': False}
            await b.send_json({'type': 'grid', 'res': 'low'})
            assert (await recv_json(b, 'error'))['message'] == 'Resolution low does not fit on this device'
            # fp16 storage needs 93 B per cell: 'low' fits beside a's fp32 'low'
            await b.send_json({'type': 'grid', 'res': 'low', 'storage': 'fp16'})
            g = await recv_json(b, 'grid')
            assert g['storage'] == 'fp16'
            await b.send_json({'type': 'grid', 'res': 'low', 'storage': 'fp99'})
            assert (await recv_json(b, 'error'))['message'] == 'Unknown storage fp99'
            await a.close()
            for _ in range(100):
                if engine.used == fake_footprint((32, 16, 16), 'fp16'):
                    break
                await asyncio.sleep(0.02)
            assert engine.used == fake_footprint((32, 16, 16), 'fp16')
            c = await cs.ws_connect(url)
            h = await recv_json(c, 'hello')
            assert {r['name']: r['fits'] for r in h['resolutions']}['low'] is True
            await b.close()
            await c.close()
    # budget 0.85 * 2.6e6 B = 2.21 MB: one fp32 'low' (1.38 MB) fits, two do not;
    # an fp16 'low' (0.76 MB) fits beside it
    run_with_server(body, memory=2_600_000)


def test_reset_restarts_steps():
    async def body(url, engine):
        async with aiohttp.ClientSession() as cs:
            async with cs.ws_connect(url) as ws:
                await recv_json(ws, 'hello')
                await ws.send_json({'type': 'grid', 'res': 'test'})
                await recv_json(ws, 'grid')
                for _ in range(3):
                    await recv_frame(ws)
                await ws.send_json({'type': 'params', 'running': False, 'uIn': 0.03})
                while (await recv_frame(ws))['header']['running']:
                    pass
                await ws.send_json({'type': 'reset'})
                for _ in range(3):
                    f = await recv_frame(ws)
                assert f['header']['forces'][:2] == [0.0, 0.03]
                sv = next(iter(engine.sessions.values())).solver
                assert sv.resets == 2
    run_with_server(body)
