"""Front-end benchmark: K browser sessions against the multi-session WebSocket
server in this process, speaking the viewer's protocol (web/js/net.js,
server/app.py:104-148): grid, params, gzip solid upload, reset, respawn, then
frames acknowledged with `ready` as the viewer does (app.py:216-217).

Each client gets its own tunnel (resolution and storage form per client,
sphere body, 10 steps per tick, a slice every frame, particles). Reports the
server's aggregate MLUPS, frames per second per client, and the p50/p99
interval between a client's frames.

    python -m windtunnel.frontend_bench [--clients 8] [--res high] [--seconds 20] [--particles 20000]
"""
import argparse
import asyncio
import json
import sys
import time

import numpy as np


async def _client(url, res, storage, dims, mask_bytes, seconds, particles, stats):
    import aiohttp
    from . import wire
    async with aiohttp.ClientSession() as cs:
        async with cs.ws_connect(url, max_msg_size=1 << 30) as ws:
            await ws.receive_json()                                   # hello
            await ws.send_json({'type': 'grid', 'res': res, 'storage': storage})
            while True:
                m = await ws.receive()
                if m.type == aiohttp.WSMsgType.TEXT and json.loads(m.data)['type'] in ('grid', 'error'):
                    assert json.loads(m.data)['type'] == 'grid', m.data
                    break
            nx, ny, nz = dims
            await ws.send_json({'type': 'params', 'uIn': 0.05, 'tau': 0.52, 'smag': 0.16, 'steps': 10,
                                'particles': particles, 'maxAge': 2.2 * nx / 0.5, 'lref': nx / 6,
                                'seedCy': ny / 2, 'seedCz': nz / 2, 'seedHy': ny / 4, 'seedHz': nz / 4,
                                'slice': {'axis': 2, 'index': nz // 2, 'field': 0, 'vmin': 0.0, 'vmax': 1.5}})
            await ws.send_bytes(wire.command_bytes({'type': 'solid', 'dims': list(dims)}, mask_bytes))
            await ws.send_json({'type': 'reset'})
            await ws.send_json({'type': 'respawn'})
            t_end = time.monotonic() + seconds
            last = None
            gaps, steps, frames = [], 0, 0
            while time.monotonic() < t_end:
                m = await ws.receive(timeout=30)
                if m.type != aiohttp.WSMsgType.BINARY:
                    continue
                now = time.monotonic()
                h = wire.decode_frame(m.data)['header']
                await ws.send_json({'type': 'ready'})
                if h['numP'] != particles:                            # frames from before the set-up
                    continue
                if last is not None:
                    gaps.append(now - last)
                last = now
                steps += h['steps']
                frames += 1
            stats.append({'frames': frames, 'steps': steps, 'gaps': gaps, 'cells': nx * ny * nz})


async def _main(clients, res, seconds, particles, forms):
    import lbm
    from aiohttp import web
    from . import frontend, geometry
    dims = lbm.RESOLUTIONS[res]
    mask = geometry.sphere(dims, (dims[0] // 4, dims[1] // 2, dims[2] // 2), dims[1] // 4)[0]
    app = frontend.make_app()
    runner = web.AppRunner(app)
    await runner.setup()
    site = web.TCPSite(runner, '127.0.0.1', 0)
    await site.start()
    port = site._server.sockets[0].getsockname()[1]
    engine = app[frontend.ENGINE]
    stats = []
    try:
        t0 = time.monotonic()
        c0 = engine.cell_steps
        await asyncio.gather(*[_client(f'http://127.0.0.1:{port}/ws', res, forms[k % len(forms)], dims,
                                       mask.tobytes(), seconds, particles, stats) for k in range(clients)])
        wall = time.monotonic() - t0
        cells = engine.cell_steps - c0
    finally:
        await runner.cleanup()
    gaps = np.concatenate([np.array(s['gaps']) for s in stats if s['gaps']]) * 1e3
    return {'clients': clients, 'res': res, 'dims': list(dims), 'forms': list(forms), 'particles': particles,
            'seconds': seconds, 'server_mlups': cells / wall / 1e6,
            'frames_per_s_per_client': float(np.mean([s['frames'] for s in stats])) / seconds,
            'steps_per_s_per_client': float(np.mean([s['steps'] for s in stats])) / seconds,
            'frame_interval_ms': {'p50': float(np.percentile(gaps, 50)), 'p99': float(np.percentile(gaps, 99))},
            'engine_errors': len(engine.errors), 'batched_launches': engine.batched_launches}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--clients', type=int, default=8)
    ap.add_argument('--res', default='high')
    ap.add_argument('--seconds', type=float, default=20.0)
    ap.add_argument('--particles', type=int, default=20000)
    ap.add_argument('--forms', default='fp32,fp16')
    args = ap.parse_args(argv)
    out = asyncio.run(_main(args.clients, args.res, args.seconds, args.particles, args.forms.split(',')))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
