"""Multi-session WebSocket front end: the reference viewer's protocol over many
resident tunnels.

The reference server (server/app.py:50-259) runs one solver thread (`Engine`)
that owns ONE tunnel shared by every connected browser. Here every WebSocket
connection is a session with its own tunnel, parameters, particles and frame
stream, and one solver thread advances all of them:

  * Protocol, per connection, as app.py:104-148 and 205-227: on connect the
    server sends `hello` (device, memory, resolutions with `fits`, default);
    text commands `grid` {res}, `params` {...}, `respawn`, `reset`, `ready`;
    binary commands [u32 header length][JSON header][gzip payload] of type
    `solid` (u8 mask, dropped unless dims and size match) and `surface`
    (f32 verts then normals). Frames are wire.frame() bytes, sent only to a
    session that has acknowledged its previous frame with `ready`.
    Extension: `grid` may carry `storage` ('fp32' default, 'fp16', 'bf16',
    'fp64').
  * Admission (app.py:72-76 generalised): a grid is granted when its exact
    footprint plus every other session's fits in 85 % of device memory;
    `hello.resolutions[].fits` is computed against what is left for this
    connection. A refused grid gets the reference's error message.
  * Stepping: every tick, running sessions with the same storage form and
    `steps` count advance together, one batched launch per lattice step
    (windtunnel.batch), forces fused into the last step; sessions alone in
    their group use Solver.step. Results are those of each session alone.
  * Pacing as app.py:199-202: ticks at 60 Hz while any session runs, 12 Hz
    otherwise; paused sessions still receive frames at 12 Hz.

    python -m windtunnel.frontend [--port 8765] [--host 0.0.0.0] [--web-dir DIR]
"""
import argparse
import asyncio
import json
import os
import queue
import sys
import threading
import time

from aiohttp import WSMsgType, web

from . import wire

FRAME_HZ = 60        # app.py:19
PAUSED_HZ = 12       # app.py:20
HEADROOM = 0.85      # app.py:76
# app.py:62-64
DEFAULT_PARAMS = dict(uIn=0.1, tau=0.51, smag=0.16, steps=10, maxAge=1000.0, lref=40.0,
                      seedX=2.0, seedCy=20.0, seedCz=20.0, seedHy=10.0, seedHz=10.0,
                      slice=None, cp=False, running=True, particles=5000)
LATTICE_KEYS = ('uIn', 'tau', 'smag')
STORAGES = ('fp32', 'fp16', 'bf16', 'fp64')
ENGINE = web.AppKey('engine', object)


class Session:
    """One WebSocket connection and the tunnel it drives."""

    def __init__(self, sid, ws):
        self.sid = sid
        self.ws = ws
        self.ready = True
        self.solver = None
        self.res = None
        self.storage = 'fp32'
        self.bytes = 0
        self.p = dict(DEFAULT_PARAMS)
        self.version = 0            # bumped whenever the batched descriptor would change
        self.steps_since = 0
        self.lattice_steps = 0      # steps since the last grid or reset
        self.perf_steps = 0
        self.perf_t = time.perf_counter()
        self.mlups = 0.0
        self.frame_no = 0
        self.last_frame = 0.0
        self.closed = False


class MultiEngine(threading.Thread):
    """The solver thread: all GPU work of every session happens here."""

    def __init__(self, loop, solver_factory=None, footprint=None, resolutions=None, memory=None,
                 device_name=None, batched=True, frame_hz=FRAME_HZ, paused_hz=PAUSED_HZ):
        super().__init__(daemon=True)
        if solver_factory is None or footprint is None or resolutions is None:
            import lbm
            solver_factory = solver_factory or lbm.Solver
            footprint = footprint or lbm.footprint
            resolutions = resolutions or lbm.RESOLUTIONS
            if memory is None or device_name is None:
                dev = lbm.get_preferred_device()
                if not dev.is_cuda:
                    raise RuntimeError('the front end needs a HIP GPU')
                memory = dev.total_memory if memory is None else memory
                device_name = dev.name if device_name is None else device_name
        self.loop = loop
        self.factory = solver_factory
        self.footprint = footprint
        self.resolutions = dict(resolutions)
        self.memory = int(memory)
        self.budget = int(self.memory * HEADROOM)
        self.device_name = device_name or 'unknown'
        self.batched = batched
        self.frame_hz, self.paused_hz = frame_hz, paused_hz
        self.cmds = queue.Queue()
        self.sessions = {}
        self.used = 0
        self._next_sid = 0
        self._sid_lock = threading.Lock()
        self._halt = threading.Event()
        self._batches = {}
        self.ticks = 0
        self.cell_steps = 0
        self.batched_launches = 0
        self.errors = []

    # ---- called from the event loop ------------------------------------------
    def open_session(self, ws):
        with self._sid_lock:
            sid = self._next_sid
            self._next_sid += 1
        s = Session(sid, ws)
        self.cmds.put(('open', None, None, s))
        return s

    def close_session(self, s):
        s.closed = True
        self.cmds.put(('close', None, None, s))

    def stop(self):
        self._halt.set()

    # ---- admission -------------------------------------------------------------
    def fits(self, dims, storage, s):
        """The session's grid fits beside every other resident session."""
        need = self.footprint(dims, storage)
        return need + (self.used - s.bytes) <= self.budget

    def hello(self, s):
        """app.py:78-82, `fits` checked against the memory this session may still use."""
        res = [{'name': k, 'dims': list(v), 'fits': self.fits(v, s.storage, s)} for k, v in self.resolutions.items()]
        high = self.resolutions.get('high')
        default = 'high' if high and self.fits(high, s.storage, s) else 'test'
        return {'type': 'hello', 'gpu': self.device_name, 'cuda': True, 'memory': self.memory,
                'resolutions': res, 'default': default, 'res': s.res, 'sessions': len(self.sessions),
                'storages': list(STORAGES)}

    # ---- sending (thread-safe) ---------------------------------------------------
    def send_json(self, s, obj):
        async def go():
            try:
                await s.ws.send_json(obj)
            except Exception:
                pass
        asyncio.run_coroutine_threadsafe(go(), self.loop)

    def send_frame(self, s, data):
        async def go():
            try:
                await s.ws.send_bytes(data)
            except Exception:
                pass
        asyncio.run_coroutine_threadsafe(go(), self.loop)

    # ---- commands (app.py:104-148, per session) -----------------------------------
    def _release(self, s):
        if s.solver is not None:
            s.solver = None
            self.used -= s.bytes
            s.bytes = 0
            s.res = None
            s.version += 1
            self._batches.clear()

    def handle(self, kind, msg, payload, s):
        if kind == 'open':
            self.sessions[s.sid] = s
            return
        if kind == 'close':
            self._release(s)
            self.sessions.pop(s.sid, None)
            return
        if s.sid not in self.sessions:
            return
        t = msg.get('type')
        if t == 'grid':
            res = msg.get('res')
            dims = self.resolutions.get(res)
            storage = msg.get('storage', s.storage)
            if storage not in STORAGES:
                self.send_json(s, {'type': 'error', 'message': f'Unknown storage {storage}'})
                return
            if not dims or not self.fits(dims, storage, s):
                self.send_json(s, {'type': 'error', 'message': f'Resolution {res} does not fit on this device'})
                return
            if res != s.res or storage != s.storage or s.solver is None:
                self._release(s)
                try:
                    s.solver = self.factory(dims, None, storage=storage)
                    s.solver.reset(s.p['uIn'])
                    s.res, s.storage = res, storage
                    s.lattice_steps = 0
                    s.bytes = self.footprint(dims, storage)
                    self.used += s.bytes
                except Exception as e:  # out of memory etc. (app.py:119-122)
                    s.solver, s.res = None, None
                    self.send_json(s, {'type': 'error', 'message': f'Could not allocate {res} grid: {e}'})
                    return
            self.send_json(s, {'type': 'grid', 'res': s.res, 'dims': list(self.resolutions[s.res]),
                               'storage': s.storage})
            return
        if s.solver is None:
            return
        sv = s.solver
        if t == 'params':
            old = s.p.get('particles')
            before = tuple(s.p.get(k) for k in LATTICE_KEYS)
            s.p.update({k: v for k, v in msg.items() if k != 'type'})
            if tuple(s.p.get(k) for k in LATTICE_KEYS) != before:
                s.version += 1
            if s.p['particles'] != old and sv.num_p:
                sv.set_particles(int(s.p['particles']), s.p)
        elif t == 'respawn':
            sv.set_particles(int(s.p['particles']), s.p)
        elif t == 'reset':
            sv.reset(s.p['uIn'])
            s.steps_since = 0
            s.lattice_steps = 0
            s.version += 1
        elif t == 'solid':
            arr = wire.solid_from_command(msg, payload, (sv.nx, sv.ny, sv.nz))
            if arr is not None:
                sv.set_solid(arr)
        elif t == 'surface':
            vn = wire.surface_from_command(msg, payload)
            if vn is not None:
                sv.set_surface(*vn)

    # ---- stepping ---------------------------------------------------------------
    def _advance(self, running):
        groups = {}
        for s in running:
            steps = int(s.p.get('steps', 10))
            if self.batched and hasattr(s.solver, 'T'):
                key = (s.storage, steps)
            else:
                key = ('solo', s.sid, steps)
            groups.setdefault(key, []).append(s)
        for key, members in groups.items():
            steps = key[-1]
            try:
                if len(members) > 1:
                    self._step_batched(key, members, steps)
                else:
                    members[0].solver.step(steps, members[0].p)
            except Exception as e:
                self.errors.append(repr(e))
                print('solver error:', repr(e), flush=True)
                continue
            for s in members:
                s.steps_since += steps
                s.lattice_steps += steps
                s.perf_steps += steps
                self.cell_steps += steps * s.solver.n

    def _step_batched(self, key, members, steps):
        from ._lib import WT_FLAG_FORCES, WT_FLAG_NO_MAC
        from .batch import Batch
        sig = tuple((s.sid, s.version, s.solver.T.cur) for s in members)
        b = self._batches.get(key)
        if b is None or b.sig != sig:
            b = Batch([s.solver.T for s in members],
                      [tuple(float(s.p[k]) for k in LATTICE_KEYS) for s in members])
            b.set_gates([False] * len(members))     # WT_FLAG_NO_MAC steps store no field
        for k in range(steps):
            # as Solver.step: only the last step's macroscopic field is read
            b.step(flags=WT_FLAG_FORCES if k == steps - 1 else WT_FLAG_NO_MAC)
        self.batched_launches += steps
        sums = b.last_sums()
        for i, s in enumerate(members):
            s.solver.set_forces(sums[i])
            s.solver.advance_particles(steps, s.p)
        b.sig = tuple((s.sid, s.version, s.solver.T.cur) for s in members)
        self._batches[key] = b

    # ---- frames (app.py:173-195, per session) --------------------------------------
    def _frame(self, s, now):
        sv = s.solver
        running = bool(s.p.get('running'))
        if not s.ready or not (running or now - s.last_frame > 1 / self.paused_hz):
            return
        s.frame_no += 1
        every = max(3, sv.num_verts // 20000)
        out = sv.frame(s.p, bool(s.p.get('cp')) and s.frame_no % every == 0)
        if now - s.perf_t > 1.0:
            s.mlups = s.perf_steps * sv.n / 1e6 / (now - s.perf_t)
            s.perf_steps, s.perf_t = 0, now
        data = wire.frame(sv, out, s.steps_since, s.mlups, running)
        s.steps_since = 0
        s.ready = False
        s.last_frame = now
        self.send_frame(s, data)

    def run(self):
        while not self._halt.is_set():
            t0 = time.perf_counter()
            try:
                while True:
                    self.handle(*self.cmds.get_nowait())
            except queue.Empty:
                pass
            except Exception as e:
                self.errors.append(repr(e))
                print('command failed:', repr(e), flush=True)
            live = [s for s in self.sessions.values() if s.solver is not None and not s.closed]
            if not live:
                time.sleep(0.02)
                continue
            running = [s for s in live if s.p.get('running') and int(s.p.get('steps', 10)) > 0]
            self._advance(running)
            self.ticks += 1
            now = time.perf_counter()
            for s in live:
                try:
                    self._frame(s, now)
                except Exception as e:
                    self.errors.append(repr(e))
                    print('frame error:', repr(e), flush=True)
            spare = 1 / (self.frame_hz if running else self.paused_hz) - (time.perf_counter() - t0)
            if spare > 0:
                self._halt.wait(spare)


async def ws_handler(request):
    """app.py:205-227, one session per connection."""
    engine = request.app[ENGINE]
    ws = web.WebSocketResponse(max_msg_size=1024 ** 3, heartbeat=20)
    await ws.prepare(request)
    s = engine.open_session(ws)
    await ws.send_json(engine.hello(s))
    try:
        async for msg in ws:
            if msg.type == WSMsgType.TEXT:
                data = json.loads(msg.data)
                if data.get('type') == 'ready':
                    s.ready = True
                else:
                    engine.cmds.put(('json', data, None, s))
            elif msg.type == WSMsgType.BINARY:
                header, payload = wire.parse_command(msg.data)
                engine.cmds.put(('bin', header, payload, s))
    finally:
        engine.close_session(s)
    return ws


def make_app(web_dir=None, **engine_kwargs):
    """aiohttp application: /ws plus, when `web_dir` exists, the viewer's static files."""
    async def on_startup(app):
        engine = MultiEngine(asyncio.get_running_loop(), **engine_kwargs)
        app[ENGINE] = engine
        engine.start()

    async def on_cleanup(app):
        engine = app[ENGINE]
        engine.stop()
        await asyncio.get_running_loop().run_in_executor(None, engine.join, 5.0)

    app = web.Application()
    app.on_startup.append(on_startup)
    app.on_cleanup.append(on_cleanup)
    app.router.add_get('/ws', ws_handler)
    if web_dir and os.path.isdir(web_dir):
        async def index(request):
            return web.FileResponse(os.path.join(web_dir, 'index.html'), headers={'Cache-Control': 'no-cache'})
        app.router.add_get('/', index)
        app.router.add_static('/', web_dir)
    return app


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--port', type=int, default=int(os.environ.get('PORT', '8765')))
    ap.add_argument('--host', default='0.0.0.0')
    ap.add_argument('--web-dir', default=os.environ.get('WEB_DIR'))
    args = ap.parse_args(argv)
    app = make_app(args.web_dir)
    print(f'wind tunnel sessions on port {args.port}', flush=True)
    web.run_app(app, host=args.host, port=args.port, print=None)
    return 0


if __name__ == '__main__':
    sys.exit(main())
