"""A tunnel session: one resident tunnel with its particle rake, slice planes and
force stream, driven step by step without host synchronization.

This is the per-session unit of the serving loop and of the benchmark
workloads. Per step it runs the fused step kernel (force fused on frame steps),
advects the particles with dt = 1 (advect_kernel, lbm.py:233-263), and on frame
steps extracts slice planes (slice_kernel, lbm.py:278-318) and queues
asynchronous copies of slices and forces into pinned host buffers.

The advection of step k runs inside the launch of step k + 1 (wt_step_advect)
instead of its own launch after step k (fuse=False keeps separate launches).
Results are identical either way; the fused launch measured faster at 33.5M
cells and 2^20 particles (FUSED_FORMS below).
"""
import numpy as np
import torch

from ._lib import WT_FLAG_FORCES, WT_FLAG_NO_MAC
from .coefficients import ForceAverage
from .tunnel import Tunnel

# storage forms whose sessions fuse the particle advection into the next step's
# launch by default (Session(fuse=...) overrides). Against a separate advection
# launch at 33.5M cells and 2^20 particles (same process, 200-step loops, two
# runs of 3 and 6 alternating repeats): FP16 0.4-0.6 % faster per step fused
# (0.5491 vs 0.5516 ms); FP32 0.05-0.2 % faster separate (0.9742 vs 0.9747 ms),
# so FP32 and FP64 sessions keep the step kernel and the advection apart.
FUSED_FORMS = ('fp16', 'bf16')
# frames whose forces the device-side log holds between two reads of the average
FRAME_LOG = 4096

# slice field ranges of the viewer (web/js/main.js:15-20)
FIELD_RANGES = {0: (0.0, 1.6), 1: (-1.2, 1.2), 2: (0.0, 12.0), 3: (-0.4, 1.4)}


def inlet_rake(dims, ground=True):
    """Inlet-wide particle rake: seedX = 2 and the y/z limits of
    web/js/main.js:267-270 applied to the whole inlet plane."""
    nx, ny, nz = dims
    lo = 1.6 if ground else 1.5
    y0, y1 = lo, ny - 2.5
    z0, z1 = 1.5, nz - 2.5
    return (2.0, (y0 + y1) / 2, (z0 + z1) / 2, (y1 - y0) / 2, (z1 - z0) / 2)


class Session:
    def __init__(self, dims, mask, form='fp32', uin=0.05, tau=0.51, smag=0.16, particles=0,
                 slices=(), frame_every=10, lref=40.0, seed=0, device='cuda', ground=True, frontal=1,
                 fuse=None, side_copies=True):
        self.dims = tuple(int(d) for d in dims)
        self.uin, self.tau, self.smag = float(uin), float(tau), float(smag)
        self.frame_every = int(frame_every)
        self.lref = float(lref)
        self.frontal = max(1, int(frontal))
        self.T = Tunnel(dims, form, device)
        self.T.init(uin)
        self.T.set_solid(mask)
        nx, ny, nz = self.dims
        self.box = inlet_rake(dims, ground)
        # maxAge = 2.2 nx / (U steps) with one step per advect (web/js/main.js:284)
        self.max_age = 2.2 * nx / self.uin
        self.num_p = int(particles)
        self.seed = int(seed) * 1000003 + 1
        self.step_count = 0
        if self.num_p:
            rng = np.random.default_rng(seed)
            x, cy, cz, hy, hz = self.box
            d = np.zeros((self.num_p, 4), np.float32)
            d[:, 0] = x + rng.random(self.num_p) * 2.0
            d[:, 1] = cy + (rng.random(self.num_p) * 2 - 1) * hy
            d[:, 2] = cz + (rng.random(self.num_p) * 2 - 1) * hz
            d[:, 3] = rng.random(self.num_p) * self.max_age * 0.5
            # slot storage: slot j holds particle ids[j]; every sort_every steps the
            # slots are re-gathered in cell order so that advection reads and
            # writes contiguous slots and samples neighbouring cells
            self.pos = torch.from_numpy(d).to(device)
            self.spd = torch.ones(self.num_p, dtype=torch.float32, device=device)
            self.ids = torch.arange(self.num_p, dtype=torch.int32, device=device)
            self._spare = (torch.empty_like(self.pos), torch.empty_like(self.spd), torch.empty_like(self.ids))
            self.order = torch.empty(self.num_p, dtype=torch.int32, device=device)
            self.order_ws = None
            # every 100 steps: 25 and 50 measured 0.3-0.6 % slower per step, 200 and 400 the same
            self.sort_every = 100
        # Fused launches: step k writes macs[k % 2], and the particles of step k are
        # advected inside the launch of step k + 1, from macs[k % 2] (one kernel
        # boundary per step instead of two). flush() advects a pending step alone.
        self.fuse = bool((form in FUSED_FORMS if fuse is None else fuse) and self.num_p)
        self.pending = False
        self.advected = 0             # steps whose particles have been advected
        self._last_sort = None
        if self.num_p:
            # the first sort runs here, at set-up: its one-time cost (the sort library's
            # first call, ~25 ms on the host) then falls outside every step
            self._sort(None)
        if self.fuse:
            self.macs = (self.T.mac, self.T.mac.clone())
        # slice planes: (axis, index, field)
        self.slices = [tuple(s) for s in slices]
        self.slice_dev, self.slice_host = [], []
        for axis, index, field in self.slices:
            w, h = {2: (nx, ny), 1: (nx, nz), 0: (nz, ny)}[axis]
            self.slice_dev.append(torch.zeros(2 * w * h, dtype=torch.uint8, device=device))
            self.slice_host.append(torch.zeros(2 * w * h, dtype=torch.uint8).pin_memory())
        self.forces_host = torch.zeros(4, dtype=torch.float64).pin_memory()
        # the force of every frame, logged on the device (one 32-byte copy per frame)
        # so that the viewer's Cd/Cl average costs no synchronization while running
        self.frame_forces = torch.zeros((FRAME_LOG, 4), dtype=torch.float64, device=device)
        # device-to-host copies of slices and forces go out on a side stream (default
        # priority), so the copy engine overlaps them with the next steps; the next
        # frame's slices and the next force step wait for them (an in-stream event,
        # long satisfied by then). side_copies=False keeps them on the step's stream.
        self.copy_stream = torch.cuda.Stream(device=self.T.device) if side_copies else None
        if self.copy_stream is not None:
            # the first dispatch on a stream sets up its hardware queue: ~2 ms during
            # which the host issues nothing (kernel traces show the GPU idle that long
            # at a session's first frame), so make it here, at set-up
            with torch.cuda.stream(self.copy_stream):
                self.forces_host.copy_(self.T.sums, non_blocking=True)
            self.copy_stream.synchronize()
        self._copied = None
        self.frames = 0
        self._averaged = 0
        self.average = ForceAverage(self.uin, self.frontal)

    @property
    def n(self):
        return self.T.n

    def is_frame(self, k=None):
        k = self.step_count if k is None else k
        return self.frame_every > 0 and (k + 1) % self.frame_every == 0

    def step(self, force=None, stream=None, events=None):
        """One lattice step (+ particles); frame work on frame steps. Asynchronous.
        events: optional (start, end) CUDA events recorded around the step kernel."""
        frame = self.is_frame()
        want_force = frame if force is None else force
        main = stream if stream is not None else torch.cuda.current_stream(self.T.device)
        flags = WT_FLAG_FORCES if want_force else 0
        if want_force and self._copied is not None:
            main.wait_event(self._copied)        # the reduction rewrites the sums being copied
        if self.fuse:
            b = self.step_count & 1
            self.T.set_mac(self.macs[b])
            if self.pending:
                self._sort_if_due(main)
                self.seed += 1
                if events is not None:
                    events[0].record(main)
                self.T.step_advect(self.uin, self.tau, self.smag, flags, self.macs[b ^ 1], self.pos, self.spd,
                                   self.ids, 1.0, self.max_age, self.seed, self.box, stream=main)
                if events is not None:
                    events[1].record(main)
                self.advected += 1
            else:
                if events is not None:
                    events[0].record(main)
                self.T.step(1, self.uin, self.tau, self.smag, flags=flags, stream=main)
                if events is not None:
                    events[1].record(main)
            self.pending = True
        else:
            if not self.num_p and not frame:
                flags |= WT_FLAG_NO_MAC      # no particle or slice reads this step's field
            if events is not None:
                events[0].record(main)
            self.T.step(1, self.uin, self.tau, self.smag, flags=flags, stream=main)
            if events is not None:
                events[1].record(main)
            if self.num_p:
                self._advect(main)
        if frame:
            self.emit_frame(main)
        elif want_force:
            self._copy_out(main, slices=False)
        self.step_count += 1

    def _sort_if_due(self, stream):
        """Re-gather the particle slots in cell order every sort_every advections."""
        if self._last_sort is None or self.advected - self._last_sort >= self.sort_every:
            self._sort(stream)

    def _sort(self, stream):
        self.order_ws = self.T.particle_order(self.pos, self.order, self.order_ws, stream=stream)
        pos2, spd2, ids2 = self._spare
        self.T.particle_gather(self.pos, self.spd, self.ids, self.order, pos2, spd2, ids2, stream=stream)
        self._spare = (self.pos, self.spd, self.ids)
        self.pos, self.spd, self.ids = pos2, spd2, ids2
        self._last_sort = self.advected

    def _advect(self, stream=None):
        """Particle advection with dt = 1 on slot storage re-sorted by cell every sort_every steps."""
        self._sort_if_due(stream)
        self.seed += 1
        self.T.advect_slots(self.pos, self.spd, self.ids, 1.0, self.uin, self.max_age, self.seed, self.box,
                            stream=stream)
        self.advected += 1

    def flush(self, stream=None):
        """Advect the particles of the last step if a fused launch has not yet."""
        if self.fuse and self.pending:
            self._advect(stream)
            self.pending = False

    def particles_by_id(self):
        """(pos, spd) on the host in particle-id order, every step advected."""
        self.flush()
        torch.cuda.synchronize(self.T.device)
        ids = self.ids.cpu().numpy().astype(np.int64)
        pos, spd = np.empty((self.num_p, 4), np.float32), np.empty(self.num_p, np.float32)
        pos[ids] = self.pos.cpu().numpy()
        spd[ids] = self.spd.cpu().numpy()
        return pos, spd

    def emit_frame(self, stream=None):
        s = stream if stream is not None else torch.cuda.current_stream(self.T.device)
        with torch.cuda.stream(s):
            if self._copied is not None:
                s.wait_event(self._copied)      # the last frame's copies still read slice_dev
            for (axis, index, field), dev in zip(self.slices, self.slice_dev):
                vmin, vmax = FIELD_RANGES.get(field, (0.0, 1.0))
                self.T.slice(dev, axis, index, field, self.uin, self.lref, vmin, vmax, stream=s)
            self.frame_forces[self.frames % FRAME_LOG].copy_(self.T.sums, non_blocking=True)
        self.frames += 1
        self._copy_out(s, slices=True)

    def _copy_out(self, s, slices):
        """Queue the copies of the slices (if asked) and the force sums into pinned
        host memory after the work queued so far on stream s."""
        cs = self.copy_stream or s
        if cs is not s:
            ready = torch.cuda.Event()
            ready.record(s)
            cs.wait_event(ready)
        with torch.cuda.stream(cs):
            if slices:
                for dev, host in zip(self.slice_dev, self.slice_host):
                    host.copy_(dev, non_blocking=True)
            self.forces_host.copy_(self.T.sums, non_blocking=True)
            if cs is not s:
                self._copied = torch.cuda.Event()
                self._copied.record(cs)

    def viewer_coefficients(self):
        """(Cd, Cl) as the reference viewer shows them: every frame's force folded
        into windtunnel.coefficients.ForceAverage (web/js/main.js:336-351), frames
        of frame_every steps. None while settling or after divergence. Synchronizes."""
        if self.frames - self._averaged > FRAME_LOG:
            raise RuntimeError(f'{self.frames - self._averaged} frames since the last read; the log holds '
                               f'{FRAME_LOG}')
        torch.cuda.synchronize(self.T.device)
        log = self.frame_forces.cpu().numpy()
        for k in range(self._averaged, self.frames):
            f = log[k % FRAME_LOG]
            self.average.add(f[0], f[1], self.frame_every)
        self._averaged = self.frames
        a = self.average
        return (a.cd, a.cl) if a.settled and not a.diverged else None

    def coefficients(self):
        """Cd, Cl from the last copied forces (web/js/main.js:336-339); call after a sync."""
        f = self.forces_host.numpy()
        q = 0.5 * self.uin * self.uin * self.frontal
        return f[0] / q, f[1] / q
