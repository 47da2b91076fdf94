"""Many resident tunnels advanced together: one launch of the batched step kernel
per lattice step for all tunnels of one storage form, at the cells per thread
tuned for batches (or, on request, each tunnel's own).

The descriptor table (per-tunnel buffers, sizes, parameters, tile offsets) is
encoded on the host once per batch composition (wt_batch_encode) and copied to
the device in stream order, with no synchronization; each step is then a
kernel launch only (wt_batch_step), so a step of the whole batch can be
captured in a HIP graph. Graph replays advance two steps (parity 0 then 1),
returning every tunnel to the same buffer parity.

A step runs either one workgroup per (tunnel, tile) item, or, with
persistent=True, a grid of one full occupancy whose workgroups stride over the
items (the persistent kernel). Both compute every cell identically.
"""
import ctypes

import torch

from . import _lib, autotune
from ._lib import BatchInfo, Grid, Params, check, stream_handle

# the kernel a Batch steps with unless told otherwise
PERSISTENT = False


class Batch:
    def __init__(self, tunnels, params, cpt=None, own_shapes=False, persistent=None, persist_blocks=0):
        """tunnels: list of Tunnel (same storage form); params: list of (uin, tau, smag).

        Every tunnel runs at cpt cells per thread (default: the value tuned for
        tunnels of its size class in batches of this storage form,
        autotune.batch_cpt), unless own_shapes, when each keeps the cells per
        thread of its own launch shape. persistent picks
        the persistent kernel (default: PERSISTENT), run with persist_blocks
        workgroups if positive, else at the device's full occupancy."""
        if not tunnels:
            raise ValueError('empty batch')
        forms = {t.form for t in tunnels}
        if len(forms) != 1:
            raise ValueError(f'a batch holds one storage form, got {sorted(forms)}')
        self.lib = _lib.load()
        self.tunnels = list(tunnels)
        form_name = _lib.FORM_NAMES[self.tunnels[0].form]
        # cells per thread of each tunnel inside the batch (None: their own shapes)
        self.cpts = None if own_shapes else [int(cpt) if cpt is not None else autotune.batch_cpt(form_name, t.n)
                                            for t in self.tunnels]
        self.params = [tuple(float(v) for v in p) for p in params]
        self.persistent = PERSISTENT if persistent is None else bool(persistent)
        self.persist_blocks = int(persist_blocks)
        if self.persist_blocks < 0:
            raise ValueError(f'persist_blocks must be >= 0, got {persist_blocks}')
        nt = len(self.tunnels)
        nbytes = check(self.lib.wt_batch_workspace_bytes(nt), 'wt_batch_workspace_bytes')
        self.workspace = torch.empty(nbytes, dtype=torch.uint8, device=self.tunnels[0].device)
        off = check(self.lib.wt_batch_gate_offset(nt), 'wt_batch_gate_offset')
        # gate byte per tunnel: under WT_FLAG_NO_MAC only gated tunnels store their field
        self.gates = self.workspace[off:off + nt]
        # force x, y, z and mass of every tunnel after a step of parity 0 / 1
        self.sums = torch.zeros((2, nt, 4), dtype=torch.float64, device=self.tunnels[0].device)
        self.info = BatchInfo()
        self.parity = 0
        self.graph = None
        self.prepare()

    def prepare(self, stream=None):
        nt = len(self.tunnels)
        # each tunnel's tiles hold 256 * cpt cells inside the one launch: the batch's
        # cells per thread, or with own_shapes that of the tunnel's launch shape
        grids = (Grid * nt)(*[t.grid for t in self.tunnels])
        if self.cpts is not None:
            for g, c in zip(grids, self.cpts):
                g.cpt = c
        prm = (Params * nt)(*[Params(*p) for p in self.params])
        cur = (ctypes.c_int32 * nt)(*[t.cur for t in self.tunnels])
        # the descriptor table is written into pinned host memory and copied in stream
        # order: no synchronization, so a change of membership never drains the work in
        # flight (the pinned block returns to torch's host allocator only after the copy)
        image = torch.empty(self.workspace.numel(), dtype=torch.uint8).pin_memory()
        check(self.lib.wt_batch_encode(grids, prm, cur, nt, ctypes.c_void_p(self.sums.data_ptr()),
                                       ctypes.c_void_p(image.data_ptr()), ctypes.byref(self.info)),
              'wt_batch_encode')
        if self.persistent:
            self.info.tile_order |= _lib.WT_BATCH_PERSISTENT
            self.info.persist_blocks = self.persist_blocks
        s = stream if stream is not None else torch.cuda.current_stream(self.workspace.device)
        with torch.cuda.stream(s):
            self.workspace.copy_(image, non_blocking=True)
        self.parity = 0
        self.graph = None

    def set_gates(self, store, stream=None):
        """Gate bytes for the following steps (asynchronous, in stream order): under
        WT_FLAG_NO_MAC tunnel k stores its macroscopic field only if store[k] is true.
        prepare() resets every gate to store."""
        if len(store) != len(self.tunnels):
            raise ValueError(f'{len(store)} gates for {len(self.tunnels)} tunnels')
        host = torch.tensor([1 if v else 0 for v in store], dtype=torch.uint8).pin_memory()
        s = stream if stream is not None else torch.cuda.current_stream(self.workspace.device)
        with torch.cuda.stream(s):
            self.gates.copy_(host, non_blocking=True)

    @property
    def cells(self):
        return sum(t.n for t in self.tunnels)

    def last_sums(self):
        """Device view (ntunnels, 4) of the sums written by the most recent step."""
        return self.sums[self.parity ^ 1]

    def step(self, flags=0, stream=None):
        """One lattice step of every tunnel (asynchronous)."""
        check(self.lib.wt_batch_step(ctypes.c_void_p(self.workspace.data_ptr()), ctypes.byref(self.info),
                                     self.parity, int(flags), stream_handle(stream)), 'wt_batch_step')
        self.parity ^= 1
        for t in self.tunnels:
            t.cur ^= 1

    def capture(self, flags=0):
        """Capture two batch steps (parity 0 and 1) in a HIP graph."""
        if self.parity != 0:
            raise RuntimeError('capture at even parity')
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                for p in (0, 1):
                    check(self.lib.wt_batch_step(ctypes.c_void_p(self.workspace.data_ptr()),
                                                 ctypes.byref(self.info), p, int(flags), stream_handle(s)),
                          'wt_batch_step')
        torch.cuda.current_stream().wait_stream(s)
        self.graph = g
        return g

    def replay(self):
        """Two steps of every tunnel from the captured graph."""
        if self.graph is None:
            raise RuntimeError('no captured graph')
        self.graph.replay()
        # two flips leave every tunnel's cur unchanged
