"""Benchmark of the fused D3Q19 wind-tunnel step on one MI355X.

Workload (the single-GPU configuration the headline metric is quoted on):
a 512 x 256 x 256 = 33.5M-cell tunnel, car-like body = seeded union of 12
ellipsoids on a moving ground belt, uin 0.05, tau 0.51, Cs 0.16, 2^20 particles
seeded on the inlet plane and advected every step, three mid-span slice planes
(speed, pressure, vorticity) and the force extracted every 10 steps. A "step"
is one lattice update of the whole tunnel plus the particle advection. By
default each storage form runs 20 warm-up and 2000 timed steps; the batch of
48 tunnels runs 1000 steps (forces every step; nothing reads the macroscopic
field, so its steps do not store it; reported beside it: the same steps
storing it every step) and the serving mix a 30 s window at steady-state
occupancy (120 sessions resident from the start, then Poisson arrivals),
admitted under a 30 ms tick budget against a stated p99 step latency of
33.3 ms. A precision block reports each storage form's error against the FP64
form: fields at step 200 of the 384x192x192 sphere and Cd/Cl averaged over
steps 6000-10000 of the developed 128x64x64 sphere flow.

value = MLUPS with FP32 populations (all nx*ny*nz cells per step, as
server/app.py:180 counts them); the same workload with FP16-deviation storage
is reported beside it. roofline: algorithmic bytes of the step kernel
(169 B/cell FP32: 19 populations read and written, the class byte, the 16 B
float4 macroscopic store; 85 B/cell FP16: 2-byte deviations and an 8 B binary16
macroscopic store) over its mean launch time, from HIP events on
the stream the kernel runs on, bracketing every 5th launch of the timed
region (session steps 0 mod 5: force-free; a pair of events on every step slows
the loop by ~0.7 %). FP32 sessions advect each step's particles with their
own launch after the step (k_advect); FP16 sessions advect them inside the next
step's launch (k_step_adv: 0.4-0.6 % faster per step for FP16, for FP32 the
separate launches are 0.05-0.2 % faster), so that kernel is priced at 40 more
bytes per particle (id 4, position 16 read + 16 written, speed 4 written); the
last step's particles are advected by their own launch at the end of the
timed region.
cpu_baseline: the CPU restatement of the same step (oracle/, C++ with OpenMP)
timed on this host on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, 'wind-tunnel_amd'))
sys.path.insert(0, REPO)

PEAK_HBM_GBS = 8000.0
BYTES_PER_CELL = {'fp32': 169, 'fp16': 85, 'bf16': 85, 'fp64': 321}
BYTES_PER_PARTICLE = 40
EVENT_EVERY, EVENT_PHASE = 5, 0      # HIP events around steps numbered 0 mod 5 (force-free)
FORM_IDS = {'fp32': 0, 'fp16': 1, 'bf16': 2, 'fp64': 3}


DIMS = (512, 256, 256)
UIN, TAU, SMAG = 0.05, 0.51, 0.16
BODY_SEED = 2025
PARTICLES = 1 << 20
# serving leg: sessions resident at the start, as a steady state of the stream holds
# them (about 120-140 at 2 sessions/s), and the seconds before recording starts
SERVE_PREFILL, SERVE_WARM_S = 120, 10.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_mask(dims):
    from windtunnel import geometry
    ells = geometry.car_body(dims, BODY_SEED)
    mask, counts = geometry.ellipsoids(ells, dims, ground=True)
    lref = float(ells[:, 0].max() + ells[:, 3].max() - (ells[:, 0].min() - ells[:, 3].max()))
    return mask, counts, lref


def make_session(form, mask, counts, lref):
    from windtunnel.session import Session
    nx, ny, nz = DIMS
    slices = [(2, nz // 2, 0), (2, nz // 2, 1), (2, nz // 2, 2)]
    return Session(DIMS, mask, form, UIN, TAU, SMAG, particles=PARTICLES, slices=slices, frame_every=10,
                   lref=lref, seed=7, frontal=counts['frontal'])


def run_gpu(s, form, steps, warmup):
    import torch
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    for _ in range(warmup):
        s.step()
    s.flush()
    torch.cuda.synchronize()
    # steps whose session step number is 0 mod 5 (frames, the only force steps,
    # are 9 mod 10), never the first timed launch (a fused session's is a plain
    # step: nothing pending)
    c0 = s.step_count
    sampled = [k for k in range(1, steps) if (c0 + k) % EVENT_EVERY == EVENT_PHASE and not s.is_frame(c0 + k)]
    if not sampled:                        # runs of a few steps: the last one
        sampled = [steps - 1]
    ev = {k: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for k in sampled}
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    h0 = time.perf_counter()
    t_start.record(stream)
    for k in range(steps):
        s.step(stream=stream, events=ev.get(k))
    s.flush(stream)                        # the last step's particles (fused sessions)
    t_end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - h0
    gpu_ms = t_start.elapsed_time(t_end)
    plain = [ev[k][0].elapsed_time(ev[k][1]) for k in sampled]
    step_ms = sum(plain) / max(1, len(plain))
    srt = sorted(plain)
    n = s.n
    nbytes = BYTES_PER_CELL[form] * n + (BYTES_PER_PARTICLE * s.num_p if s.fuse else 0)
    res = {
        'form': form, 'wall_s': wall, 'gpu_ms': gpu_ms, 'mlups': n * steps / wall / 1e6,
        'ms_per_step': wall * 1e3 / steps, 'step_kernel_ms': step_ms,
        'kernel_launches_timed': len(plain), 'kernel_ms_min': srt[0], 'kernel_ms_median': srt[len(srt) // 2],
        'kernel_ms_max': srt[-1],
        'kernel_mlups': n / (step_ms * 1e-3) / 1e6, 'fused_advect': s.fuse, 'bytes_per_launch': nbytes,
        'kernel': (f"k_step_adv<{FORM_IDS[form]}, {s.T.shape[0]}, {s.T.cache_policy}> "
                   "(step + previous step's advection)") if s.fuse
                  else f'k_step<{FORM_IDS[form]}, {s.T.shape[0]}, {s.T.cache_policy}> (fused stream/collide)',
        'achieved_gbs': nbytes / (step_ms * 1e-3) / 1e9,
    }
    return res


def cpu_baseline(mask, target_s=12.0):
    """CPU restatement of the same FP32 step (oracle, C++/OpenMP) on this host."""
    from oracle import oracle as O
    threads = int(os.environ.get('OMP_NUM_THREADS') or os.cpu_count() or 1)
    O.set_threads(threads)
    t = O.OracleTunnel(DIMS, O.FP32)
    t.init(UIN)
    t.set_solid(mask)
    t0 = time.perf_counter()
    t.step(1, UIN, TAU, SMAG)
    one = time.perf_counter() - t0
    k = max(1, min(50, int(target_s / max(one, 1e-3))))
    t0 = time.perf_counter()
    t.step(k, UIN, TAU, SMAG)
    dt = time.perf_counter() - t0
    n = t.n
    # one core on the same tunnel (memory already touched by the run above)
    O.set_threads(1)
    t0 = time.perf_counter()
    t.step(1, UIN, TAU, SMAG)
    one_core = n / (time.perf_counter() - t0) / 1e6
    del t
    # config 1: 128x64x64, sphere D=16 at (40,32,32), uin 0.05, tau 0.52, Cs 0.17, no ground
    from windtunnel import geometry
    d1 = (128, 64, 64)
    m1 = geometry.sphere(d1, (40, 32, 32), 16)[0]
    c1 = {}
    for nthr, steps in ((threads, 500), (1, 60)):
        O.set_threads(nthr)
        t1 = O.OracleTunnel(d1, O.FP32)
        t1.init(0.05)
        t1.set_solid(m1)
        t0 = time.perf_counter()
        t1.step(steps, 0.05, 0.52, 0.17)
        c1[nthr] = (steps, d1[0] * d1[1] * d1[2] * steps / (time.perf_counter() - t0) / 1e6)
    O.set_threads(threads)
    return {'value': round(n * k / dt / 1e6, 2), 'unit': 'MLUPS', 'cores': threads, 'kind': 'port',
            'sample': f'{k} FP32 lattice steps (no particles/slices) of the same 512x256x256 tunnel and '
                      f'body, after 1 untimed step; oracle/wt_oracle.cpp with {threads} OpenMP threads',
            'one_core': round(one_core, 2),
            'config1_128x64x64': {'mlups_all_cores': round(c1[threads][1], 2), 'steps_all_cores': c1[threads][0],
                                  'mlups_one_core': round(c1[1][1], 2), 'steps_one_core': c1[1][0]}}


def load_traffic(form, fused):
    path = os.path.join(REPO, 'profiles', f'step{"_adv" if fused else ""}_{form}_512x256x256_traffic.json')
    if os.path.exists(path):
        with open(path) as fh:
            return json.load(fh).get('hbm_bytes_per_launch')
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=2000)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--forms', default='fp32,fp16')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-precision', action='store_true')
    ap.add_argument('--c4-steps', type=int, default=1000)
    ap.add_argument('--serve-seconds', type=float, default=30.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py measures one GPU: there is no multi-GPU path (asked for --gpus {args.gpus}). '
              'Independent tunnels shard one per GPU with no communication; one tunnel larger than a GPU '
              'would shard as x-slabs exchanging a one-cell halo of 5 populations per face.', file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print('bench.py needs a HIP GPU', file=sys.stderr)
        sys.exit(1)
    from windtunnel import _lib
    _lib.load(build_if_missing=True)
    log(f'building geometry {DIMS} ...')
    mask, counts, lref = build_mask(DIMS)
    log(f'body: {counts}')
    # every session is set up first (state, geometry, particles, first sort), so the
    # forms then run back to back: the GPU does not sit idle between them, which kernel
    # traces show costs the following steps 5-10 % while its clocks recover
    forms = args.forms.split(',')
    sessions = {form: make_session(form, mask, counts, lref) for form in forms}
    results = {}
    for form in forms:
        log(f'GPU run {form}: warmup {args.warmup}, {args.steps} steps ...')
        results[form] = run_gpu(sessions[form], form, args.steps, args.warmup)
    for form in forms:
        log(json.dumps(results[form]))
    del sessions
    torch.cuda.empty_cache()
    c4 = None
    if args.c4_steps > 0:
        from windtunnel.workloads import run_config4
        log(f'config-4 batch: 48 tunnels, {args.c4_steps} steps ...')
        c4 = run_config4(steps=args.c4_steps, warmup=20, graphs=True)
        log(json.dumps(c4))
        # the same steps storing the macroscopic field every step, as the reference's step does
        c4m = run_config4(steps=args.c4_steps, warmup=20, graphs=True, store_mac=True)
        log(json.dumps(c4m))
        c4['mac_every_step'] = c4m
        # the same steps by the persistent batched kernel (one grid of full occupancy
        # claiming the 48 tunnels' tiles in order)
        c4p = run_config4(steps=args.c4_steps, warmup=20, graphs=True, persistent=True)
        log(json.dumps(c4p))
        c4['persistent_kernel'] = c4p
        torch.cuda.empty_cache()
    serving = None
    if args.serve_seconds > 0:
        from windtunnel.serve_bench import run as serve_run
        from windtunnel.serving import P99_TARGET_MS, TICK_TARGET_MS
        log(f'serving mix: {SERVE_PREFILL} sessions resident, then Poisson 2/s; {args.serve_seconds:.0f} s '
            f'recorded after {SERVE_WARM_S:.0f} s ...')
        serving = serve_run(duration=SERVE_WARM_S + args.serve_seconds, rate=2.0, seed=0, graphs=False,
                            time_limit=SERVE_WARM_S + args.serve_seconds, progress=False,
                            tick_target_ms=TICK_TARGET_MS, prefill=SERVE_PREFILL, record_after=SERVE_WARM_S)
        log(json.dumps(serving))
        torch.cuda.empty_cache()
    precision = None
    if not args.no_precision:
        from windtunnel import precision as P
        log('precision against the FP64 form: fields at step 200 (384x192x192 sphere), '
            'Cd/Cl over steps 6000-10000 (128x64x64 sphere) ...')
        fe = P.field_errors(scale=3, step=200)
        co = P.coefficients(scale=1, avg_from=6000, avg_to=10000)
        precision = {
            'field_errors_step200': {'workload': 'sphere D=48 at (120,96,96) in 384x192x192, uin 0.05, tau 0.52, '
                                                 'Cs 0.17, 200 steps from rest; fluid cells',
                                     **{f: {k: round(v, 8) for k, v in e.items()} for f, e in fe['forms'].items()}},
            'cd_cl_time_average': {'workload': 'sphere D=16 at (40,32,32) in 128x64x64, uin 0.05, tau 0.52, Cs 0.17 '
                                               '(Re 120), force every step, mean over steps 6000-10000',
                                   **{f: {k: round(v, 7) for k, v in r.items() if k != 'finite'}
                                      for f, r in co['forms'].items()}},
        }
        log(json.dumps(precision))
        torch.cuda.empty_cache()
    cpu = None
    if not args.no_cpu:
        log('CPU baseline ...')
        cpu = cpu_baseline(mask)
        log(json.dumps(cpu))
    main_form = 'fp32' if 'fp32' in results else next(iter(results))
    r = results[main_form]
    line = {
        'metric': 'MLUPS @33M cells fp32 vs fp16-store; p50/p99 step latency, mixed-size serving',
        'value': round(r['mlups'], 1), 'unit': 'MLUPS', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(r['ms_per_step'], 4), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': 'single tunnel 512x256x256 (33.5M cells), 12-ellipsoid car body on a moving '
                               'belt, 2^20 particles advected every step, 3 slices + forces every 10 steps',
                   'cells': DIMS[0] * DIMS[1] * DIMS[2], 'storage': main_form, 'uin': UIN, 'tau': TAU,
                   'smag': SMAG, 'particles': PARTICLES, 'parallelism': 'single GPU'},
        'roofline': {'bound': 'hbm', 'achieved': round(r['achieved_gbs'], 1), 'peak': PEAK_HBM_GBS,
                     'unit': 'GB/s', 'frac': round(r['achieved_gbs'] / PEAK_HBM_GBS, 4),
                     'traffic': load_traffic(main_form, r['fused_advect']),
                     'kernel': r['kernel'], 'bytes_per_launch': r['bytes_per_launch'],
                     'bytes_per_cell': BYTES_PER_CELL[main_form],
                     'bytes_per_particle': BYTES_PER_PARTICLE if r['fused_advect'] else 0,
                     'kernel_ms': round(r['step_kernel_ms'], 4)},
        'cpu_baseline': cpu,
    }
    for form, rr in results.items():
        if form == main_form:
            continue
        line[form] = {'value': round(rr['mlups'], 1), 'ratio_vs_fp32': round(rr['mlups'] / r['mlups'], 3),
                      'kernel': rr['kernel'], 'kernel_ms': round(rr['step_kernel_ms'], 4),
                      'roofline_frac': round(rr['achieved_gbs'] / PEAK_HBM_GBS, 4),
                      'achieved_gbs': round(rr['achieved_gbs'], 1), 'bytes_per_launch': rr['bytes_per_launch'],
                      'bytes_per_cell': BYTES_PER_CELL[form], 'traffic': load_traffic(form, rr['fused_advect'])}
    # each result's error against the FP64 form: fields early in a flow and Cd/Cl
    # averaged over a developed window (the 20-step runs above are a start-up transient)
    if precision:
        line['precision'] = precision
    if c4:
        line['config4_batch'] = {'value': round(c4['mlups'], 1), 'unit': 'MLUPS', 'tunnels': c4['tunnels'],
                                 'cells': c4['cells'], 'steps': c4['steps'], 'ms_per_step': round(c4['ms_per_step'], 4),
                                 'launch': 'one batched launch per step (graph replay of 2 steps), forces every step',
                                 'macroscopic_field': 'not stored: nothing in the workload reads it (WT_FLAG_NO_MAC)',
                                 'value_storing_field_every_step': round(c4['mac_every_step']['mlups'], 1),
                                 'value_persistent_kernel': round(c4['persistent_kernel']['mlups'], 1)}
    if serving:
        line['serving'] = {'step_latency_ms_p50': round(serving['step_latency_ms']['p50'], 3),
                           'step_latency_ms_p99': round(serving['step_latency_ms']['p99'], 3),
                           'p99_target_ms': P99_TARGET_MS, 'tick_target_ms': TICK_TARGET_MS,
                           'p99_held': serving['step_latency_ms']['p99'] <= P99_TARGET_MS,
                           'active_sessions': serving['active_sessions'],
                           'queue_delay_s_p50': round(serving['queue_delay_s']['p50'], 3),
                           'queue_delay_s_p99': round(serving['queue_delay_s']['p99'], 3),
                           'mlups': round(serving['mlups'], 1), 'window_s': round(serving['window_s'], 1),
                           'prefill': SERVE_PREFILL,
                           'sessions_admitted': serving['sessions_admitted'],
                           'sessions_done': serving['sessions_done'],
                           'workload': 'Poisson 2 sessions/s, 0.5M-33M cells log-uniform, 200-5000 steps, '
                                       'half fp32 half fp16, forces every step, slice every 20 steps'}
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
