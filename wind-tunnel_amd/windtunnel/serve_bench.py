"""Serving benchmark: Poisson arrivals of tunnel sessions (log-uniform 0.5M-33M
cells, 200-5000 steps, half FP32 / half FP16-deviation storage, forces every
step, a slice every 20 steps), admission capped by HBM, continuous batching.

    python -m windtunnel.serve_bench [--duration 600] [--rate 2] [--seed 0] [--graphs] [--time-limit S]
                                     [--tick-target-ms T] [--prefill N] [--record-after S]

--tick-target-ms holds the queue while the predicted tick time (every active
session's step) would exceed T, trading queueing delay for step latency.
--prefill starts with N sessions resident as in a steady state of the same
stream; --record-after measures only from S seconds on.

Prints one JSON line: p50/p99 step latency, queueing delay, aggregate MLUPS.
"""
import argparse
import json
import sys

from .serving import Server, mixed_workload


def run(duration=600.0, rate=2.0, seed=0, graphs=False, time_limit=None, budget_gb=None, progress=True,
        tick_target_ms=None, prefill=0, record_after=0.0):
    reqs = mixed_workload(seed=seed, duration=duration, rate=rate, prefill=prefill)
    budget = None if budget_gb is None else int(budget_gb * 1e9)
    srv = Server(budget_bytes=budget, use_graphs=graphs, tick_target_ms=tick_target_ms)

    def prog(s, t, nact, nwait, npend):
        print(json.dumps({'t': round(t, 1), 'active': nact, 'waiting': nwait, 'pending': npend,
                          'done': len(s.finished), 'mlups_sofar': round(s.cell_steps / max(t, 1e-9) / 1e6, 1)}),
              file=sys.stderr, flush=True)

    st = srv.run(reqs, time_limit=time_limit, progress=(10.0, prog) if progress else None, record_after=record_after)
    st.update({'sessions_offered': len(reqs), 'duration_s': duration, 'rate_per_s': rate, 'seed': seed,
               'graphs': graphs, 'budget_gb': srv.budget / 1e9, 'prefill': prefill, 'record_after_s': record_after})
    return st


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--duration', type=float, default=600.0)
    ap.add_argument('--rate', type=float, default=2.0)
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--graphs', action='store_true')
    ap.add_argument('--time-limit', type=float, default=None)
    ap.add_argument('--budget-gb', type=float, default=None)
    ap.add_argument('--tick-target-ms', type=float, default=None)
    ap.add_argument('--prefill', type=int, default=0)
    ap.add_argument('--record-after', type=float, default=0.0)
    args = ap.parse_args(argv)
    st = run(args.duration, args.rate, args.seed, args.graphs, args.time_limit, args.budget_gb,
             tick_target_ms=args.tick_target_ms, prefill=args.prefill, record_after=args.record_after)
    print(json.dumps(st), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
