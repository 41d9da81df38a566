"""Parameter-sweep driver: replays seeded requests with their arrival times
through the sweep server (continuous batching) or one simulation at a time,
and reports throughput and latency percentiles.

Latency of a request = its finish time minus its arrival time (queueing
included).  Throughput = completed cell-updates / (last finish - first
arrival).
"""
import json
import time

import numpy as np

from . import FDTDField, Server


def _cells(deck):
    d = json.loads(deck)
    cc = d["CompCell"]
    n = [int(np.floor(s / h + 0.5)) + 1 if s > 0 else 1 for s, h in zip(cc["size"], cc["stepSize"])]
    return float(n[0] * n[1] * n[2])


def _summary(lat, cell_updates, span, n):
    lat = np.asarray(lat)
    return {"requests": n, "span_s": round(span, 3), "req_per_s": round(n / span, 3),
            "mcells_per_s": round(cell_updates / span / 1e6, 1),
            "p50_s": round(float(np.percentile(lat, 50)), 4), "p99_s": round(float(np.percentile(lat, 99)), 4),
            "mean_s": round(float(lat.mean()), 4)}


def run_server(requests, hbm_budget, max_active=16, chunk_steps=64, graph_steps=32, poll_s=0.002, batch=False, progress=0):
    """requests: [(arrival_s, deck_json)] sorted by arrival.  batch: drained
    simulations take their chunks together through shared launches."""
    srv = Server(hbm_budget, max_active=max_active, chunk_steps=chunk_steps, graph_steps=graph_steps, batch=batch)
    try:
        t0 = time.perf_counter()
        pending = {}
        lat = []
        cell_updates = 0.0
        i = 0
        last_finish = t0
        failed = []
        while i < len(requests) or pending:
            now = time.perf_counter() - t0
            while i < len(requests) and requests[i][0] <= now:
                arr, deck = requests[i]
                t_sub = time.perf_counter() - t0
                pending[srv.submit(deck)] = (arr, deck, t_sub)
                i += 1
            for rid in list(pending):
                state, _, fin = srv.poll(rid)
                if state in (Server.DONE, Server.FAILED):
                    arr, deck, t_sub = pending.pop(rid)
                    if state == Server.FAILED:
                        failed.append(srv.error(rid))
                        continue
                    sim = srv.take(rid)
                    cell_updates += _cells(deck) * sim.nsteps
                    sim.close()
                    done_at = t_sub + fin  # the server's clock: last step finished on the device
                    lat.append(done_at - arr)
                    last_finish = max(last_finish, t0 + done_at)
                    if progress and len(lat) % progress == 0:
                        print(f"server: {len(lat)}/{len(requests)} done at {time.perf_counter() - t0:.1f} s", flush=True)
            time.sleep(poll_s)
        if failed:
            raise RuntimeError(f"{len(failed)} sweep requests failed: {failed[0]}")
        out = _summary(lat, cell_updates, last_finish - t0 - requests[0][0], len(requests))
        out["peak_active"] = srv.stats()["peak_active"]
        return out
    finally:
        srv.close()


def run_one_at_a_time(requests, progress=0):
    """The plain GPU path: each request runs alone, in arrival order."""
    t0 = time.perf_counter()
    lat = []
    cell_updates = 0.0
    for arr, deck in requests:
        now = time.perf_counter() - t0
        if now < arr:
            time.sleep(arr - now)
        sim = FDTDField(deck)
        sim.step(sim.nsteps)
        sim.sync()
        cell_updates += _cells(deck) * sim.nsteps
        sim.close()
        lat.append(time.perf_counter() - t0 - arr)
        if progress and len(lat) % progress == 0:
            print(f"one at a time: {len(lat)}/{len(requests)} done at {time.perf_counter() - t0:.1f} s", flush=True)
    span = time.perf_counter() - t0 - requests[0][0]
    return _summary(lat, cell_updates, span, len(requests))
