"""Sweep server and HIP-graph stepping on the GPU: every result must be the
same alone and in the mix (bit for bit: the same kernels run on the same
data; only their interleaving with other simulations differs)."""
import json
import time

import numpy as np
import pytest

import chiml_amd
from chiml_amd import decks

pytestmark = pytest.mark.gpu


def _mix():
    out = []
    out.append(decks.config1_2d(n=48, steps=90, pml=8, cyl_radius_cells=6, src_cell=(16, 24), ndet=2,
                                precision="fp32"))
    out.append(decks.config2_3d(n=24, steps=70, pml=4, d=0.02, sphere_cells=4, flux_cells=10, nfreq=5,
                                precision="fp32"))
    d = json.loads(decks.config3_ml2d(n=48, steps=120, pml=6, slab=(8, 4), precision="fp32", det_cells=2))
    fc = decks.ev2fdtd(2.0)
    d["SourceList"][0]["PulseList"] = [{"profile": "gaussian", "fcen": fc, "fwidth": fc, "cutoff": 3.0, "t_0": 0.5,
                                        "Field_Intensity": 1e10}]
    out.append(json.dumps(d))
    out.append(decks.config2_3d(n=32, steps=50, pml=8, d=0.02, sphere_cells=5, flux_cells=12, nfreq=4,
                                precision="fp64"))
    out.append(decks.config1_2d(n=64, steps=75, pml=10, cyl_radius_cells=9, src_cell=(20, 32), ndet=3,
                                precision="fp64"))
    # an odd-size grid: the generic per-cell kernels instead of the vector ones
    out.append(decks.config1_2d(n=45, steps=64, pml=6, cyl_radius_cells=5, src_cell=(14, 22), ndet=2,
                                precision="fp32"))
    # N-level emitters in 3D: the team density kernel
    d = decks.empty_deck((24, 24, 24), 0.02, 60, 4, precision="fp32")
    d["ObjectList"].append(decks.n_level_emitter(4, [0.04, 0.04, 0.04], (0.01, 0.01, 0.01), seed=44))
    fc = decks.ev2fdtd(2.0)
    d["SourceList"].append({"pol": "Ez", "size": [0.0, 0.0, 0.0], "loc": [-0.09, 0.01, 0.01],
                            "PulseList": [{"profile": "gaussian", "fcen": fc, "fwidth": fc, "cutoff": 3.0, "t_0": 0.3,
                                           "Field_Intensity": 1e10}]})
    out.append(json.dumps(d))
    return out


def _results(sim):
    r = {}
    for c in range(6):
        r[f"f{c}"] = sim.field(c)
    for k, dtc in enumerate(sim.dtcArr()):
        r[f"d{k}"] = dtc.series()[1]
    for k, fx in enumerate(sim.fluxArr()):
        r[f"x{k}"] = fx.getFlux()
    for k, q in enumerate(sim.qeArr()):
        r[f"p{k}"] = np.concatenate([q.population(0), q.population(1)])
    return r


def _alone(deck, graph=0):
    sim = chiml_amd.FDTDField(deck)
    if graph:
        sim.set_graph(graph)
    sim.step(sim.nsteps)
    sim.sync()
    r = _results(sim)
    sim.close()
    return r


def _same(a, b, label):
    assert a.keys() == b.keys(), label
    for k in a:
        assert a[k].shape == b[k].shape, (label, k)
        assert np.array_equal(a[k], b[k]), (label, k, float(np.max(np.abs(a[k] - b[k]))))


@pytest.mark.parametrize("graph", [8, 32])
def test_graph_steps_equal_plain_launches(graph):
    for i, deck in enumerate(_mix()):
        print(f"graph={graph} deck {i}", flush=True)
        _same(_alone(deck), _alone(deck, graph), f"deck {i}")


def _served_mix():
    # a long request first: it is still stepping while the next ones are built,
    # so the server holds at least two at once however fast the short ones run
    long = decks.config2_3d(n=32, steps=1500, pml=8, d=0.02, sphere_cells=5, flux_cells=12, nfreq=4,
                            precision="fp64")
    return [long] + _mix()


@pytest.mark.parametrize("graph,batch", [(0, False), (16, False), (0, True)])
def test_server_results_equal_alone(graph, batch):
    mix = _served_mix()
    alone = [_alone(d) for d in mix]
    # budget that holds about two of the larger requests at once: the rest queue
    budget = 2 * max(chiml_amd.estimate_bytes(d) for d in mix)
    srv = chiml_amd.Server(budget, max_active=3, chunk_steps=10, graph_steps=graph, batch=batch)
    try:
        ids = [srv.submit(d) for d in mix[:3]]
        time.sleep(0.05)  # the rest arrive while the first ones run
        ids += [srv.submit(d) for d in mix[3:]]
        srv.drain()
        st = srv.stats()
        print("server stats", st, flush=True)
        for k, rid in enumerate(ids):
            state, adm, fin = srv.poll(rid)
            assert state == chiml_amd.Server.DONE, srv.error(rid)
            assert 0 <= adm <= fin
            sim = srv.take(rid)
            _same(alone[k], _results(sim), f"request {k}")
            sim.close()
        assert st["peak_active"] >= 2
        assert srv.stats()["bytes_held"] == 0
    finally:
        srv.close()


def test_server_drain_completes_with_untaken_results_over_budget():
    # results not yet taken fill the budget: each later request is admitted
    # once nothing runs, so drain() returns and every result can be taken after
    mix = [decks.config1_2d(n=32, steps=20 + 5 * k, pml=6, cyl_radius_cells=4, src_cell=(10, 16), ndet=1,
                            precision="fp32") for k in range(4)]
    budget = min(chiml_amd.estimate_bytes(d) for d in mix)  # room for one request at a time
    srv = chiml_amd.Server(budget, max_active=2, chunk_steps=7, graph_steps=0)
    try:
        ids = [srv.submit(d) for d in mix]
        srv.drain()
        for k, rid in enumerate(ids):
            state, _, _ = srv.poll(rid)
            assert state == chiml_amd.Server.DONE, srv.error(rid)
        for k, rid in enumerate(ids):
            sim = srv.take(rid)
            _same(_alone(mix[k]), _results(sim), f"request {k}")
            sim.close()
        assert srv.stats()["bytes_held"] == 0
    finally:
        srv.close()


def _batch_run(mix, chunks):
    """Steps every deck of `mix` to its end through chiml_sim_step_batch, in
    batches of the given chunk lengths (each batch holds the simulations with
    at least that many steps left; the rest finish alone)."""
    sims = [chiml_amd.FDTDField(d) for d in mix]
    done = [0] * len(sims)
    k = 0
    while True:
        left = [s.nsteps - done[i] for i, s in enumerate(sims)]
        if max(left) == 0:
            break
        ch = chunks[k % len(chunks)]
        k += 1
        grp = [i for i in range(len(sims)) if left[i] >= ch]
        if grp:
            chiml_amd.step_batch([sims[i] for i in grp], ch)
            for i in grp:
                done[i] += ch
        else:
            for i in range(len(sims)):
                if left[i]:
                    sims[i].step(left[i])
                    done[i] += left[i]
    out = []
    for s in sims:
        s.sync()
        out.append(_results(s))
        s.close()
    return out


def test_batched_steps_equal_alone():
    """Shared launches (2D and 3D, FP32 and FP64, emitters, flux boxes, and an
    odd-size deck that steps alone inside the batch call) give every
    simulation exactly its results alone."""
    mix = _mix()
    alone = [_alone(d) for d in mix]
    got = _batch_run(mix, [7, 13, 1, 20])
    for k, (a, b) in enumerate(zip(alone, got)):
        print(f"batched deck {k}", flush=True)
        _same(a, b, f"deck {k}")
