"""chiml_dev_step_batch_noalloc (through chiml_sim_step_batch_noalloc): the
batched step whose parameter tables live in the library's static device
array.  Two calls in a row, the second over fewer simulations, must give
what chiml_dev_step_batch gives on the same decks, bit for bit."""
import json

import numpy as np
import pytest

import chiml_amd
from chiml_amd import decks

pytestmark = pytest.mark.gpu


def _decks():
    out = [
        decks.config1_2d(n=48, steps=60, pml=8, cyl_radius_cells=6, src_cell=(16, 24), ndet=2, precision="fp32"),
        decks.config2_3d(n=24, steps=60, pml=4, d=0.02, sphere_cells=4, flux_cells=10, nfreq=5, precision="fp32"),
        decks.config1_2d(n=64, steps=60, pml=10, cyl_radius_cells=9, src_cell=(20, 32), ndet=3, precision="fp64"),
        decks.config2_3d(n=32, steps=60, pml=8, d=0.02, sphere_cells=5, flux_cells=12, nfreq=4, precision="fp64"),
    ]
    d = json.loads(decks.config3_ml2d(n=48, steps=60, pml=6, slab=(8, 4), precision="fp32", det_cells=2))
    fc = decks.ev2fdtd(2.0)
    d["SourceList"][0]["PulseList"] = [{"profile": "gaussian", "fcen": fc, "fwidth": fc, "cutoff": 3.0, "t_0": 0.5,
                                        "Field_Intensity": 1e10}]
    out.append(json.dumps(d))
    return out


def _results(sim):
    r = [sim.field(c) for c in range(6)]
    r += [dtc.series()[1] for dtc in sim.dtcArr()]
    r += [fx.getFlux() for fx in sim.fluxArr()]
    r += [np.concatenate([q.population(0), q.population(1)]) for q in sim.qeArr()]
    return r


def _run(noalloc):
    sims = [chiml_amd.FDTDField(d) for d in _decks()]
    chiml_amd.step_batch(sims, 25, noalloc=noalloc)           # all five
    chiml_amd.step_batch(sims[:3], 20, noalloc=noalloc)       # then three of them
    for s in sims:
        s.sync()
    out = [_results(s) for s in sims]
    for s in sims:
        s.close()
    return out


def test_noalloc_batch_equals_batch():
    ref = _run(False)
    got = _run(True)
    for k, (a, b) in enumerate(zip(ref, got)):
        assert len(a) == len(b)
        for j, (x, y) in enumerate(zip(a, b)):
            assert x.shape == y.shape, (k, j)
            assert np.array_equal(x, y), (k, j, float(np.max(np.abs(x - y))))
        assert np.max(np.abs(a[2])) > 0 or np.max(np.abs(a[0])) > 0, k  # the decks did move


def test_noalloc_batch_rejects_oversized_call():
    sims = [chiml_amd.FDTDField(_decks()[0]) for _ in range(65)]
    with pytest.raises(chiml_amd.ChimlError, match="more than 64"):
        chiml_amd.step_batch(sims, 1, noalloc=True)
    for s in sims:
        s.close()
