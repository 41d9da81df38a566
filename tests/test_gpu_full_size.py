"""Checks at the benchmark's full sizes: 3D 256^3 (the Yee + CPML + ADE +
flux deck) and 3D 192^3 with 368,640 6-level emitters.

The oracle finishes only a few 256^3 steps in seconds, so the full-size
checks are a short oracle comparison plus properties that do not depend on
the size: linearity in the source amplitude (bit-exact in FP64), FP32
against FP64, and the trace and hermiticity of the density matrices.
"""
import json

import numpy as np
import pytest

from chiml_amd import decks
from tests.helpers import rel_linf

pytestmark = pytest.mark.gpu


def _run(deck, steps):
    from chiml_amd import FDTDField
    sim = FDTDField(deck)
    sim.step(steps)
    sim.sync()
    return sim


def _scaled(deck_json, k):
    d = json.loads(deck_json)
    for s in d["SourceList"]:
        for p in s["PulseList"]:
            p["Field_Intensity"] = p["Field_Intensity"] * k
    return json.dumps(d)


def test_256_fp64_matches_oracle():
    """20 steps of the 256^3 deck: the plane-wave sheet spans every x/y slab
    and band of the interior walk, so a mis-indexed region shows at once."""
    from oracle.pyoracle import Oracle
    deck = decks.config2_3d(n=256, steps=20, precision="fp64")
    print("256^3 fp64 vs oracle, 20 steps", flush=True)
    sim = _run(deck, 20)
    orc = Oracle(deck)
    orc.step(20)
    for c in range(6):
        o = orc.field(c)
        err = rel_linf(sim.field(c), o)
        print(f"field {c}: max|ref|={np.abs(o).max():.3e} rel_linf={err:.3e}", flush=True)
        assert np.abs(o).max() > 0.0, c
        assert err <= 1e-9, (c, err)
    sim.close()


def test_256_fp64_exactly_linear_in_source_amplitude():
    """Twice the source amplitude gives exactly twice every field and four
    times the flux, bit for bit: scaling by a power of two commutes with every
    rounding of the update as long as nothing underflows, which FP64 never
    does here.  Any read of stale or uninitialised memory breaks it."""
    base = decks.config2_3d(n=256, steps=1000, precision="fp64")
    print("256^3 fp64 linearity, 1000 steps", flush=True)
    s1 = _run(base, 1000)
    f1 = [s1.field(c) for c in range(6)]
    x1 = s1.fluxArr()[0].getFlux()
    s1.close()
    assert np.abs(x1).max() > 0.0
    s2 = _run(_scaled(base, 2.0), 1000)
    for c in range(6):
        assert np.abs(f1[c]).max() > 0.0, c
        nd = int(np.count_nonzero(s2.field(c) != 2.0 * f1[c]))
        print(f"field {c}: cells not exactly doubled: {nd}", flush=True)
        assert nd == 0, c
    assert np.array_equal(s2.fluxArr()[0].getFlux(), 4.0 * x1)
    s2.close()


def test_256_fp32_linear_to_rounding():
    """The same in FP32 holds only to rounding: values ahead of the wavefront
    pass through FP32's gradual underflow, where doubling is not exact, and
    from there the two runs' roundings decorrelate (each run is itself bit
    for bit reproducible).  The gap stays at FP32's noise level: 1e-4 on the
    fields (the FP32 field bound) and 1e-5 on the flux."""
    base = decks.config2_3d(n=256, steps=1000, precision="fp32")
    print("256^3 fp32 linearity, 1000 steps", flush=True)
    s1 = _run(base, 1000)
    f1 = [s1.field(c) for c in range(6)]
    x1 = s1.fluxArr()[0].getFlux()
    s1.close()
    assert np.abs(x1).max() > 0.0
    s2 = _run(_scaled(base, 2.0), 1000)
    for c in range(6):
        err = rel_linf(s2.field(c), 2.0 * f1[c])
        print(f"field {c}: rel_linf={err:.3e}", flush=True)
        assert np.abs(f1[c]).max() > 0.0, c
        assert err <= 1e-4, (c, err)  # the FP32 field bound (Hx, a weak scattered component, reaches ~2e-5)
    err = rel_linf(s2.fluxArr()[0].getFlux(), 4.0 * x1)
    print("flux", err, flush=True)
    assert err <= 1e-5
    s2.close()


def test_256_fp32_flux_within_1e5_of_fp64():
    """The whole 1000-step 256^3 run: FP32 fields, FP64 DFT accumulators; the
    flux spectrum stays within 1e-5 of the FP64 run (the FP64 path is pinned
    to the oracle above and on the small decks)."""
    print("256^3 fp32 vs fp64 flux, 1000 steps", flush=True)
    s32 = _run(decks.config2_3d(n=256, steps=1000, precision="fp32"), 1000)
    x32 = s32.fluxArr()[0].getFlux()
    s32.close()
    s64 = _run(decks.config2_3d(n=256, steps=1000, precision="fp64"), 1000)
    x64 = s64.fluxArr()[0].getFlux()
    s64.close()
    err = rel_linf(x32, x64)
    print("flux rel", err, "max|flux|", float(np.abs(x64).max()), flush=True)
    assert np.abs(x64).max() > 0.0
    assert err <= 1e-5


def test_192_six_level_density_invariants():
    """368,640 6-level emitters over 800 steps: every sampled density matrix
    keeps trace 1 and stays Hermitian, and the ground-state population series
    is real and within [0, 1]."""
    print("192^3 6-level layer, 800 steps", flush=True)
    sim = _run(decks.config4_ml3d(n=192, steps=800), 800)
    q = sim.qeArr()[0]
    rng = np.random.default_rng(7)
    moved = 0.0
    for e in rng.choice(192 * 192 * 10, size=40, replace=False):
        rho = q.density(0, int(e), 6)
        tr = np.trace(rho)
        assert abs(tr - 1.0) <= 1e-10, (e, tr)
        assert np.max(np.abs(rho - rho.conj().T)) <= 1e-10, e
        d = np.real(np.diag(rho))
        assert np.all(d >= -1e-10) and np.all(d <= 1.0 + 1e-10), (e, d)
        moved = max(moved, 1.0 - d[0])
    p0 = q.population(0)
    assert np.max(np.abs(p0.imag)) <= 1e-10
    assert np.all(p0.real >= -1e-10) and np.all(p0.real <= 1.0 + 1e-10)
    print("largest ground-state depletion", moved, flush=True)
    assert moved > 0.0  # the field reached the layer
    sim.close()
