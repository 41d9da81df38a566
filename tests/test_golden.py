"""Golden fixtures (tests/golden): the oracle must keep reproducing them (CPU)
and the HIP path must match them (GPU, FP64 decks: 1e-9 normalised L-inf)."""
import json
import os

import numpy as np
import pytest

from tests.helpers import rel_linf

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(HERE, "manifest.json")) as _f:
    MANIFEST = json.load(_f)
CASES = sorted(MANIFEST["cases"])


def _load(case):
    entry = MANIFEST["cases"][case]
    with open(os.path.join(HERE, case, "deck.json")) as f:
        deck = f.read()
    arrays = {k: np.fromfile(os.path.join(HERE, case, k + ".bin"), dtype="<f8").reshape(v["shape"])
              for k, v in entry["arrays"].items()}
    return deck, entry["steps"], arrays


def _outputs(sim_like, arrays, is_oracle):
    """The arrays of a fixture, read from the oracle or from the device simulation."""
    out = {}
    for k in arrays:
        if k.startswith("field"):
            out[k] = sim_like.field(int(k[5:]))
        elif k.startswith("det"):
            d = int(k[3:k.index("_")])
            t, v = sim_like.detector(d) if is_oracle else sim_like.dtcArr()[d].series()
            out[k] = t if k.endswith("_t") else v
        elif k == "flux0":
            f = sim_like.flux(0) if is_oracle else sim_like.fluxArr()[0].getFlux()
            out[k] = np.stack([f.real, f.imag], axis=1)
        elif k.startswith("pop"):
            j = int(k[3:])
            p = sim_like.pops(0, j) if is_oracle else sim_like.qeArr()[0].population(j)
            out[k] = np.stack([p.real, p.imag], axis=1)
        elif k.startswith("rho"):
            e = int(k[3:])
            n = arrays[k].shape[0]
            r = sim_like.density(0, 0, e, n) if is_oracle else sim_like.qeArr()[0].density(0, e, n)
            out[k] = np.stack([r.real, r.imag], axis=-1)
    return out


@pytest.mark.parametrize("case", CASES)
def test_oracle_reproduces_golden(case):
    from oracle.pyoracle import Oracle

    deck, steps, arrays = _load(case)
    o = Oracle(deck)
    o.step(steps)
    got = _outputs(o, arrays, True)
    for k, ref in arrays.items():
        assert got[k].shape == ref.shape, k
        assert rel_linf(got[k], ref) <= 1e-13, (k, rel_linf(got[k], ref))


def test_golden_cases_are_nontrivial():
    for case in CASES:
        _, _, arrays = _load(case)
        if "ricker" in case:
            assert all(np.all(a == 0.0) for k, a in arrays.items() if k.startswith("field"))
        else:
            assert any(np.max(np.abs(a)) > 0 for k, a in arrays.items() if k.startswith("field"))


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_gpu_matches_golden(case):
    from chiml_amd import FDTDField

    deck, steps, arrays = _load(case)
    print(f"golden case {case}", flush=True)
    sim = FDTDField(deck)
    sim.step(steps)
    sim.sync()
    got = _outputs(sim, arrays, False)
    for k, ref in arrays.items():
        assert got[k].shape == ref.shape, (k, got[k].shape, ref.shape)
        tol = 1e-12 if k.endswith("_t") else 1e-9
        err = rel_linf(got[k], ref)
        print(f"  {k}: rel_linf={err:.3e}", flush=True)
        assert err <= tol, (k, err)
