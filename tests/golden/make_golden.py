"""Writes the golden fixtures: small decks and the FP64 oracle's outputs.

Each case is a JSON deck (chiml_amd.decks, SplitMix64 seeds recorded in the
deck) plus raw little-endian FP64 arrays listed in manifest.json.  The CPU
tests check that the oracle still reproduces them; the GPU tests check the
HIP path against them.  Run from the repository root:

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from chiml_amd import decks  # noqa: E402
from oracle.pyoracle import Oracle  # noqa: E402


def cases():
    """(name, deck json, steps, outputs) -- outputs name the arrays to store."""
    out = []
    for profile in ("gaussian", "ricker"):
        out.append((f"tmz32_cyl_{profile}",
                    decks.config1_2d(n=32, steps=50, profile=profile, pml=6, cyl_radius_cells=5, src_cell=(10, 16),
                                     ndet=3, seed=1),
                    50, {"fields": (2, 3, 4), "detectors": True}))
    d = json.loads(decks.config1_2d(n=32, steps=50, profile="gaussian", pml=6, cyl_radius_cells=5, src_cell=(10, 16),
                                    ndet=3, seed=1))
    d["ObjectList"] = []
    out.append(("tmz32_vacuum_gaussian", json.dumps(d), 50, {"fields": (2, 3, 4), "detectors": True}))
    out.append(("yee16_sphere_dft",
                decks.config2_3d(n=16, steps=60, pml=4, d=0.02, sphere_cells=4, flux_cells=8, nfreq=8,
                                 precision="fp64"),
                60, {"fields": range(6), "flux": True}))
    d = json.loads(decks.config3_ml2d(n=48, steps=200, pml=6, slab=(8, 2), precision="fp64", det_cells=2, seed=3))
    fc = decks.ev2fdtd(2.0)  # a short, strong resonant pulse that reaches the slab within 200 steps
    d["SourceList"][0]["PulseList"] = [{"profile": "gaussian", "fcen": fc, "fwidth": fc, "cutoff": 3.0, "t_0": 0.5,
                                        "Field_Intensity": 1e10}]
    out.append(("ml2_slab8x2", json.dumps(d), 200, {"fields": (2,), "pops": 2, "density": (2, 16)}))
    d = decks.empty_deck((16, 16, 16), 0.02, 80, 4, precision="fp64")
    em = decks.n_level_emitter(6, [0.0, 0.0, 0.0], (0.01, 0.01, 0.01), seed=4)  # on a grid point
    d["ObjectList"].append(em)
    d["SourceList"].append({"pol": "Ez", "size": [0.0, 0.0, 0.0], "loc": [-0.06, 0.0, 0.0],
                            "PulseList": [decks.gaussian_pulse(decks.A_UNIT / 600e-9, 2 * decks.A_UNIT / 600e-9,
                                                               3.0, 1e9)]})
    out.append(("ml6_single", json.dumps(d), 80, {"fields": (2,), "pops": 2, "density": (6, 1)}))
    return out


def run_case(deck, steps, want):
    o = Oracle(deck)
    o.step(steps)
    arrays = {}
    for c in want.get("fields", ()):
        arrays[f"field{c}"] = o.field(c)
    if want.get("detectors"):
        for k in range(o.num_detectors()):
            t, v = o.detector(k)
            arrays[f"det{k}_t"] = t
            arrays[f"det{k}_v"] = v
    if want.get("flux"):
        f = o.flux(0)
        arrays["flux0"] = np.stack([f.real, f.imag], axis=1)
    for j in range(want.get("pops", 0)):
        p = o.pops(0, j)
        arrays[f"pop{j}"] = np.stack([p.real, p.imag], axis=1)
    if "density" in want:
        n, ne = want["density"]
        for e in range(ne):
            r = o.density(0, 0, e, n)
            arrays[f"rho{e}"] = np.stack([r.real, r.imag], axis=-1)
    return arrays


def main():
    manifest = {"format": "raw little-endian float64, C order", "cases": {}}
    for name, deck, steps, want in cases():
        cdir = os.path.join(HERE, name)
        os.makedirs(cdir, exist_ok=True)
        with open(os.path.join(cdir, "deck.json"), "w") as f:
            f.write(deck)
        arrays = run_case(deck, steps, want)
        entry = {"steps": steps, "arrays": {}}
        for k, a in arrays.items():
            a = np.ascontiguousarray(a, dtype="<f8")
            a.tofile(os.path.join(cdir, k + ".bin"))
            entry["arrays"][k] = {"shape": list(a.shape), "sha256": hashlib.sha256(a.tobytes()).hexdigest()}
        manifest["cases"][name] = entry
        print(name, {k: v["shape"] for k, v in entry["arrays"].items()})
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
