"""The benchmark configurations at their stated sizes (or, where the oracle
cannot finish the stated size in seconds, geometrically scaled) against the
FP64 oracle, plus whole-run properties where only the GPU finishes.

- 2D TMz, 256^2, r = 40 cylinder, 2000 steps: FP64 at 1e-9 on fields and
  detector series, FP32 detector series at 1e-5 of each series' peak, and a
  Ricker source that leaves every field and detector exactly zero.
- 3D Yee + CPML + Drude/Lorentz sphere + flux box scaled to 128^3 (every
  length in cells halved, d doubled, so the sheet, sphere and flux box keep
  their physical places), 400 steps: the pulse crosses the flux box and the
  sphere.  FP64 at 1e-9; FP32 flux and detectors at 1e-5.
- 2D Maxwell-Liouville at 160^2 with a 64 x 16 slab of 2-level emitters and
  the pulse moved forward to end inside 8000 steps: FP64 at 1e-9, FP32 fields
  with FP64 densities at 1e-5 on the transmitted Ez series and the
  populations.  The stated 1024^2 x 20,000-step run keeps trace 1,
  hermiticity and populations in [0, 1].
"""
import json
import os

import numpy as np
import pytest

from chiml_amd import decks
from oracle.pyoracle import Oracle, set_threads
from tests.helpers import rel_linf

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-9
FP32_TOL = 1e-5


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def _gpu(deck, steps):
    from chiml_amd import FDTDField
    sim = FDTDField(deck)
    sim.step(steps)
    sim.sync()
    return sim


def _oracle(deck, steps):
    set_threads(_threads())
    o = Oracle(deck)
    o.step(steps)
    return o


def _with_precision(deck_json, precision):
    d = json.loads(deck_json)
    d["GPU"]["precision"] = precision
    return json.dumps(d)


def _detectors(sim, orc, tol, label):
    worst = 0.0
    for k in range(orc.num_detectors()):
        t_o, v_o = orc.detector(k)
        t_g, v_g = sim.dtcArr()[k].series()
        assert v_g.shape == v_o.shape, (label, k)
        assert np.allclose(t_g, t_o, rtol=1e-12, atol=0), (label, k)
        assert np.abs(v_o).max() > 0.0, (label, k)
        err = rel_linf(v_g, v_o)
        worst = max(worst, err)
        assert err <= tol, (label, k, err)
    print(f"{label}: {orc.num_detectors()} detector series, worst rel_linf {worst:.3e}", flush=True)


def _fields(sim, orc, comps, tol, label):
    for c in comps:
        err = rel_linf(sim.field(c), orc.field(c))
        print(f"{label}: field {c} rel_linf {err:.3e}", flush=True)
        assert err <= tol, (label, c, err)


# ------------------------------------------------------------ config 1 ------
@pytest.fixture(scope="module")
def config1_oracle():
    deck = decks.config1_2d()
    print("config 1 oracle: 256^2, 2000 steps", flush=True)
    return deck, _oracle(deck, 2000)


def test_config1_full_size_fp64(config1_oracle):
    deck, orc = config1_oracle
    sim = _gpu(deck, 2000)
    _fields(sim, orc, (2, 3, 4), FP64_TOL, "config 1 fp64")
    _detectors(sim, orc, FP64_TOL, "config 1 fp64")
    sim.close()


def test_config1_full_size_fp32_detectors(config1_oracle):
    deck, orc = config1_oracle
    sim = _gpu(_with_precision(deck, "fp32"), 2000)
    _detectors(sim, orc, FP32_TOL, "config 1 fp32")
    sim.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_config1_full_size_ricker_is_zero(precision):
    """The reference's Ricker turn-off test is false for every t >= 0, so the
    2000-step run leaves every field and detector sample exactly zero."""
    sim = _gpu(decks.config1_2d(profile="ricker", precision=precision), 2000)
    for c in (2, 3, 4):
        assert np.all(sim.field(c) == 0.0), c
    for det in sim.dtcArr():
        t, v = det.series()
        assert v.size >= 2000
        assert np.all(v == 0.0)
    sim.close()


# ------------------------------------------------------------ config 2 ------
def _config2_128(precision):
    return decks.config2_3d(n=128, steps=400, pml=8, d=0.04, sphere_cells=20, flux_cells=60, nfreq=40,
                            precision=precision, ndet=12)


@pytest.fixture(scope="module")
def config2_oracle():
    print("config 2 oracle: 128^3 scaled deck, 400 steps", flush=True)
    return _oracle(_config2_128("fp64"), 400)


def test_config2_scaled_fp64(config2_oracle):
    orc = config2_oracle
    sim = _gpu(_config2_128("fp64"), 400)
    _fields(sim, orc, range(6), FP64_TOL, "config 2 fp64")
    _detectors(sim, orc, FP64_TOL, "config 2 fp64")
    fo = orc.flux(0)
    fg = sim.fluxArr()[0].getFlux()
    assert np.abs(fo).max() > 0.0
    err = rel_linf(fg, fo)
    print(f"config 2 fp64: flux rel_linf {err:.3e}", flush=True)
    assert err <= FP64_TOL
    sim.close()


def _family_detectors(sim, orc, tol, label):
    """Detectors k = 6 j + c read component c (Ex..Hz) at location j.  FP32
    rounding is relative to the field's magnitude, so a component that is
    nearly zero there (Ey and Hx on the sheet's axis, set by scattering only)
    carries the rounding of the strong ones: each series' error is scaled by
    the largest peak of its family (E or H) at the same location.  The
    error relative to its own peak is printed beside it."""
    n = orc.num_detectors()
    ser = [(orc.detector(k)[1], sim.dtcArr()[k].series()[1]) for k in range(n)]
    for k, (vo, vg) in enumerate(ser):
        assert vg.shape == vo.shape, (label, k)
        fam = [j for j in range(n) if j // 6 == k // 6 and (j % 6 < 3) == (k % 6 < 3)]
        scale = max(float(np.abs(ser[j][0]).max()) for j in fam)
        err = float(np.abs(vg - vo).max()) / scale
        print(f"{label}: detector {k} (comp {k % 6}) err/family peak {err:.3e}, "
              f"err/own peak {rel_linf(vg, vo):.3e}", flush=True)
        assert err <= tol, (label, k, err)


def test_config2_scaled_fp32_flux_and_detectors(config2_oracle):
    orc = config2_oracle
    sim = _gpu(_config2_128("fp32"), 400)
    _family_detectors(sim, orc, FP32_TOL, "config 2 fp32")
    err = rel_linf(sim.fluxArr()[0].getFlux(), orc.flux(0))
    print(f"config 2 fp32: flux rel_linf {err:.3e}", flush=True)
    assert err <= FP32_TOL
    sim.close()


# ------------------------------------------------------------ config 3 ------
def _config3_small(precision):
    d = json.loads(decks.config3_ml2d(n=160, steps=8000, pml=20, slab=(64, 16), precision=precision, det_cells=8))
    for p in d["SourceList"][0]["PulseList"]:
        p["cutoff"] = 2.5
        p["t_0"] = 2.5 / p["fwidth"]  # the pulse ends inside the run
    return json.dumps(d)


@pytest.fixture(scope="module")
def config3_oracle():
    print("config 3 oracle: 160^2, 64 x 16 emitters, 8000 steps", flush=True)
    return _oracle(_config3_small("fp64"), 8000)


def _pops(sim, orc, tol, label):
    q = sim.qeArr()[0]
    for j in range(2):
        pg, po = q.population(j), orc.pops(0, j)
        assert pg.shape == po.shape
        err = rel_linf(pg, po)
        print(f"{label}: population {j} rel_linf {err:.3e}", flush=True)
        assert err <= tol, (label, j, err)
    po = orc.pops(0, 1).real
    assert po.max() > 1e-7  # the pulse moved population (8.5e-7 at the peak)


def test_config3_small_fp64(config3_oracle):
    orc = config3_oracle
    sim = _gpu(_config3_small("fp64"), 8000)
    _fields(sim, orc, (2, 3, 4), FP64_TOL, "config 3 fp64")
    _detectors(sim, orc, FP64_TOL, "config 3 fp64")
    _pops(sim, orc, FP64_TOL, "config 3 fp64")
    sim.close()


def test_config3_small_fp32(config3_oracle):
    orc = config3_oracle
    sim = _gpu(_config3_small("fp32"), 8000)
    _detectors(sim, orc, FP32_TOL, "config 3 fp32")
    _pops(sim, orc, FP32_TOL, "config 3 fp32")
    sim.close()


def test_config3_full_size_density_invariants():
    """1024^2, 512 x 64 emitters, 20,000 steps in FP32 fields: sampled density
    matrices keep trace 1 and stay Hermitian with populations in [0, 1], and
    the population series stay real, in [0, 1] and sum to 1."""
    print("config 3 full size: 1024^2, 20000 steps", flush=True)
    sim = _gpu(decks.config3_ml2d(), 20000)
    q = sim.qeArr()[0]
    rng = np.random.default_rng(11)
    moved = 0.0
    for e in rng.choice(512 * 64, size=48, replace=False):
        rho = q.density(0, int(e), 2)
        assert abs(np.trace(rho) - 1.0) <= 1e-10, e
        assert np.max(np.abs(rho - rho.conj().T)) <= 1e-10, e
        d = np.real(np.diag(rho))
        assert np.all(d >= -1e-10) and np.all(d <= 1.0 + 1e-10), (e, d)
        moved = max(moved, d[1])
    p0, p1 = q.population(0), q.population(1)
    assert p0.size == 20000 and p1.size == 20000
    for p in (p0, p1):
        assert np.max(np.abs(p.imag)) <= 1e-10
        assert np.all(p.real >= -1e-10) and np.all(p.real <= 1.0 + 1e-10)
    assert np.max(np.abs(p0.real + p1.real - 1.0)) <= 1e-10
    print(f"config 3 full size: largest excited population {moved:.3e}, series max {p1.real.max():.3e}", flush=True)
    assert p1.real.max() > 0.0
    sim.close()
