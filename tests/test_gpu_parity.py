"""GPU path vs the FP64 oracle on the same decks (parity tests).

FP64 decks must agree to round-off (1e-9 normalised L-inf after hundreds of
steps); FP32 decks to the north-star bound on detector and flux series
(1e-5 relative to the FP64 oracle) where stated.
"""
import json

import numpy as np
import pytest

from chiml_amd import decks
from oracle.pyoracle import Oracle
from tests.helpers import rel_linf

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-9
# FP32 fields, detector series, flux and populations against the FP64 oracle
# (the bound the FP32 path is held to; measured 1e-7..5e-6 on these decks)
FP32_TOL = 1e-5


def _pair(deck, steps):
    from chiml_amd import FDTDField
    sim = FDTDField(deck)
    sim.step(steps)
    sim.sync()
    orc = Oracle(deck)
    orc.step(steps)
    return sim, orc


def _fields(sim, orc, comps, tol):
    for c in comps:
        g = sim.field(c)
        o = orc.field(c)
        err = rel_linf(g, o)
        print(f"field {c}: max|ref|={np.abs(o).max():.3e} rel_linf={err:.3e}", flush=True)
        assert err <= tol, (c, err)


@pytest.mark.parametrize("profile", ["gaussian", "ricker"])
def test_2d_tmz_cylinder_fp64(profile):
    deck = decks.config1_2d(n=64, steps=300, profile=profile, pml=10, cyl_radius_cells=12, src_cell=(20, 32), ndet=4)
    sim, orc = _pair(deck, 300)
    if profile == "ricker":
        for c in (2, 3, 4):
            assert np.all(sim.field(c) == 0.0)
        return
    _fields(sim, orc, (2, 3, 4), FP64_TOL)
    for d in range(orc.num_detectors()):
        t_o, v_o = orc.detector(d)
        t_g, v_g = sim.dtcArr()[d].series()
        assert v_g.shape == v_o.shape
        assert np.allclose(t_g, t_o, rtol=1e-12, atol=0)
        assert rel_linf(v_g, v_o) <= FP64_TOL


@pytest.mark.parametrize("precision,tol", [("fp64", FP64_TOL), ("fp32", FP32_TOL)])
def test_2d_magnetic_cylinder(precision, tol):
    """A cylinder with mu = 2.5 and eps = 4: the Hx/Hy updates take dt/mu from
    the H material map."""
    d = json.loads(decks.config1_2d(n=64, steps=300, pml=10, cyl_radius_cells=12, src_cell=(20, 32), ndet=2))
    d["GPU"]["precision"] = precision
    for o in d["ObjectList"]:
        o["mu"] = 2.5
    print(f"mu 2d precision={precision}", flush=True)
    sim, orc = _pair(json.dumps(d), 300)
    _fields(sim, orc, (2, 3, 4), tol)


@pytest.mark.parametrize("precision,tol", [("fp64", FP64_TOL), ("fp32", FP32_TOL)])
def test_2d_oblique_h_sources(precision, tol):
    """Oblique Hx and Hy sheet sources (per-point time offsets, no dt factor)
    beside the dielectric cylinder; the same adds in the oracle."""
    d = json.loads(decks.config1_2d(n=64, steps=300, pml=10, cyl_radius_cells=12, src_cell=(20, 32), ndet=2))
    d["GPU"]["precision"] = precision
    pul = d["SourceList"][0]["PulseList"]
    d["SourceList"] = [{"pol": "Hx", "size": [0.8, 0.4, 0.0], "loc": [-0.6, 0.3, 0.0], "phi": 30.0, "PulseList": pul},
                       {"pol": "Hy", "size": [0.6, 0.2, 0.0], "loc": [0.7, -0.5, 0.0], "phi": 200.0, "PulseList": pul}]
    deck = json.dumps(d)
    print(f"oblique 2d precision={precision}", flush=True)
    sim, orc = _pair(deck, 300)
    _fields(sim, orc, (2, 3, 4), tol)
    for k in range(orc.num_detectors()):
        assert rel_linf(sim.dtcArr()[k].series()[1], orc.detector(k)[1]) <= tol


@pytest.mark.parametrize("precision,tol", [("fp64", FP64_TOL), ("fp32", FP32_TOL)])
def test_2d_detectors_on_every_component(precision, tol):
    """A small 2D deck whose detectors read Ez, Hx, Hy and E power (2D decks
    this small sample their point detectors inside the E kernel)."""
    d = json.loads(decks.config1_2d(n=64, steps=300, pml=10, cyl_radius_cells=12, src_cell=(20, 32), ndet=2,
                                    precision=precision))
    base = d["DetectorList"][0]
    for typ in ("Hx", "Hy", "Epow"):
        det = dict(base)
        det["type"] = typ
        d["DetectorList"].append(det)
    deck = json.dumps(d)
    sim, orc = _pair(deck, 300)
    assert orc.num_detectors() == 5
    for k in range(orc.num_detectors()):
        t_o, v_o = orc.detector(k)
        t_g, v_g = sim.dtcArr()[k].series()
        assert v_g.shape == v_o.shape, k
        assert np.abs(v_o).max() > 0.0, k
        print(f"{precision} detector {k}: rel_linf={rel_linf(v_g, v_o):.3e}", flush=True)
        assert rel_linf(v_g, v_o) <= tol, (k, rel_linf(v_g, v_o))


def test_3d_aniso_drude_lorentz_flux_fp64():
    deck = decks.config2_3d(n=28, steps=120, pml=6, d=0.02, sphere_cells=6, flux_cells=12, nfreq=8, precision="fp64")
    sim, orc = _pair(deck, 120)
    _fields(sim, orc, range(6), FP64_TOL)
    fg = sim.fluxArr()[0].getFlux()
    fo = orc.flux(0)
    print("flux rel", rel_linf(fg, fo))
    assert rel_linf(fg, fo) <= 1e-8


def test_3d_fp32_close_to_fp64_oracle():
    deck = decks.config2_3d(n=28, steps=120, pml=6, d=0.02, sphere_cells=6, flux_cells=12, nfreq=8, precision="fp32")
    sim, orc = _pair(deck, 120)
    _fields(sim, orc, range(6), FP32_TOL)
    fg = sim.fluxArr()[0].getFlux()
    fo = orc.flux(0)
    err = rel_linf(fg, fo)
    print("fp32 flux rel", err)
    assert err <= 1e-5


def test_2d_ml_two_level_fp64():
    d = json.loads(decks.config3_ml2d(n=64, steps=400, pml=10, slab=(16, 6), precision="fp64", det_cells=2))
    deck = json.dumps(d)
    sim, orc = _pair(deck, 400)
    _fields(sim, orc, (2, 3, 4), FP64_TOL)
    q = sim.qeArr()[0]
    for j in range(2):
        pg = q.population(j)
        po = orc.pops(0, j)
        assert pg.shape == po.shape, (pg.shape, po.shape)
        assert rel_linf(pg, po) <= 1e-10


def _split_deck(variant, precision):
    """Decks whose 3D updates take the interior + CPML-slab kernels (x PML
    thickness and nx multiples of 4), with variants that move cells, sources
    and slabs around."""
    d = json.loads(decks.config2_3d(n=32, steps=100, pml=8, d=0.02, sphere_cells=5, flux_cells=14, nfreq=6,
                                    precision=precision))
    dx = 0.02
    if variant == "wide":  # nx = 40, ny = nz = 32
        d["CompCell"]["size"][0] = 39 * dx
    elif variant == "thinz":  # a 2-cell CPML along z: thin z slabs
        d["PML"]["thickness"][2] = 2 * dx
    elif variant == "pmlsrc":  # extra H point sources in the low-x slab and in a corner
        # (E sources on CPML cells filled with material land on D and are dropped, as in the reference)
        for pol, loc in (("Hy", [-14 * dx, 0.0, 0.0]), ("Hz", [-13 * dx, 13 * dx, -14 * dx])):
            d["SourceList"].append({"pol": pol, "size": [0.0, 0.0, 0.0], "loc": loc,
                                    "PulseList": [decks.gaussian_pulse(5.0, 5.0, 3.0, 1.0)]})
    elif variant == "circ":  # R and L sources (two E components each) and a centred phi = 90.3 source
        pul = [decks.gaussian_pulse(5.0, 5.0, 3.0, 1.0)]
        d["SourceList"] += [
            {"pol": "R", "size": [0.0, 4 * dx, 6 * dx], "loc": [-5 * dx, 2 * dx, 0.0], "psi": 30.0, "PulseList": pul},
            {"pol": "L", "size": [3 * dx, 3 * dx, 0.0], "loc": [4 * dx, -3 * dx, 5 * dx], "PulseList": pul},
            {"pol": "Ez", "size": [2 * dx, 0.0, 0.0], "loc": [0.0, 6 * dx, -6 * dx], "phi": 90.3, "PulseList": pul}]
    elif variant == "oblique":  # oblique Ez and Hz patches (half-cell delayed) and an Ex that adds nothing;
        # the Ez patch stays off the material-filled CPML (its E points there are D-derived)
        pul = [decks.gaussian_pulse(5.0, 5.0, 3.0, 1.0)]
        d["SourceList"] = [
            {"pol": "Ez", "size": [8 * dx, 4 * dx, 0.0], "loc": [2 * dx, 2 * dx, 6 * dx], "phi": 45.0, "PulseList": pul},
            {"pol": "Hz", "size": [6 * dx, 6 * dx, 0.0], "loc": [4 * dx, -2 * dx, 8 * dx], "phi": 60.0, "PulseList": pul},
            {"pol": "Ex", "size": [6 * dx, 6 * dx, 0.0], "loc": [0.0, 0.0, 9 * dx], "phi": 30.0, "PulseList": pul}]
    elif variant == "mu":  # a magnetic dielectric block (mu = 1.8) around the sphere, clear of the CPML:
        # the H map takes the generic update kernels
        d["ObjectList"].append({"shape": "block", "material": "custom", "eps": 1.5, "mu": 1.8, "pols": [],
                                "size": [10 * dx, 12 * dx, 8 * dx], "loc": [2 * dx, -dx, 3 * dx], "Basis_Set": []})
    return json.dumps(d)


@pytest.mark.parametrize("variant", ["cube", "wide", "thinz", "pmlsrc", "circ", "oblique", "mu"])
@pytest.mark.parametrize("precision,tol_f,tol_flux", [("fp64", FP64_TOL, 1e-8), ("fp32", FP32_TOL, FP32_TOL)])
def test_3d_split_kernels(variant, precision, tol_f, tol_flux):
    deck = _split_deck(variant, precision)
    print(f"split variant={variant} precision={precision}", flush=True)
    sim, orc = _pair(deck, 100)
    _fields(sim, orc, range(6), tol_f)
    err = rel_linf(sim.fluxArr()[0].getFlux(), orc.flux(0))
    print("flux rel", err, flush=True)
    assert err <= tol_flux


@pytest.mark.parametrize("variant", ["cube", "wide", "thinz", "pmlsrc"])
def test_3d_fused_step_fp32(variant, monkeypatch):
    """The opt-in fused H + E kernel (CHIML_FUSED=1, read when the simulation
    is created) on the split-kernel decks, against the FP64 oracle and against
    the split sweeps of the same FP32 deck."""
    deck = _split_deck(variant, "fp32")
    print(f"fused variant={variant}", flush=True)
    from chiml_amd import FDTDField
    split = FDTDField(deck)
    split.step(100)
    split.sync()
    monkeypatch.setenv("CHIML_FUSED", "1")
    sim, orc = _pair(deck, 100)
    _fields(sim, orc, range(6), FP32_TOL)
    for c in range(6):
        assert rel_linf(sim.field(c), split.field(c)) <= 1e-5, c
    err = rel_linf(sim.fluxArr()[0].getFlux(), orc.flux(0))
    print("flux rel", err, flush=True)
    assert err <= 1e-5


def _n_level_block(nlev, precision):
    d = decks.empty_deck((20, 20, 20), 0.02, 90, 4, precision=precision)
    em = decks.n_level_emitter(nlev, [0.04, 0.04, 0.04], (0.01, 0.01, 0.01), seed=40 + nlev)
    d["ObjectList"].append(em)
    fc = decks.ev2fdtd(2.0)
    d["SourceList"].append({"pol": "Ez", "size": [0.0, 0.0, 0.0], "loc": [-0.09, 0.01, 0.01],
                            "PulseList": [{"profile": "gaussian", "fcen": fc, "fwidth": fc, "cutoff": 3.0,
                                           "t_0": 0.3, "Field_Intensity": 1e10}]})
    return json.dumps(d)


@pytest.mark.parametrize("nlev", [3, 4, 6])
def test_3d_n_level_emitter_block_fp64(nlev):
    """A 3x3x3 block of N-level emitters (seeded Hermitian H0 and dipoles)
    driven by a point source: fields, populations and one density matrix."""
    deck = _n_level_block(nlev, "fp64")
    print(f"n-level block N={nlev}", flush=True)
    sim, orc = _pair(deck, 90)
    _fields(sim, orc, range(6), FP64_TOL)
    q = sim.qeArr()[0]
    for j in range(2):
        pg, po = q.population(j), orc.pops(0, j)
        assert pg.shape == po.shape
        assert rel_linf(pg, po) <= 1e-10, (j, rel_linf(pg, po))
    rg = q.density(0, 13, nlev)
    ro = orc.density(0, 0, 13, nlev)
    assert rel_linf(rg, ro) <= 1e-10


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_3d_six_level_block_eight_population_detectors(precision):
    """Eight population detectors on one 6-level group: the six diagonals and
    the (1, 0) / (0, 1) coherences.  The first four are summed in the density
    kernel, the rest by the gather kernel (the (1, 0) entry read as the
    conjugate of the stored (0, 1) with Hermitian packing)."""
    d = json.loads(_n_level_block(6, precision))
    levs = [k * 7 for k in range(6)] + [1, 6]
    for o in d["ObjectList"]:
        if "dtc_levs" in o:
            o["dtc_levs"] = levs
    print(f"8 populations precision={precision}", flush=True)
    sim, orc = _pair(json.dumps(d), 90)
    q = sim.qeArr()[0]
    tol = 1e-10 if precision == "fp64" else FP32_TOL
    for j in range(len(levs)):
        pg, po = q.population(j), orc.pops(0, j)
        assert pg.shape == po.shape
        scale = max(np.abs(po).max(), 1e-300)
        err = np.abs(pg - po).max() / scale
        print(f"population {j} (flat {levs[j]}): peak {scale:.3e} err {err:.3e}", flush=True)
        assert err <= tol, (j, err)


@pytest.mark.parametrize("nlev", [2, 4, 6])
def test_3d_n_level_emitter_block_fp32_fields(nlev):
    """The same block with FP32 fields and FP64 density matrices (the 3D
    emitter workload's precision): fields within the FP32 field bound of the
    FP64 oracle, populations within 1e-5 of theirs."""
    deck = _n_level_block(nlev, "fp32")
    print(f"n-level block N={nlev} fp32 fields", flush=True)
    sim, orc = _pair(deck, 90)
    # FP32 rounding is relative to the field's scale: Hz, nearly zero by
    # symmetry here, carries the rounding of Hx and Hy, 1e4 times larger
    for comps in ((0, 1, 2), (3, 4, 5)):
        ref = [orc.field(c) for c in comps]
        scale = max(np.abs(o).max() for o in ref)
        for c, o in zip(comps, ref):
            err = np.abs(sim.field(c) - o).max() / scale
            print(f"field {c}: max|ref|={np.abs(o).max():.3e} err/scale={err:.3e}", flush=True)
            assert err <= FP32_TOL, (c, err)
    q = sim.qeArr()[0]
    for j in range(2):
        pg, po = q.population(j), orc.pops(0, j)
        assert pg.shape == po.shape
        print(f"population {j}: rel_linf={rel_linf(pg, po):.3e}", flush=True)
        assert rel_linf(pg, po) <= FP32_TOL, (j, rel_linf(pg, po))


@pytest.mark.parametrize("dim", [2, 3])
def test_power_detectors_over_several_fields_fp64(dim):
    """Epow / Hpow detectors sum conv/2 (a^2 + b^2) over the E (H) components,
    each with its Yee offset (parallelFDTDField.cpp:796-826; three fields in
    3D, two for Hpow in 2D): interior points, a box, and points on the domain
    edge whose offset cell is the zero ghost beyond it."""
    n, d = (24, 0.05) if dim == 3 else (48, 0.05)
    n3 = (n, n, n) if dim == 3 else (n, n, 1)
    half = (n - 1) * d / 2
    z0 = 0.0
    deck = decks.empty_deck(n3, d, 80, 4, precision="fp64")
    fc = 1.0 / (8 * d)
    for pol, loc in (("Ez", [d, 0.0, z0]), ("Ex", [-d, 2 * d, z0]) if dim == 3 else ("Ez", [-3 * d, 2 * d, z0])):
        deck["SourceList"].append({"pol": pol, "size": [0, 0, 0], "loc": loc,
                                   "PulseList": [decks.gaussian_pulse(fc, fc, 3.0, 1.0)]})
    sz_box = [2 * d, d, d if dim == 3 else 0.0]
    for t in ("Epow", "Hpow"):
        for loc, size in (([2 * d, -d, z0], [0, 0, 0]), ([0.0, 3 * d, z0], sz_box),
                          ([-half, 0.0, z0], [0, 0, 0]), ([half, 0.0, z0], [0, 0, 0])):
            deck["DetectorList"].append({"type": t, "dtc_class": "cout", "fname": "p", "SI": False,
                                         "size": size, "loc": loc})
    deck = json.dumps(deck)
    print(f"power detectors {dim}D", flush=True)
    sim, orc = _pair(deck, 80)
    for k, dtc in enumerate(sim.dtcArr()):
        tg, vg = dtc.series()
        to, vo = orc.detector(k)
        assert vg.shape == vo.shape, (k, vg.shape, vo.shape)
        assert np.array_equal(tg, to), k
        assert np.abs(vo).max() > 0.0, k
        err = rel_linf(vg, vo)
        print(f"detector {k}: max|ref|={np.abs(vo).max():.3e} rel_linf={err:.3e}", flush=True)
        assert err <= FP64_TOL, (k, err)


def test_3d_hemisphere_hemiellipsoid_and_cut_ellipsoid_fp64():
    """The half shapes, ellipsoid cuts, a paraboloid and a torus section through the whole product path
    (material ids, dispersive boxes, split kernels) against the oracle."""
    d = decks.empty_deck((24, 24, 24), 0.02, 80, 4, precision="fp64")
    base = {"material": "custom", "Basis_Set": [],
            "unit_vectors": [{"uvec": [0.8, 0.6, 0.0]}, {"uvec": [-0.6, 0.8, 0.0]}, {"uvec": [0.0, 0.0, 1.0]}]}
    d["ObjectList"].append(dict(base, shape="hemisphere", eps=3.0, pols=[], radius=0.11, loc=[-0.07, 0.05, 0.0]))
    d["ObjectList"].append(dict(base, shape="hemiellipsoid", eps=1.0, pols=decks.gold_like_pols()[:2],
                                size=[0.16, 0.12, 0.1], loc=[0.08, -0.06, 0.03]))
    d["ObjectList"].append(dict(base, shape="ellipsoid", eps=2.2, pols=[], size=[0.18, 0.1, 0.14],
                                loc=[0.02, 0.1, -0.08], cut_pos=[0.6, 1.0, 1.0], cut_global=[1.0, 1.0, 0.7]))
    d["ObjectList"].append(dict(base, shape="paraboloid", eps=2.6, pols=[], a=0.05, b=0.04, length=0.12,
                                loc=[-0.1, -0.1, 0.1]))
    d["ObjectList"].append(dict(base, shape="torus", eps=1.8, pols=[], rad_Ring=0.12, rad_CrossSec=0.03,
                                phiMin=30.0, phiMax=250.0, loc=[0.0, 0.0, -0.02]))
    fc = 1.0 / (8 * 0.02)
    d["SourceList"].append({"pol": "Ez", "size": [0.0, 0.0, 0.0], "loc": [0.0, 0.0, 0.05],
                            "PulseList": [decks.gaussian_pulse(fc, fc, 3.0, 1.0)]})
    deck = json.dumps(d)
    print("half shapes and cut ellipsoid", flush=True)
    sim, orc = _pair(deck, 80)
    _fields(sim, orc, range(6), FP64_TOL)


@pytest.mark.parametrize("precision,tol", [("fp64", FP64_TOL), ("fp32", FP32_TOL)])
def test_3d_cone_ters_tip_prisms_tetrahedron(precision, tol):
    """Rotated cone (dispersive), TERS tip, triangular and trapezoidal prisms and a dispersive
    tetrahedron through the product path against the oracle."""
    d = decks.empty_deck((24, 24, 24), 0.02, 80, 4, precision=precision)
    base = {"material": "custom", "Basis_Set": [],
            "unit_vectors": [{"uvec": [0.8, 0.6, 0.0]}, {"uvec": [-0.6, 0.8, 0.0]}, {"uvec": [0.0, 0.0, 1.0]}]}
    d["ObjectList"].append(dict(base, shape="cone", eps=1.0, pols=decks.gold_like_pols()[:2], radSmall=0.03,
                                radLarge=0.09, height=0.12, loc=[-0.06, 0.05, 0.02]))
    d["ObjectList"].append(dict(base, shape="ters_tip", eps=2.4, pols=[], rad_curve=0.04, base=0.16, length=0.2,
                                loc=[0.07, -0.05, -0.01]))
    d["ObjectList"].append(dict(base, shape="triangle_prism", eps=1.9, pols=[], base=0.12, height=0.1, length=0.08,
                                loc=[0.1, 0.1, 0.1]))
    d["ObjectList"].append(dict(base, shape="trapezoid_prism", eps=2.1, pols=[], base1=0.14, base2=0.06, height=0.06,
                                length=0.08, loc=[-0.12, -0.1, -0.12]))
    d["ObjectList"].append(dict(base, shape="tetrahedron", eps=1.0, pols=decks.gold_like_pols()[:2], sideLen=0.14,
                                loc=[0.11, -0.12, 0.11]))
    fc = 1.0 / (8 * 0.02)
    d["SourceList"].append({"pol": "Ez", "size": [0.0, 0.0, 0.0], "loc": [0.0, 0.0, 0.05],
                            "PulseList": [decks.gaussian_pulse(fc, fc, 3.0, 1.0)]})
    print(f"cone and TERS tip precision={precision}", flush=True)
    sim, orc = _pair(json.dumps(d), 80)
    # each component against its family's peak: Hz is 1% of Hx/Hy here, so its own peak
    # would hold FP32 rounding of the larger components to a 100x tighter bound
    for fam in ((0, 1, 2), (3, 4, 5)):
        peak = max(np.abs(orc.field(c)).max() for c in fam)
        for c in fam:
            err = np.abs(sim.field(c) - orc.field(c)).max() / peak
            print(f"field {c}: family peak {peak:.3e} err {err:.3e}", flush=True)
            assert err <= tol, (c, err)


def test_3d_basis_set_emitters_x_y_dipoles_fp64():
    """A 3x3x3 block of 4-level emitters whose dipoles come from the
    spherical-harmonic basis {(0,0), (1,-1), (1,0), (1,1)} (x and y matrix
    elements between m = +-1 states), driven by Ex and Ey point sources: the
    product's own basis-dipole construction against the oracle's, which is
    pinned to quadrature."""
    nlev = 4
    d = decks.empty_deck((20, 20, 20), 0.02, 90, 4, precision="fp64")
    em = decks.n_level_emitter(nlev, [0.04, 0.04, 0.04], (0.01, 0.01, 0.01), seed=77)
    del em["H0_eV"], em["dipole_Debye"]
    em["Basis_Set"] = [{"l": 0, "m": 0}, {"l": 1, "m": -1}, {"l": 1, "m": 0}, {"l": 1, "m": 1}]
    em["couplings"] = [12.0] * (nlev * nlev)
    d["ObjectList"].append(em)
    fc = decks.ev2fdtd(2.0)
    pul = {"profile": "gaussian", "fcen": fc, "fwidth": fc, "cutoff": 3.0, "t_0": 0.3, "Field_Intensity": 1e10}
    d["SourceList"].append({"pol": "Ex", "size": [0.0, 0.0, 0.0], "loc": [-0.09, 0.01, 0.01], "PulseList": [pul]})
    d["SourceList"].append({"pol": "Ey", "size": [0.0, 0.0, 0.0], "loc": [0.01, -0.09, 0.03], "PulseList": [pul]})
    deck = json.dumps(d)
    sim, orc = _pair(deck, 90)
    _fields(sim, orc, range(6), FP64_TOL)
    q = sim.qeArr()[0]
    for j in range(2):
        pg, po = q.population(j), orc.pops(0, j)
        assert pg.shape == po.shape
        assert rel_linf(pg, po) <= 1e-10, (j, rel_linf(pg, po))
    rg = q.density(0, 13, nlev)
    ro = orc.density(0, 0, 13, nlev)
    assert np.max(np.abs(ro - np.diag(np.diag(ro)))) > 0.0  # coherences were driven
    assert rel_linf(rg, ro) <= 1e-10


def test_2d_tmz_fp32_close_to_fp64_oracle():
    """FP32 fields against the FP64 oracle: detector series within 1e-5 of the
    series' peak, fields within 1e-5."""
    deck = decks.config1_2d(n=64, steps=300, profile="gaussian", pml=8, cyl_radius_cells=12, src_cell=(20, 32),
                            ndet=4, precision="fp32")
    sim, orc = _pair(deck, 300)
    _fields(sim, orc, (2, 3, 4), FP32_TOL)
    for d in range(orc.num_detectors()):
        v_o = orc.detector(d)[1]
        v_g = sim.dtcArr()[d].series()[1]
        err = rel_linf(v_g, v_o)
        print(f"detector {d}: rel_linf={err:.3e}", flush=True)
        assert err <= 1e-5


def _freq_deck(precision, dim=3):
    """A small 3D deck (aniso background, dispersive sphere, plane-wave sheet)
    with frequency detectors of every kind the path reads: a field over a box,
    the E and H powers, a wavelength list and a sampling interval of 2 steps;
    a point detector listed after them keeps its own index."""
    if dim == 3:
        d = json.loads(decks.config2_3d(n=28, steps=160, pml=6, d=0.02, sphere_cells=5, flux_cells=12, nfreq=6,
                                        precision=precision))
        n, dd = 28, 0.02
    else:
        d = json.loads(decks.config1_2d(n=48, steps=160, pml=8, cyl_radius_cells=6, src_cell=(16, 24), ndet=0,
                                        precision=precision))
        n, dd = 48, 0.05
    p = lambda c: decks._pos(c, n, dd)  # noqa: E731
    fc = 1.0 / (12 * dd)
    z = p(17) if dim == 3 else 0.0
    d["DetectorList"] = [
        {"type": "Ex" if dim == 3 else "Ez", "dtc_class": "freq", "fname": "fq/e", "size": [3 * dd, 2 * dd, dd if dim == 3 else 0.0],
         "loc": [p(15), p(12), z], "fcen": fc, "fwidth": fc, "nfreq": 5},
        {"type": "Epow", "dtc_class": "freq", "fname": "fq/ep", "size": [2 * dd, 2 * dd, 0.0], "loc": [p(13), p(15), z],
         "lamL": 1.0 / (1.4 * fc), "lamR": 1.0 / (0.7 * fc), "nfreq": 4, "Time_Interval": 2 * decks.dt_of(dd)},
        {"type": "Hpow", "dtc_class": "freq", "fname": "fq/hp", "size": [dd, 3 * dd, 0.0], "loc": [p(14), p(14), z],
         "fcen": fc, "fwidth": 0.5 * fc, "nfreq": 3, "SI": False},
        {"type": "Ez", "dtc_class": "cout", "fname": "pt", "size": [0, 0, 0], "loc": [p(15), p(13), z]},
    ]
    return json.dumps(d)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_freq_detectors_3d(precision):
    """Running DFTs of the frequency detectors (parallelDetectorFREQ_Base::output,
    src/DTC/parallelDTC_FREQ.hpp:258-263) on the GPU against the oracle: the
    accumulators per grid, point and frequency, the samples taken, and the
    table's values; FP32 fields keep FP64 accumulators."""
    deck = _freq_deck(precision)
    print(f"freq detectors 3D {precision}", flush=True)
    sim, orc = _pair(deck, 160)
    tol = FP64_TOL if precision == "fp64" else FP32_TOL
    fds = sim.dtcFreqArr()
    assert len(fds) == orc.num_freq_detectors() == 3
    for k, fd in enumerate(fds):
        fr, ng, npts, nt = fd.info()
        nf, ng_o, npts_o, nt_o = orc.freq_info(k)
        assert (fr.size, ng, npts, nt) == (nf, ng_o, npts_o, nt_o), k
        fr_o, vals_o = orc.freq_values(k)
        np.testing.assert_allclose(fr, fr_o, rtol=1e-15)
        F, F_o = fd.fields(), orc.freq_fields(k)
        assert np.abs(F_o).max() > 0.0
        err_f = rel_linf(F, F_o)
        err_v = rel_linf(fd.values(), vals_o)
        print(f"freq detector {k}: {ng} grids x {npts} points x {nf} freqs, nt {nt}: fields {err_f:.3e}, values {err_v:.3e}",
              flush=True)
        assert err_f <= tol and err_v <= tol, (k, err_f, err_v)
    # the point detector after them is point detector 0
    t_o, v_o = orc.detector(0)
    t_g, v_g = sim.dtcArr()[0].series()
    assert rel_linf(v_g, v_o) <= tol


def test_eleven_freq_detectors_3d():
    """Eleven frequency detectors (twelve DFT groups with the flux boxes):
    the three detectors of the deck above, repeated at shifted points."""
    d = json.loads(_freq_deck("fp64"))
    fq = [x for x in d["DetectorList"] if x["dtc_class"] == "freq"]
    extra = []
    for r in range(1, 4):
        for x in fq:
            y = json.loads(json.dumps(x))
            y["fname"] = f"{x['fname']}{r}"
            y["loc"] = [y["loc"][0] + 0.02 * r, y["loc"][1] - 0.02 * (r % 2), y["loc"][2]]
            extra.append(y)
    d["DetectorList"] = fq + extra[:8] + [x for x in d["DetectorList"] if x["dtc_class"] != "freq"]
    print("eleven freq detectors", flush=True)
    sim, orc = _pair(json.dumps(d), 160)
    fds = sim.dtcFreqArr()
    assert len(fds) == orc.num_freq_detectors() == 11
    for k, fd in enumerate(fds):
        F, F_o = fd.fields(), orc.freq_fields(k)
        err_f = rel_linf(F, F_o)
        err_v = rel_linf(fd.values(), orc.freq_values(k)[1])
        print(f"freq detector {k}: fields {err_f:.3e}, values {err_v:.3e}", flush=True)
        assert err_f <= FP64_TOL and err_v <= FP64_TOL, (k, err_f, err_v)


def test_freq_detector_2d_is_zero():
    """In 2D the reference's frequency storage samples no point (its z extent
    is 0), so the table is zero; the product keeps that."""
    deck = _freq_deck("fp64", dim=2)
    sim, orc = _pair(deck, 160)
    for k, fd in enumerate(sim.dtcFreqArr()):
        fr, ng, npts, nt = fd.info()
        assert npts == 0 and orc.freq_info(k)[2] == 0
        assert np.all(fd.values() == 0.0)
    t_o, v_o = orc.detector(0)
    t_g, v_g = sim.dtcArr()[0].series()
    assert np.abs(v_o).max() > 0.0 and rel_linf(v_g, v_o) <= FP64_TOL
