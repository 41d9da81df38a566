"""Output files written after a GPU run (FDTDField.toFile) read back against
the series the API returns and, for the flux conversions, the oracle.

Formats: TXT detector src/DTC/parallelDTC_TXT.cpp:14-52, BIN detector
src/DTC/parallelDTC_BIN.cpp:12-53, flux table src/DTC/parallelFlux.hpp:445-534,
population file src/ML/QEPopDtc.cpp:36-61; file names
src/INPUTS/parallelInputs.cpp:650, 680, 760.
"""
import json
import math
import struct

import numpy as np
import pytest

from chiml_amd import decks

pytestmark = pytest.mark.gpu

N, D, STEPS = 48, 0.05, 150
A = decks.A_UNIT


def _deck():
    d = json.loads(decks.config1_2d(n=N, steps=STEPS, pml=8, cyl_radius_cells=6, src_cell=(16, 24), ndet=0,
                                    precision="fp64"))
    p = lambda c: decks._pos(c, N, D)  # noqa: E731
    d["DetectorList"] = [
        {"type": "Ez", "dtc_class": "txt", "fname": "out/pt", "size": [0, 0, 0], "loc": [p(30), p(20), 0.0]},
        {"type": "Ez", "dtc_class": "bin", "fname": "out/box", "size": [2 * D, D, 0], "loc": [p(21), p(30.5), 0.0],
         "SI": False},
        {"type": "Hx", "dtc_class": "txt", "fname": "out/hx", "size": [0, 0, 0], "loc": [p(25), p(26), 0.0],
         "SI": False, "Time_Interval": 3.5 * decks.dt_of(D)},  # floor(3.5) = 3 steps
        {"type": "Ez", "dtc_class": "cout", "fname": "out/quiet", "size": [0, 0, 0], "loc": [p(25), p(26), 0.0]},
    ]
    fc = 0.1
    box = {"size": [11 * D, 11 * D, 0.0], "loc": [p(16), p(24), 0.0], "fcen": fc, "fwidth": fc, "nfreq": 5}
    d["FluxList"] = [dict(box, name="flux/w1"), dict(box, name="flux/w2", weight=2.0),
                     dict(box, name="flux/si", SI=True)]
    return json.dumps(d)


def _rows(path):
    return [line.split("\t") for line in path.read_text().splitlines()]


def test_detector_and_flux_files(tmp_path):
    from chiml_amd import FDTDField
    from oracle.pyoracle import Oracle
    deck = _deck()
    print("2D 48^2 fp64 writers deck", flush=True)
    sim = FDTDField(deck)
    sim.run()
    sim.toFile(tmp_path)
    dets = sim.dtcArr()
    # TXT point detector, SI (default): the corner is the deck location less
    # half a cell on an even grid, z is one cell (2D), and SI scales both by
    # a twice; times are t a / c
    t, v = dets[0].series()
    rows = _rows(tmp_path / "out" / "pt_field_0.dat")
    assert rows[0] == ["# time", "x", "y", "z", "field"]
    assert len(rows) == 1 + t.size
    want_corner = np.array([decks._pos(30, N, D) - D / 2, decks._pos(20, N, D) - D / 2, D]) * A * A
    for s, r in enumerate(rows[1:]):
        assert float(r[0]) == pytest.approx(t[s], rel=1e-5, abs=0.0)
        assert np.allclose([float(x) for x in r[1:4]], want_corner, rtol=1e-5, atol=0.0)
        assert float(r[4]) == v[s, 0]  # 18 digits give the double back
        assert len(r[4]) >= 24
    assert np.abs(v).max() > 0.0
    # BIN box detector: header sz, loc; then t and the values, z, y, x order
    t, v = dets[1].series()
    raw = (tmp_path / "out" / "box_field_1.dat").read_bytes()
    sz = struct.unpack("<3i", raw[:12])
    assert sz == (3, 2, 1)
    body = np.frombuffer(raw[24:], dtype="<f8").reshape(t.size, 1 + 6)
    assert np.array_equal(body[:, 0], t)
    assert np.array_equal(body[:, 1:], v)  # one z plane: z, y, x order equals y, z, x order
    # TXT with a time interval of 3 steps, no SI
    t, v = dets[2].series()
    rows = _rows(tmp_path / "out" / "hx_field_2.dat")
    assert len(rows) == 1 + t.size and t.size == STEPS // 3 + 1  # outputs at steps 0, 3, ..., 150
    assert [float(x) for x in rows[1][1:4]] == pytest.approx(
        [decks._pos(25, N, D) - D / 2, decks._pos(26, N, D) - D / 2, D], rel=1e-5)
    assert np.array_equal(np.array([float(r[4]) for r in rows[1:]]), v[:, 0])
    # COUT detectors write no file
    assert not (tmp_path / "out" / "quiet_field_3.dat").exists()
    # flux tables: faces add up to the total getFlux returns; the weight
    # multiplies once; SI replaces it by the unit conversion (as the oracle)
    fx = [f.getFlux() for f in sim.fluxArr()]
    assert np.abs(fx[0]).max() > 0.0
    assert np.array_equal(fx[1], 2.0 * fx[0])
    orc = Oracle(deck)
    orc.step(STEPS)
    for k in range(3):
        o = orc.flux(k)
        assert np.allclose(fx[k], o, rtol=1e-9, atol=1e-9 * np.abs(o).max()), k
    for k, name in enumerate(["w1", "w2", "si"]):
        rows = _rows(tmp_path / "flux" / f"{name}.dat")
        assert rows[0][0].startswith("#freq")
        assert len(rows) == 1 + fx[k].size
        conv_f = 2.99792458e8 / A if name == "si" else 1.0
        for i, r in enumerate(rows[1:]):
            vals = [float(x) for x in r if x.strip()]
            assert len(vals) == 4 + 3 * 4 + 3  # freq, incident (0), 4 faces, total
            assert vals[1:4] == [0.0, 0.0, 0.0]
            assert vals[0] == pytest.approx(conv_f * 2 * np.pi * (0.1 - 0.05 + i * 0.1 / 4), rel=1e-6)
            faces = np.array(vals[4:16]).reshape(4, 3)
            tot = complex(vals[-2], vals[-1])
            assert tot == pytest.approx(fx[k][i], rel=1e-13, abs=1e-13 * np.abs(fx[k]).max())
            assert complex(faces[:, 1].sum(), faces[:, 2].sum()) == pytest.approx(tot, rel=1e-12,
                                                                                 abs=1e-12 * np.abs(fx[k]).max())
    sim.close()


def test_population_files(tmp_path):
    from chiml_amd import FDTDField
    d = json.loads(decks.config3_ml2d(n=64, steps=200, pml=8, slab=(16, 4), det_cells=0, precision="fp64"))
    d["ObjectList"][-1]["dtc_pop_fname_base"] = "pops/run_"
    d["ObjectList"][-1]["levDTC_timeInt"] = 2
    print("2D 64^2 two-level populations writer", flush=True)
    sim = FDTDField(json.dumps(d))
    sim.run()
    sim.toFile(str(tmp_path))
    q = sim.qeArr()[0]
    for j, lev in enumerate([0, 3]):
        p = q.population(j)
        rows = _rows(tmp_path / "pops" / f"run_0_level_{lev}.dat")
        assert len(rows) == p.size and p.size == (200 + 1) // 2
        for tt, r in enumerate(rows):
            assert float(r[0]) == pytest.approx(tt * 2 * sim.dt, rel=1e-8)
            assert complex(float(r[1]), float(r[2])) == pytest.approx(p[tt], rel=1e-15, abs=1e-15)
            assert float(r[3]) == pytest.approx(abs(p[tt]), rel=1e-15)
    assert q.population(0)[0] == pytest.approx(1.0)
    sim.close()


def test_freq_detector_files(tmp_path):
    """Frequency-detector tables and maps written after a 3D run read back as
    the values the API returns (12 significant digits), with the file names
    <fname>_field_<d>.dat and, for maps, one file per frequency."""
    from chiml_amd import FDTDField
    d = json.loads(decks.config2_3d(n=24, steps=80, pml=4, d=0.02, sphere_cells=4, flux_cells=10, nfreq=4,
                                    precision="fp64"))
    p = lambda c: decks._pos(c, 24, 0.02)  # noqa: E731
    fc = 1.0 / (12 * 0.02)
    d["DetectorList"] = [
        {"type": "Epow", "dtc_class": "freq", "fname": "fq/tab", "size": [0.04, 0.02, 0.0], "loc": [p(12), p(12), p(10)],
         "fcen": fc, "fwidth": fc, "nfreq": 3},
        {"type": "Hy", "dtc_class": "freq", "fname": "fq/map", "size": [0.02, 0.02, 0.02], "loc": [p(12), p(12), p(11)],
         "fcen": fc, "fwidth": fc, "nfreq": 3, "output_map": True},
    ]
    d["FluxList"] = []
    print("3D 24^3 freq detector files", flush=True)
    sim = FDTDField(json.dumps(d))
    sim.run()
    sim.toFile(tmp_path)
    tab, mp = sim.dtcFreqArr()
    fr, ng, npts, nt = tab.info()
    vals = tab.values()
    rows = _rows(tmp_path / "fq" / "tab_field_0.dat")
    assert rows[0][0].startswith("#")
    assert len(rows) == 1 + fr.size
    for k, r in enumerate(rows[1:]):
        x = [float(v) for v in r]
        assert x[0] == pytest.approx(fr[k], rel=1e-11)
        assert x[1:13] == [0.0] * 12  # three incident fields and their total
        for gi in range(ng):
            re, im, ab = x[13 + 3 * gi: 16 + 3 * gi]
            assert re == pytest.approx(vals[gi, k].real, rel=1e-11, abs=1e-300)
            assert im == pytest.approx(vals[gi, k].imag, rel=1e-11, abs=1e-300)
            assert ab == pytest.approx(abs(vals[gi, k]), rel=1e-11, abs=1e-300)
    frm, _, npm, ntm = mp.info()
    F = mp.fields()
    for k in range(frm.size):
        rows = _rows(tmp_path / "fq" / ("map_field_1.dat." + "%f" % frm[k]))
        assert len(rows) == 1 + 8
        for jj, r in enumerate(rows[1:]):
            # H component, SI: conv = I0 / a; value conv/2 F / nt per point
            c = 2.9979245e8  # the reference's speed of light
            conv = 1.0 / (4e-7 * math.pi * c * c) * c  # I0 / a = eps0 c at the default I0
            want = conv / 2.0 * F[0, jj, k] / ntm
            ab, re, im = (float(v) for v in r[6:9])
            assert re == pytest.approx(want.real, rel=1e-9, abs=1e-18)
            assert im == pytest.approx(want.imag, rel=1e-9, abs=1e-18)


def test_flux_save_and_load(tmp_path):
    """A flux box with save writes <name>_fields.dat with the face fields
    (the reference's saveFields layout); a second run that loads it starts
    from those fields negated: on the GPU as in the oracle, and the files of
    the two agree."""
    from chiml_amd import FDTDField
    from oracle.pyoracle import Oracle
    from tests.test_oracle_pins import _save_load_deck, read_flux_fields
    print("flux save / load 30^3", flush=True)
    a = FDTDField(_save_load_deck(save=True))
    a.run()
    a.toFile(tmp_path / "gpu")
    pa = tmp_path / "gpu" / "f_fields.dat"
    oa = Oracle(_save_load_deck(save=True))
    oa.step(60)
    po = tmp_path / "oracle_fields.dat"
    oa.flux_save(0, po)
    nf, sz, fg = read_flux_fields(pa)
    nf_o, sz_o, fo = read_flux_fields(po)
    assert (nf, sz) == (nf_o, sz_o)
    scale = max(np.abs(fo[w][v]).max() for w in range(4) for v in range(6))
    for w in range(4):
        for v in range(6):
            assert fg[w][v].shape == fo[w][v].shape
            assert np.abs(fg[w][v] - fo[w][v]).max() <= 1e-9 * scale, (w, v)
    b = FDTDField(_save_load_deck(sphere=True, load=pa))
    b.run()
    ob = Oracle(_save_load_deck(sphere=True, load=pa))
    ob.step(60)
    got, want = b.fluxArr()[0].getFlux(), ob.flux(0)
    print("scattered flux", np.abs(want).max(), "rel err", np.abs(got - want).max() / np.abs(want).max(), flush=True)
    assert np.abs(want).max() > 0
    assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max()
