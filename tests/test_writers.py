"""Output writers against the reference's formats, on the CPU.

The reference writes with iostreams: default float notation at the stream's
precision (C's %g), setw right-aligned.  The expected bytes here come from
Python's own float formatting of the same values (an implementation of C's
%g independent of the C++ library) and from struct for the binary layout.
Formats: TXT detector src/DTC/parallelDTC_TXT.cpp:14-52, BIN detector
src/DTC/parallelDTC_BIN.cpp:12-53, flux table src/DTC/parallelFlux.hpp:445-534,
population file src/ML/QEPopDtc.cpp:36-61.
"""
import struct

import numpy as np
import pytest

import chiml_amd

# values that exercise %g: integers, trailing zeros, exponents both ways,
# negative zero, subnormals, long mantissas
EDGE = np.array([0.0, -0.0, 1.0, -2.5, 1.5e-7, 3.0e20, 123456789.123456789, -9.87654321e-300, 5e-324,
                 0.1, 1.0 / 3.0, -2.0 / 3.0, 1e16, 1e-5, 1e-4, 100000.0, 1000000.0])


def g(v, p, w=0):
    return ("%.*g" % (p, v)).rjust(w)


def test_detector_txt_bytes(tmp_path):
    rng = np.random.default_rng(3)
    ns, npts = 5, 7
    t = np.array([0.0, 1.25e-16, 2.5e-16, 3.75e-16, 5e-16])
    v = rng.standard_normal((ns, npts)) * 10.0 ** rng.integers(-30, 30, (ns, npts))
    v[0, : min(npts, EDGE.size)] = EDGE[:npts]
    corner = np.array([-1.2e-7, 4e-8, 1e-9])
    p = tmp_path / "d.dat"
    chiml_amd.write_detector_txt(p, t, v, corner)
    want = "# time\tx\ty\tz\tfield\n"
    for s in range(ns):
        want += g(t[s], 6) + "".join("\t" + g(c, 6) for c in corner)
        want += "".join("\t" + g(x, 18, 24) for x in v[s]) + "\n"
    assert p.read_bytes() == want.encode()


def test_detector_txt_values_round_trip(tmp_path):
    """18 significant digits give every double back exactly."""
    rng = np.random.default_rng(4)
    v = rng.standard_normal((3, 50)) * 1e-3
    p = tmp_path / "d.dat"
    chiml_amd.write_detector_txt(p, np.arange(3.0), v, np.zeros(3))
    rows = [line.split("\t") for line in p.read_text().splitlines()[1:]]
    back = np.array([[float(x) for x in r[4:]] for r in rows])
    assert np.array_equal(back, v)


@pytest.mark.parametrize("sz", [(3, 2, 1), (2, 3, 4), (1, 1, 1)])
def test_detector_bin_bytes(tmp_path, sz):
    rng = np.random.default_rng(5)
    ns = 4
    npts = sz[0] * sz[1] * sz[2]
    t = np.arange(ns) * 0.5
    v = rng.standard_normal((ns, npts))
    loc = (7, 8, 0)
    p = tmp_path / "d.bin"
    chiml_amd.write_detector_bin(p, t, v, sz, loc)
    want = struct.pack("<3i", *sz) + struct.pack("<3i", *loc)
    for s in range(ns):
        want += struct.pack("<d", t[s])
        # input point order: y outermost, z, x fastest; file order: z, y, x
        vv = v[s].reshape(sz[1], sz[2], sz[0])
        for kk in range(sz[2]):
            for jj in range(sz[1]):
                want += struct.pack(f"<{sz[0]}d", *vv[jj, kk, :])
    assert p.read_bytes() == want


def _flux_header(nfaces, front_back):
    h = "#" + "freq\tabs(incd)\treal(incd)\timag(incd)"
    if nfaces > 1:
        h += ("\tabs(right)\treal(right)\timag(right)\tabs(left)\treal(left)\timag(left)\tabs(top)\treal(top)"
              "\timag(top)\tabs(bot)\treal(bot)\timag(bot)")
        if front_back:
            h += "\tabs(front)\treal(front)\timag(front)\tabs(back)\treal(back)\timag(back)"
    return h + "freq\tabs(total)\treal(total)\timag(total)\n"


@pytest.mark.parametrize("nfaces,front_back", [(6, True), (4, False), (1, False)])
def test_flux_table_bytes(tmp_path, nfaces, front_back):
    rng = np.random.default_rng(6)
    nfreq = 9
    freqs = 2 * np.pi * np.linspace(0.05, 0.15, nfreq)
    faces = (rng.standard_normal((nfaces, nfreq)) + 1j * rng.standard_normal((nfaces, nfreq))) * 1e-9
    faces[0, 0] = 0.0
    faces[-1, 1] = -1.0 + 0.0j
    p = tmp_path / "f.dat"
    chiml_amd.write_flux(p, freqs, faces, front_back)
    want = _flux_header(nfaces, front_back)
    for k in range(nfreq):
        want += g(freqs[k], 15, 18) + "\t" + g(0.0, 15, 16) + "\t" + g(0.0, 15, 16) + "\t" + g(0.0, 15, 16) + "\t"
        tot = 0j
        for v in range(nfaces):
            x = faces[v, k]
            tot += x
            want += g(abs(x), 15, 16) + "\t" + g(x.real, 15, 16) + "\t" + g(x.imag, 15, 16) + "\t"
        if nfaces > 1:
            want += g(abs(tot), 15, 16) + "\t" + g(tot.real, 15, 16) + "\t" + g(tot.imag, 15, 16)
        want += "\n"
    assert p.read_bytes() == want.encode()


def test_pops_bytes(tmp_path):
    rng = np.random.default_rng(7)
    ns, interval, dt = 12, 3, 0.0123456789
    pops = rng.random(ns) + 1j * rng.standard_normal(ns) * 1e-12
    pops[0] = 1.0
    p = tmp_path / "qe_0_level_0.dat"
    chiml_amd.write_pops(p, pops, interval, dt)
    want = ""
    for tt in range(ns):
        x = pops[tt]
        want += g(tt * interval * dt, 9, 9) + "\t" + g(x.real, 16, 16) + "\t" + g(x.imag, 16, 16) + "\t"
        want += g(abs(x), 16, 16) + "\n"
    assert p.read_bytes() == want.encode()


def test_empty_series(tmp_path):
    chiml_amd.write_detector_txt(tmp_path / "a.dat", np.zeros(0), np.zeros(0), np.zeros(3))
    assert (tmp_path / "a.dat").read_text() == "# time\tx\ty\tz\tfield\n"
    chiml_amd.write_detector_bin(tmp_path / "b.dat", np.zeros(0), np.zeros(0), (2, 1, 1), (0, 0, 0))
    assert (tmp_path / "b.dat").read_bytes() == struct.pack("<6i", 2, 1, 1, 0, 0, 0)
    chiml_amd.write_pops(tmp_path / "c.dat", np.zeros(0, complex), 1, 0.1)
    assert (tmp_path / "c.dat").read_bytes() == b""


def test_writer_errors(tmp_path):
    with pytest.raises(chiml_amd.ChimlError, match="cannot open"):
        chiml_amd.write_pops(tmp_path / "missing_dir" / "x.dat", np.ones(2, complex), 1, 0.1)
    with pytest.raises(chiml_amd.ChimlError, match="size must be positive"):
        chiml_amd.write_detector_bin(tmp_path / "b.dat", np.zeros(1), np.zeros(1), (0, 1, 1), (0, 0, 0))


def _freq_header(lead, ngrid):
    h = "#" + lead
    for nn in range(1, ngrid):
        h += f"\tabs(field {nn})\treal(field {nn})\timag(field {nn})"
    if ngrid > 1:
        h += "\tabs(total)\treal(total)\timag(total)"
    return h + "\n"


@pytest.mark.parametrize("ngrid,nincd", [(1, 1), (3, 3), (2, 3)])
def test_freq_table_bytes(tmp_path, ngrid, nincd):
    """parallelDetectorFREQ_Base::toFile with incident fields
    (src/DTC/parallelDTC_FREQ.hpp:398-444): the header names one grid fewer
    than there are; every value is setw(16) at precision 12, real / imag /
    abs; zero incident columns (and their total with more than two)."""
    rng = np.random.default_rng(8)
    nfreq = 6
    freqs = np.linspace(2.2e14, 4.4e14, nfreq)
    vals = (rng.standard_normal((ngrid, nfreq)) + 1j * rng.standard_normal((ngrid, nfreq))) * 1e-7
    vals[0, 0] = 0.0
    vals[-1, 1] = -3.5 + 0.0j
    p = tmp_path / "fq.dat"
    chiml_amd.write_freq_table(p, freqs, vals, nincd)
    # setw(16) on a longer header string pads nothing
    want = _freq_header("freq\tabs(incd)\treal(incd)\timag(incd)", ngrid)
    for k in range(nfreq):
        want += g(freqs[k], 12, 16)
        for _ in range(nincd):
            want += "\t" + g(0.0, 12, 16) + "\t" + g(0.0, 12, 16) + "\t" + g(0.0, 12, 16)
        if nincd > 2:
            want += "\t" + g(0.0, 12, 16) + "\t" + g(0.0, 12, 16) + "\t" + g(0.0, 12, 16)
        tot = 0j
        for gi in range(ngrid):
            x = vals[gi, k]
            tot += x
            want += "\t" + g(x.real, 12, 16) + "\t" + g(x.imag, 12, 16) + "\t" + g(abs(x), 12, 16)
        if ngrid > 1:
            want += "\t" + g(tot.real, 12, 16) + "\t" + g(tot.imag, 12, 16) + "\t" + g(abs(tot), 12, 16)
        want += "\n"
    assert p.read_bytes() == want.encode()


@pytest.mark.parametrize("ngrid,nincd", [(1, 1), (2, 3)])
def test_freq_maps_bytes(tmp_path, ngrid, nincd):
    """parallelDetectorFREQ_Base::toMap with incident fields
    (src/DTC/parallelDTC_FREQ.hpp:506-580): one file per frequency named
    <path>.<frequency as %f>; a line per point, y outermost, then z, x
    fastest, with x at setw(16), the incident values, and per grid abs at
    setw(16) then real and imag unpadded."""
    rng = np.random.default_rng(9)
    nfreq, sz = 3, (3, 2, 2)
    npt = sz[0] * sz[1] * sz[2]
    freqs = np.array([1.5e14, 2.25e14, 3.0e14])
    vals = (rng.standard_normal((ngrid, nfreq, npt)) + 1j * rng.standard_normal((ngrid, nfreq, npt))) * 1e-5
    base = tmp_path / "map"
    chiml_amd.write_freq_maps(base, freqs, sz, vals, nincd)
    incd = [0.0] * (3 * nincd + (3 if nincd > 2 else 0))
    for k in range(nfreq):
        p = tmp_path / ("map." + "%f" % freqs[k])
        want = _freq_header("x\ty\tz\tabs(incd)\treal(incd)\timag(incd)", ngrid)
        for yy in range(sz[1]):
            for zz in range(sz[2]):
                for xx in range(sz[0]):
                    jj = xx + zz * sz[0] + yy * sz[0] * sz[2]
                    want += str(xx).rjust(16) + "\t" + str(yy) + "\t" + str(zz)
                    want += "".join("\t" + g(v, 12, 16) for v in incd)
                    tot = 0j
                    for gi in range(ngrid):
                        x = vals[gi, k, jj]
                        tot += x
                        want += "\t" + g(abs(x), 12, 16) + "\t" + g(x.real, 12) + "\t" + g(x.imag, 12)
                    if ngrid > 1:
                        want += "\t" + g(tot.real, 12, 16) + "\t" + g(tot.imag, 12, 16) + "\t" + g(abs(tot), 12, 16)
                    want += "\n"
        assert p.read_bytes() == want.encode(), k
