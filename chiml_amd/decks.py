"""chiML input decks for the benchmark workloads.

Each function returns a JSON string in the reference's input schema
(src/INPUTS/parallelInputs.cpp:12-838, py_utilities/write_json.py) with the
two namespaced extensions this build adds: ``"GPU": {"precision": ...}`` and
per-component ``"eps_diag"`` for diagonal anisotropy; ML decks with full
Hermitian H0 / dipole matrices use ``"H0_eV"`` and ``"dipole_Debye"``.

Sizes follow the reference's point count n = floor(size/d + 0.5) + 1, so a
deck with N cells along an axis has size (N - 1) * d.  Step counts follow
nSteps = ceil(tLim/dt); tLim is set to (steps - 0.5) * dt.
"""
import json
import math

C0 = 2.9979245e8
A_UNIT = 1e-7  # default unit length a (m)


def splitmix64(seed):
    """SplitMix64 generator (Steele, Lea, Flood 2014); yields floats in [0, 1)."""
    state = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        z = z ^ (z >> 31)
        yield (z >> 11) * (1.0 / (1 << 53))


def ev2fdtd(ev, a=A_UNIT):
    """eV -> FDTD frequency units (parallelInputs.cpp:1182-1185)."""
    return ev / 4.135666e-15 * a / C0


def dt_of(d, courant=0.5):
    return courant / math.sqrt(3.0 / (d * d))


def _cell(n3, d, courant, steps, pml, precision, extra=None):
    size = [(n - 1) * d if n > 1 else 0.0 for n in n3]
    dt = dt_of(d, courant)
    cc = {
        "size": size,
        "res": int(round(1.0 / d)),
        "stepSize": [d, d, d],
        "courant": courant,
        "tLim": (steps - 0.5) * dt,
        "a": A_UNIT,
    }
    if n3[2] <= 1:
        cc["pol"] = "Ez"
    deck = {
        "CompCell": cc,
        "PML": {"thickness": [pml * d if n3[0] > 1 else 0, pml * d if n3[1] > 1 else 0, pml * d if n3[2] > 1 else 0]},
        "SourceList": [],
        "TFSF": [],
        "ObjectList": [],
        "DetectorList": [],
        "FluxList": [],
        "GPU": {"precision": precision},
    }
    if extra:
        deck.update(extra)
    return deck, size, dt


def empty_deck(n3, d, steps, pml, precision="fp64", courant=0.5):
    """Deck dict with the computational cell only (no sources or objects):
    n3 points per axis (n3[2] = 1 for 2D TMz), uniform step d, pml cells."""
    deck, _, _ = _cell(n3, d, courant, steps, pml, precision)
    return deck


def gaussian_pulse(fcen, fwidth, cutoff, field):
    """Gaussian pulse centred at t_0 = cutoff/fwidth.  t_0 is given explicitly:
    the reference's default is cutoff/fwidth**2 (parallelInputs.cpp:139)."""
    return {"profile": "gaussian", "fcen": fcen, "fwidth": fwidth, "cutoff": cutoff,
            "t_0": cutoff / fwidth, "Field_Intensity": field}


def _pos(idx, n, d):
    """Real-space coordinate of grid point idx on an axis of n points."""
    return idx * d - (n - 1) * d / 2.0


def config1_2d(n=256, steps=2000, profile="gaussian", pml=20, d=0.05, cyl_radius_cells=40, eps_cyl=4.0,
               src_cell=(64, 128), ndet=8, seed=1, precision="fp64"):
    """Plain 2D TMz FDTD: dielectric cylinder, point source, Ez point detectors."""
    deck, size, dt = _cell((n, n, 1), d, 0.5, steps, pml, precision)
    fcen = A_UNIT / 500e-9  # 500 nm centre wavelength
    if profile == "ricker":
        pulse = {"profile": "ricker", "fcen": fcen, "fwidth": 1.0, "cutoff": 4.0, "Field_Intensity": 1.0}
    else:
        pulse = gaussian_pulse(fcen, 0.5 * fcen, 4.0, 1.0)
    deck["SourceList"].append({
        "pol": "Ez", "size": [0.0, 0.0, 0.0],
        "loc": [_pos(src_cell[0], n, d), _pos(src_cell[1], n, d), 0.0],
        "PulseList": [pulse],
    })
    deck["ObjectList"].append({
        "shape": "cylinder", "material": "custom", "eps": eps_cyl, "pols": [],
        "radius": cyl_radius_cells * d, "length": 1.0, "loc": [0.0, 0.0, 0.0],
        "unit_vectors": [{"uvec": [1, 0, 0]}, {"uvec": [0, 0, 1]}, {"uvec": [0, 1, 0]}],
        "Basis_Set": [],
    })
    rng = splitmix64(seed)
    lo, hi = pml + 1, n - pml - 2
    for _ in range(ndet):
        cx = lo + int(next(rng) * (hi - lo + 1))
        cy = lo + int(next(rng) * (hi - lo + 1))
        deck["DetectorList"].append({
            "type": "Ez", "dtc_class": "cout", "fname": "dtc", "size": [0.0, 0.0, 0.0],
            "loc": [_pos(cx, n, d), _pos(cy, n, d), 0.0],
        })
    return json.dumps(deck)


def gold_like_pols(a=A_UNIT):
    """Drude (wp=9 eV, Gamma=0.07 eV) + Lorentz poles at 2.6 and 3.5 eV
    (delta-eps 1.0 and 0.5, Gamma 0.3 eV), in the getMetal convention
    (parallelInputs.cpp:1187-1221): sigma_p = f (wp/Omega)^2, gamma = pi Gamma,
    omega = 2 pi Omega; custom pols take Gamma and Omega directly."""
    wp = ev2fdtd(9.0, a)
    om_d = ev2fdtd(1e-20, a)
    return [
        {"omega": om_d, "gamma": ev2fdtd(0.07, a), "sigma_p": (wp / om_d) ** 2},
        {"omega": ev2fdtd(2.6, a), "gamma": ev2fdtd(0.3, a), "sigma_p": 1.0},
        {"omega": ev2fdtd(3.5, a), "gamma": ev2fdtd(0.3, a), "sigma_p": 0.5},
    ]


def config2_3d(n=256, steps=1000, pml=16, d=0.02, sphere_cells=40, flux_cells=120, nfreq=40,
               eps_bg=(2.0, 2.25, 2.5), precision="fp32", with_flux=True, with_source=True, ndet=0):
    """3D Yee with anisotropic background, Drude+2 Lorentz sphere, soft
    plane-wave sheet and a DFT flux box; ndet > 0 adds point detectors of Ex,
    Ey, Ez, Hx, Hy, Hz in turn on the sheet side of the sphere, between the
    sheet and the sphere's surface."""
    deck, size, dt = _cell((n, n, n), d, 0.5, steps, pml, precision)
    deck["ObjectList"].append({
        "shape": "block", "material": "custom", "eps": eps_bg[1], "eps_diag": list(eps_bg), "pols": [],
        "size": list(size), "loc": [0.0, 0.0, 0.0], "Basis_Set": [],
    })
    deck["ObjectList"].append({
        "shape": "sphere", "material": "custom", "eps": 1.0, "pols": gold_like_pols(),
        "radius": sphere_cells * d, "loc": [0.0, 0.0, 0.0], "Basis_Set": [],
    })
    if with_source:
        inner = (n - 2 * pml - 2) * d
        fcen = A_UNIT / 600e-9
        deck["SourceList"].append({
            "pol": "Ex", "size": [inner, inner, 0.0],
            "loc": [0.0, 0.0, _pos(pml + 4, n, d)],
            "PulseList": [gaussian_pulse(fcen, 2.0 * fcen, 3.0, 1.0)],
        })
    comps = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"]
    zc = max(pml + 6, n // 2 - sphere_cells - 3)
    for k in range(ndet):
        deck["DetectorList"].append({
            "type": comps[k % 6], "dtc_class": "cout", "fname": "dtc", "size": [0.0, 0.0, 0.0],
            "loc": [_pos(n // 2 + 3 * (k // 6) + 1, n, d), _pos(n // 2 + 2 * (k % 6) - 5, n, d), _pos(zc, n, d)],
        })
    if with_flux:
        fs = (flux_cells - 1) * d
        deck["FluxList"].append({
            "name": "flux", "size": [fs, fs, fs], "loc": [0.0, 0.0, 0.0],
            "lamL": 400e-9 / A_UNIT, "lamR": 800e-9 / A_UNIT, "nfreq": nfreq,
        })
    return json.dumps(deck)


def two_level_emitter(x_size, y_size, loc, e_ev=2.0, dipole_debye=10.0, mol_den=1e25, t1=1e-12, t2=100e-15,
                      z_size=0.0):
    """A block of 2-level emitters, basis {(0,0),(1,0)} (mu_z = coupling/sqrt(3))."""
    return {
        "shape": "block", "material": "custom", "eps": 1.0, "pols": [],
        "size": [x_size, y_size, z_size], "loc": list(loc),
        "Basis_Set": [{"l": 0, "m": 0}, {"l": 1, "m": 0}],
        "Energy_Levels": [{"E_cen": [0.0]}, {"E_cen": [e_ev]}],
        "couplings": [0.0, dipole_debye * math.sqrt(3.0), dipole_debye * math.sqrt(3.0), 0.0],
        "mol_den": mol_den,
        "RelaxationOperators": [{"state_i": 1, "state_f": 0, "rate": 1.0 / t1,
                                 "dephasing_rate": 1.0 / t2 - 1.0 / (2.0 * t1)}],
        "dtc_levs": [0, 3],
        "gam": [],
    }


def config3_ml2d(n=1024, steps=20000, pml=20, d=0.05, slab=(512, 64), seed=3, precision="fp32", det_cells=8):
    """2D Maxwell-Liouville: a slab of 2-level emitters under a resonant pulse."""
    deck, size, dt = _cell((n, n, 1), d, 0.5, steps, pml, precision)
    rng = splitmix64(seed)
    peak = 1e7 * (100.0 ** next(rng))  # log-uniform in [1e7, 1e9] V/m
    e_ev = 2.0
    fcen = e_ev / 4.135666e-15 * A_UNIT / C0
    fwhm_t = 50e-15 * C0 / A_UNIT
    fwidth = 2.0 * math.sqrt(2.0 * math.log(2.0)) / fwhm_t
    deck["SourceList"].append({
        "pol": "Ez", "size": [(n - 2 * pml - 2) * d, 0.0, 0.0],
        "loc": [0.0, _pos(n // 2 + slab[1] // 2 + 16, n, d), 0.0],
        "PulseList": [gaussian_pulse(fcen, fwidth, 4.0, peak)],
    })
    deck["ObjectList"].append(two_level_emitter((slab[0] - 1) * d, (slab[1] - 1) * d, (0.0, 0.0, 0.0), e_ev=e_ev))
    for k in range(det_cells):
        cx = n // 2 - det_cells // 2 + k
        deck["DetectorList"].append({
            "type": "Ez", "dtc_class": "cout", "fname": "trans", "size": [0.0, 0.0, 0.0],
            "loc": [_pos(cx, n, d), _pos(n // 2 - slab[1] // 2 - 16, n, d), 0.0],
        })
    return json.dumps(deck)


def random_hermitian_levels(nlev, seed, lo_ev=1.5, hi_ev=3.0, dip_lo=1.0, dip_hi=15.0):
    """Seeded Hermitian H0 (eV) and dipole matrices (Debye), flat column-major."""
    rng = splitmix64(seed)
    levels = sorted([0.0] + [lo_ev + (hi_ev - lo_ev) * next(rng) for _ in range(nlev - 1)])
    h_re = [0.0] * (nlev * nlev)
    h_im = [0.0] * (nlev * nlev)
    for i in range(nlev):
        h_re[i + nlev * i] = levels[i]
    for i in range(nlev):
        for j in range(i + 1, nlev):
            cr = 0.01 * (next(rng) - 0.5)
            ci = 0.01 * (next(rng) - 0.5)
            h_re[i + nlev * j] = cr
            h_im[i + nlev * j] = ci
            h_re[j + nlev * i] = cr
            h_im[j + nlev * i] = -ci
    dip = {}
    for ax in "xyz":
        re = [0.0] * (nlev * nlev)
        im = [0.0] * (nlev * nlev)
        for i in range(nlev):
            for j in range(i + 1, nlev):
                mag = dip_lo + (dip_hi - dip_lo) * next(rng)
                ph = 2 * math.pi * next(rng)
                s = mag / math.sqrt(3.0)
                re[i + nlev * j] = s * math.cos(ph)
                im[i + nlev * j] = s * math.sin(ph)
                re[j + nlev * i] = s * math.cos(ph)
                im[j + nlev * i] = -s * math.sin(ph)
        dip[ax] = {"re": re, "im": im}
    return levels, {"re": h_re, "im": h_im}, dip


def n_level_emitter(nlev, size3, loc, seed, mol_den=1e25):
    levels, h0, dip = random_hermitian_levels(nlev, seed)
    rng = splitmix64(seed + 1000)
    relax = []
    for i in range(1, nlev):
        t2 = (10e-15 + 190e-15 * next(rng))
        relax.append({"state_i": i, "state_f": 0, "rate": 1e12, "dephasing_rate": 1.0 / t2})
    return {
        "shape": "block", "material": "custom", "eps": 1.0, "pols": [],
        "size": list(size3), "loc": list(loc),
        "Basis_Set": [{"l": 0, "m": 0}] * nlev,
        "Energy_Levels": [{"E_cen": [lv]} for lv in levels],
        "couplings": [],
        "H0_eV": h0,
        "dipole_Debye": dip,
        "mol_den": mol_den,
        "RelaxationOperators": relax,
        "dtc_levs": [0, nlev + 1],
        "gam": [],
    }


def config4_ml3d(n=192, steps=5000, pml=16, d=0.02, layer_cells=10, film_cells=20, nlev=6, seed=4, precision="fp32"):
    """3D Maxwell-Liouville: a layer of N-level emitters (layer_cells planes,
    n * n * layer_cells emitters) one plane above a Drude film of film_cells
    planes.  Each block is centred between grid planes with a z size of whole
    cells, so both of its faces fall between planes and it holds exactly that
    many planes."""
    deck, size, dt = _cell((n, n, n), d, 0.5, steps, pml, precision)
    full = (n - 1) * d
    film_lo = n // 2 - film_cells // 2
    z_film = _pos(film_lo, n, d) + (film_cells - 1) * d / 2.0
    deck["ObjectList"].append({
        "shape": "block", "material": "custom", "eps": 1.0, "pols": gold_like_pols()[:1],
        "size": [full, full, film_cells * d], "loc": [0.0, 0.0, z_film], "Basis_Set": [],
    })
    layer_lo = film_lo + film_cells + 1
    z_layer = _pos(layer_lo, n, d) + (layer_cells - 1) * d / 2.0
    deck["ObjectList"].append(n_level_emitter(nlev, [full, full, layer_cells * d], (0.0, 0.0, z_layer), seed))
    fcen = A_UNIT / 600e-9
    inner = (n - 2 * pml - 2) * d
    deck["SourceList"].append({
        "pol": "Ex", "size": [inner, inner, 0.0], "loc": [0.0, 0.0, _pos(n - pml - 6, n, d)],
        "PulseList": [gaussian_pulse(fcen, 2.0 * fcen, 3.0, 1e8)],
    })
    return json.dumps(deck)


def sweep_requests(n_req=512, seed=5, rate_hz=20.0, steps_range=(1000, 10000), emitter_frac=0.25,
                   sizes_2d=(128, 256, 512, 1024), sizes_3d=(64, 96, 128, 160, 192)):
    """Seeded parameter-sweep requests: (arrival time in s, deck json) pairs.

    Poisson arrivals at rate_hz; half 2D TMz and half 3D grids; a random
    cylinder (2D) or sphere (3D) that is a dielectric (eps in [1.5, 12]) or a
    Drude metal; a fraction with a 2- or 4-level emitter block; steps uniform
    in steps_range; FP32 fields and FP64 density matrices."""
    rng = splitmix64(seed)
    t = 0.0
    out = []
    for _ in range(n_req):
        t += -math.log(1.0 - next(rng)) / rate_hz
        is3 = next(rng) < 0.5
        sizes = sizes_3d if is3 else sizes_2d
        n = sizes[min(int(next(rng) * len(sizes)), len(sizes) - 1)]
        steps = steps_range[0] + int(next(rng) * (steps_range[1] - steps_range[0] + 1))
        d = 0.02 if is3 else 0.05
        pml = 10 if is3 else 12
        n3 = (n, n, n) if is3 else (n, n, 1)
        deck, size, dt = _cell(n3, d, 0.5, steps, pml, "fp32")
        r_cells = max(3, int(n * (0.08 + 0.12 * next(rng))))
        obj = {"shape": "sphere" if is3 else "cylinder", "material": "custom", "radius": r_cells * d,
               "loc": [0.0, 0.0, 0.0], "Basis_Set": []}
        if not is3:
            obj.update({"length": 1.0, "unit_vectors": [{"uvec": [1, 0, 0]}, {"uvec": [0, 0, 1]}, {"uvec": [0, 1, 0]}]})
        if next(rng) < 0.5:
            obj.update({"eps": 1.5 + 10.5 * next(rng), "pols": []})
        else:
            obj.update({"eps": 1.0, "pols": gold_like_pols()[:1]})
        deck["ObjectList"].append(obj)
        fcen = A_UNIT / 600e-9
        src_pos = _pos(pml + 3, n, d)
        if is3:
            inner = (n - 2 * pml - 2) * d
            deck["SourceList"].append({"pol": "Ex", "size": [inner, inner, 0.0], "loc": [0.0, 0.0, src_pos],
                                       "PulseList": [gaussian_pulse(fcen, 2.0 * fcen, 3.0, 1e8)]})
        else:
            deck["SourceList"].append({"pol": "Ez", "size": [0.0, 0.0, 0.0], "loc": [src_pos, 0.0, 0.0],
                                       "PulseList": [gaussian_pulse(fcen, 2.0 * fcen, 3.0, 1e8)]})
        det = {"type": "Ez", "dtc_class": "cout", "fname": "t", "size": [0.0, 0.0, 0.0],
               "loc": [-src_pos, 0.0, 0.0]}
        deck["DetectorList"].append(det)
        if next(rng) < emitter_frac:
            nlev = 2 if next(rng) < 0.5 else 4
            em_cells = max(2, n // 16)
            ex = (em_cells - 1) * d
            loc = (d / 2 if n % 2 == 0 else 0.0,) * 2 + ((d / 2 if n % 2 == 0 else 0.0) if is3 else 0.0,)
            size3 = [ex, ex, ex if is3 else 0.0]
            loc = (loc[0] + (r_cells + 2) * d, loc[1], loc[2])
            if nlev == 2:
                em = two_level_emitter(size3[0], size3[1], loc, z_size=size3[2])
            else:
                em = n_level_emitter(4, size3, loc, seed=seed * 1000 + len(out))
            deck["ObjectList"].append(em)
        out.append((t, json.dumps(deck)))
    return out


def tfsf_deck(n3, steps, box, m=None, angles=None, psi=90.0, pml=12, d=0.05, precision="fp64", fcen_nm=500.0,
              circ=None, ndet=0, seed=7, tfsf_pml=20, objects=None, flux=None, fwidth_ratio=0.5, cutoff=4.0):
    """A TFSF surface (the reference's "TFSF" list, parallelInputs.cpp:219-408)
    centred in a vacuum cell of n3 points (n3[2] = 1 for 2D TMz): box = points
    per axis of the total-field region, m = integer propagation vector or
    angles = (theta, phi) in degrees, psi = polarisation angle in degrees, a
    Gaussian pulse centred at fcen_nm.  ndet > 0 adds seeded point detectors of
    every existing component inside and outside the box."""
    deck, size, dt = _cell(n3, d, 0.5, steps, pml, precision)
    fcen = A_UNIT / (fcen_nm * 1e-9)
    t = {"size": [(b - 1) * d if n > 1 else 0.0 for b, n in zip(box, n3)], "loc": [0.0, 0.0, 0.0],
         "psi": psi, "pmlThick": tfsf_pml, "PulseList": [gaussian_pulse(fcen, fwidth_ratio * fcen, cutoff, 1.0)]}
    if m is not None:
        t["m"] = list(m)
    else:
        t["theta"], t["phi"] = angles
    if circ is not None:
        t["circPol"] = circ
    deck["TFSF"].append(t)
    if objects:
        deck["ObjectList"].extend(objects)
    if flux:
        deck["FluxList"].append(flux)
    if ndet:
        comps = ["Ez", "Hx", "Hy"] if n3[2] <= 1 else ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"]
        rng = splitmix64(seed)
        for k in range(ndet):
            c = [pml + 2 + int(next(rng) * (nn - 2 * pml - 4)) if nn > 1 else 0 for nn in n3]
            deck["DetectorList"].append({
                "type": comps[k % len(comps)], "dtc_class": "cout", "fname": "dtc", "size": [0.0, 0.0, 0.0],
                "loc": [_pos(c[a], n3[a], d) if n3[a] > 1 else 0.0 for a in range(3)],
            })
    return json.dumps(deck)
