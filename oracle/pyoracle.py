"""ctypes wrapper around the FP64 oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, as the checker; the product path never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build() or make -C oracle)")
        L = ctypes.CDLL(path)
        L.oracle_create.restype = ctypes.c_void_p
        L.oracle_create.argtypes = [ctypes.c_char_p]
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oracle_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.oracle_flags.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.oracle_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.oracle_num_detectors.argtypes = [ctypes.c_void_p]
        L.oracle_detector.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.oracle_flux.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.oracle_pops.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.oracle_density.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.oracle_set_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.oracle_set_density.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.oracle_qe_drive.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.oracle_qe_gamma.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.oracle_hbar.restype = ctypes.c_double
        L.oracle_hbar.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oracle_wigner3j.restype = ctypes.c_double
        L.oracle_wigner3j.argtypes = [ctypes.c_int] * 6
        L.oracle_qe_params.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]
        L.oracle_basis_dipole.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_flux_save.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p]
        L.oracle_tfsf_line.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.oracle_freq_incd.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.oracle_flux_incd.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.oracle_incd_record.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.oracle_num_freq_detectors.argtypes = [ctypes.c_void_p]
        L.oracle_freq_info.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 4
        L.oracle_freq_fields.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.oracle_freq_values.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.oracle_pulse.restype = ctypes.c_double
        L.oracle_pulse.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_double]
        _LIB = L
    return _LIB


def set_threads(n):
    """Worker threads of the oracle's update lists (results do not depend on it)."""
    lib().oracle_set_threads(int(n))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """CPU FP64 restatement of the reference step loop for one JSON deck."""

    def __init__(self, deck_json):
        L = lib()
        self._h = L.oracle_create(deck_json.encode())
        if not self._h:
            raise ValueError(L.oracle_last_error().decode())
        n3 = (ctypes.c_int * 3)()
        dt = ctypes.c_double()
        ns = ctypes.c_int()
        is3d = ctypes.c_int()
        L.oracle_info(self._h, n3, ctypes.byref(dt), ctypes.byref(ns), ctypes.byref(is3d))
        self.shape = (n3[2], n3[1], n3[0])  # (z, y, x) for numpy
        self.dt = dt.value
        self.nsteps = ns.value
        self.is3d = bool(is3d.value)
        fl = (ctypes.c_int * 2)()
        L.oracle_flags(self._h, fl)
        self.dielectric_mat_in_pml = bool(fl[0])

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_destroy(self._h)
            self._h = None

    def step(self, n=1):
        if lib().oracle_step(self._h, int(n)) != 0:
            raise RuntimeError(lib().oracle_last_error().decode())

    def field(self, comp):
        """comp 0..5 = Ex,Ey,Ez,Hx,Hy,Hz; array of shape (nz, ny, nx)."""
        out = np.zeros(int(np.prod(self.shape)), dtype=np.float64)
        lib().oracle_field(self._h, int(comp), _dp(out))
        return out.reshape(self.shape)

    def num_detectors(self):
        return lib().oracle_num_detectors(self._h)

    def detector(self, d, cap=1 << 20):
        npts = ctypes.c_int()
        times = np.zeros(1, dtype=np.float64)
        ns = lib().oracle_detector(self._h, d, _dp(times), _dp(times), 0, ctypes.byref(npts))
        times = np.zeros(ns, dtype=np.float64)
        vals = np.zeros(ns * max(npts.value, 1), dtype=np.float64)
        lib().oracle_detector(self._h, d, _dp(times), _dp(vals), ns, ctypes.byref(npts))
        return times, vals.reshape(ns, npts.value)

    def num_freq_detectors(self):
        return lib().oracle_num_freq_detectors(self._h)

    def freq_info(self, d):
        """(nfreq, grids, sampled points per grid, samples taken) of frequency detector d."""
        v = [ctypes.c_int() for _ in range(4)]
        lib().oracle_freq_info(self._h, d, *[ctypes.byref(x) for x in v])
        return tuple(x.value for x in v)

    def freq_fields(self, d):
        """Accumulated DFT of frequency detector d, shape (grids, points, nfreq)."""
        nf, ng, npts, _ = self.freq_info(d)
        n = ng * npts * nf
        re = np.zeros(max(n, 1))
        im = np.zeros(max(n, 1))
        lib().oracle_freq_fields(self._h, d, _dp(re), _dp(im), n)
        return (re[:n] + 1j * im[:n]).reshape(ng, npts, nf)

    def freq_values(self, d):
        """(printed frequencies, values of shape (grids, nfreq)) of frequency detector d's table."""
        nf, ng, _, _ = self.freq_info(d)
        re = np.zeros(ng * nf)
        im = np.zeros(ng * nf)
        fr = np.zeros(nf)
        lib().oracle_freq_values(self._h, d, _dp(re), _dp(im), _dp(fr))
        return fr, (re + 1j * im).reshape(ng, nf)

    def tfsf_line(self, t, comp, which=0, nmax=1 << 16):
        """TFSF surface t's incident line of component comp (0..5): which 0 =
        E_incd/H_incd, 1 = D/B, 2 = the pulse line (None without one)."""
        re = np.zeros(nmax)
        im = np.zeros(nmax)
        n = lib().oracle_tfsf_line(self._h, t, comp, which, _dp(re), _dp(im), nmax)
        if n < 0:
            raise IndexError("no such TFSF surface or component")
        return (re[:n] + 1j * im[:n]) if n else None

    def incd_record(self, comp, nmax=1 << 22):
        """E_incd_ / H_incd_ records of component comp (0..5): 2 per step."""
        re = np.zeros(nmax)
        im = np.zeros(nmax)
        n = lib().oracle_incd_record(self._h, comp, _dp(re), _dp(im), nmax)
        return re[:n] + 1j * im[:n]

    def flux_save(self, f, path):
        """Writes flux f's face fields as the reference's saveFields does."""
        if lib().oracle_flux_save(self._h, f, str(path).encode()) != 0:
            raise RuntimeError(lib().oracle_last_error().decode())

    def freq_incd(self, d, nmax=1 << 16):
        """Incident values of frequency detector d's table, (fields, nfreq)."""
        re = np.zeros(nmax)
        im = np.zeros(nmax)
        k = lib().oracle_freq_incd(self._h, d, _dp(re), _dp(im))
        nf = self.freq_info(d)[0]
        return (re[:k * nf] + 1j * im[:k * nf]).reshape(k, nf)

    def flux_incd(self, f, nmax=4096):
        """Incident flux per frequency (getFlux's incident column)."""
        re = np.zeros(nmax)
        im = np.zeros(nmax)
        n = lib().oracle_flux_incd(self._h, f, _dp(re), _dp(im), nmax)
        return re[:n] + 1j * im[:n]

    def flux(self, f, nmax=4096):
        re = np.zeros(nmax)
        im = np.zeros(nmax)
        n = lib().oracle_flux(self._h, f, _dp(re), _dp(im), nmax)
        return re[:n] + 1j * im[:n]

    def pops(self, q, lev, nmax=1 << 20):
        re = np.zeros(nmax)
        im = np.zeros(nmax)
        n = lib().oracle_pops(self._h, q, lev, _dp(re), _dp(im), nmax)
        return re[:n] + 1j * im[:n]

    def set_field(self, comp, arr):
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).reshape(-1))
        if a.size != int(np.prod(self.shape)) or lib().oracle_set_field(self._h, int(comp), _dp(a)) != 0:
            raise ValueError("set_field: bad component or size")

    def set_density(self, q, member, emitter, rho):
        rho = np.asarray(rho, dtype=np.complex128)
        re = np.ascontiguousarray(rho.real.reshape(-1, order="F"))
        im = np.ascontiguousarray(rho.imag.reshape(-1, order="F"))
        lib().oracle_set_density(self._h, q, member, emitter, _dp(re), _dp(im))

    def qe_drive(self, q, e3, nsteps):
        """Integrates group q's densities nsteps times under a constant field e3 (no grid step)."""
        e = np.ascontiguousarray(np.asarray(e3, dtype=np.float64))
        lib().oracle_qe_drive(self._h, q, _dp(e), int(nsteps))

    def qe_gamma(self, q, nlev):
        g = np.zeros(nlev ** 4)
        lib().oracle_qe_gamma(self._h, q, _dp(g))
        return g.reshape(nlev * nlev, nlev * nlev)

    def hbar(self, q):
        return lib().oracle_hbar(self._h, q)

    def qe_params(self, q, axis, nlev):
        """(number density, hbar, H0 real part, dipole real part along axis) of
        emitter group q in code units."""
        na = ctypes.c_double()
        hb = ctypes.c_double()
        h0 = np.zeros(nlev * nlev)
        mu = np.zeros(nlev * nlev)
        lib().oracle_qe_params(self._h, q, axis, ctypes.byref(na), ctypes.byref(hb), _dp(h0), _dp(mu))
        return na.value, hb.value, h0.reshape(nlev, nlev), mu.reshape(nlev, nlev)

    def density(self, q, member, emitter, nlev):
        re = np.zeros(nlev * nlev)
        im = np.zeros(nlev * nlev)
        lib().oracle_density(self._h, q, member, emitter, _dp(re), _dp(im))
        return (re + 1j * im).reshape(nlev, nlev, order="F")


def wigner3j(j1, j2, j3, m1, m2, m3):
    return lib().oracle_wigner3j(j1, j2, j3, m1, m2, m3)


def basis_dipole(basis, axis):
    """Dipole operator matrix (x, y, z for axis 0, 1, 2) in a basis of (l, m)
    spherical harmonics, as the emitter setup builds it."""
    n = len(basis)
    lm = (ctypes.c_int * (2 * n))(*[v for b in basis for v in b])
    re = np.zeros(n * n)
    im = np.zeros(n * n)
    lib().oracle_basis_dipole(n, lm, axis, _dp(re), _dp(im))
    return (re + 1j * im).reshape(n, n)


def pulse(deck_json, src, t):
    return lib().oracle_pulse(deck_json.encode(), src, t)
