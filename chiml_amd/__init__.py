"""chiml_amd -- MI355X-native time-stepping engine for chiML decks.

Python mirror of the reference's manager interface
(src/FDTD_MANAGER/parallelFDTDField.hpp: step(), E_/H_, dtcArr(), fluxArr(),
qeArr(), dt(); src/main.cpp:47-118), backed by libchiml_amd.so through the C
ABI of include/chiml/chiml_sim.h.  Import works without a GPU; creating a
simulation needs the built library and a HIP device and fails loudly otherwise.
"""
import ctypes
import os

import numpy as np

__all__ = ["FDTDField", "Server", "load_library", "LIB_PATH", "ChimlError"]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libchiml_amd.so")
_LIB = None

FIELD_NAMES = {"Ex": 0, "Ey": 1, "Ez": 2, "Hx": 3, "Hy": 4, "Hz": 5}


class ChimlError(RuntimeError):
    pass


def load_library():
    """Loads libchiml_amd.so (raises if it was not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ChimlError(f"{LIB_PATH} is not built; run __graft_entry__.build() (or python chiml_amd/build.py)")
    L = ctypes.CDLL(LIB_PATH)
    c_int_p = ctypes.POINTER(ctypes.c_int)
    c_dbl_p = ctypes.POINTER(ctypes.c_double)
    c_size_p = ctypes.POINTER(ctypes.c_size_t)
    L.chiml_sim_create.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]
    L.chiml_sim_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.chiml_sim_sync.argtypes = [ctypes.c_void_p]
    L.chiml_sim_info.argtypes = [ctypes.c_void_p, c_int_p, c_dbl_p, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p]
    L.chiml_sim_field.argtypes = [ctypes.c_void_p, ctypes.c_int, c_dbl_p]
    L.chiml_sim_detector.argtypes = [ctypes.c_void_p, ctypes.c_int, c_dbl_p, c_dbl_p, ctypes.c_size_t, c_size_p, c_int_p]
    L.chiml_sim_flux.argtypes = [ctypes.c_void_p, ctypes.c_int, c_dbl_p, ctypes.c_size_t, c_int_p]
    L.chiml_sim_num_freq_detectors.argtypes = [ctypes.c_void_p, c_int_p]
    L.chiml_sim_freq_info.argtypes = [ctypes.c_void_p, ctypes.c_int, c_int_p, c_int_p, c_int_p, c_int_p, c_dbl_p]
    L.chiml_sim_freq_fields.argtypes = [ctypes.c_void_p, ctypes.c_int, c_dbl_p, ctypes.c_size_t]
    L.chiml_sim_freq_values.argtypes = [ctypes.c_void_p, ctypes.c_int, c_dbl_p, ctypes.c_size_t]
    L.chiml_sim_pops.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, c_dbl_p, ctypes.c_size_t, c_size_p]
    L.chiml_sim_incd.argtypes = [ctypes.c_void_p, ctypes.c_int, c_dbl_p, ctypes.c_size_t, c_size_p]
    L.chiml_sim_freq_incd.argtypes = [ctypes.c_void_p, ctypes.c_int, c_dbl_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int)]
    L.chiml_sim_flux_incd.argtypes = [ctypes.c_void_p, ctypes.c_int, c_dbl_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int)]
    L.chiml_sim_density.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_dbl_p, c_int_p]
    L.chiml_sim_cost.argtypes = [ctypes.c_void_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p]
    L.chiml_sim_kernel_bytes.argtypes = [ctypes.c_void_p, ctypes.c_char_p, c_dbl_p]
    L.chiml_sim_profile.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.chiml_sim_profile_only.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    L.chiml_sim_profile_read.argtypes = [ctypes.c_void_p, ctypes.c_char_p, c_dbl_p, c_int_p]
    L.chiml_sim_footprint.argtypes = [ctypes.c_void_p, c_size_p]
    L.chiml_sim_estimate_bytes.argtypes = [ctypes.c_char_p, c_size_p]
    L.chiml_sim_query.argtypes = [ctypes.c_void_p, c_int_p]
    L.chiml_sim_set_graph.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.chiml_sim_step_batch.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int]
    L.chiml_sim_step_batch_noalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int]
    c_ll_p = ctypes.POINTER(ctypes.c_longlong)
    L.chiml_server_create.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_void_p)]
    L.chiml_server_submit.argtypes = [ctypes.c_void_p, ctypes.c_char_p, c_ll_p]
    L.chiml_server_set_batch.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.chiml_server_poll.argtypes = [ctypes.c_void_p, ctypes.c_longlong, c_int_p, c_dbl_p, c_dbl_p]
    L.chiml_server_take.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.POINTER(ctypes.c_void_p)]
    L.chiml_server_error.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_char_p, ctypes.c_size_t]
    L.chiml_server_drain.argtypes = [ctypes.c_void_p]
    L.chiml_server_stats.argtypes = [ctypes.c_void_p, c_dbl_p, c_size_p, c_int_p, c_int_p, c_int_p]
    L.chiml_server_destroy.argtypes = [ctypes.c_void_p]
    L.chiml_write_detector_txt.argtypes = [ctypes.c_char_p, c_dbl_p, c_dbl_p, ctypes.c_size_t, ctypes.c_int, c_dbl_p]
    L.chiml_write_detector_bin.argtypes = [ctypes.c_char_p, c_dbl_p, c_dbl_p, ctypes.c_size_t, c_int_p, c_int_p]
    L.chiml_write_flux.argtypes = [ctypes.c_char_p, c_dbl_p, ctypes.c_int, ctypes.c_int, c_dbl_p, ctypes.c_int]
    L.chiml_write_freq_table_incd.argtypes = [ctypes.c_char_p, c_dbl_p, ctypes.c_int, ctypes.c_int, c_dbl_p, ctypes.c_int, c_dbl_p]
    L.chiml_write_freq_maps_incd.argtypes = [ctypes.c_char_p, c_dbl_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int), c_dbl_p,
                                             ctypes.c_int, c_dbl_p]
    L.chiml_write_flux_incd.argtypes = [ctypes.c_char_p, c_dbl_p, ctypes.c_int, ctypes.c_int, c_dbl_p, ctypes.c_int, c_dbl_p]
    L.chiml_write_freq_table.argtypes = [ctypes.c_char_p, c_dbl_p, ctypes.c_int, ctypes.c_int, c_dbl_p, ctypes.c_int]
    L.chiml_write_freq_maps.argtypes = [ctypes.c_char_p, c_dbl_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                        c_dbl_p, ctypes.c_int]
    L.chiml_write_pops.argtypes = [ctypes.c_char_p, c_dbl_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double]
    L.chiml_sim_write_outputs.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    L.chiml_sim_last_error.restype = ctypes.c_char_p
    L.chiml_sim_destroy.argtypes = [ctypes.c_void_p]
    _LIB = L
    return L


def _check(rc):
    if rc != 0:
        raise ChimlError(load_library().chiml_sim_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def write_detector_txt(path, times, values, corner):
    """The reference's TXT point-detector file (parallelDTC_TXT.cpp) from
    times[ns] and values[ns, npts] (points y, then z, then x fastest)."""
    t, v, c = _f64(times), _f64(values), _f64(corner)
    ns = t.size
    npts = v.size // ns if ns else 0
    _check(load_library().chiml_write_detector_txt(str(path).encode(), _dp(t), _dp(v), ns, npts, _dp(c)))


def write_detector_bin(path, times, values, sz, loc):
    """The reference's BIN point-detector file (parallelDTC_BIN.cpp)."""
    t, v = _f64(times), _f64(values)
    s = (ctypes.c_int * 3)(*[int(x) for x in sz])
    lo = (ctypes.c_int * 3)(*[int(x) for x in loc])
    _check(load_library().chiml_write_detector_bin(str(path).encode(), _dp(t), _dp(v), t.size, s, lo))


def write_flux(path, freqs, faces, front_back=False, incd=None):
    """The reference's flux table (parallelFlux.hpp getFlux) from the
    per-face complex fluxes faces[nfaces, nfreq] and the incident flux per
    frequency (None: no incident field)."""
    f = _f64(freqs)
    fc = np.asarray(faces, dtype=np.complex128).reshape(-1, f.size)
    reim = _f64(np.stack([fc.real, fc.imag], axis=-1))
    if incd is None:
        _check(load_library().chiml_write_flux(str(path).encode(), _dp(f), f.size, fc.shape[0], _dp(reim),
                                               1 if front_back else 0))
        return
    ic = np.asarray(incd, dtype=np.complex128).reshape(f.size)
    ir = _f64(np.stack([ic.real, ic.imag], axis=-1))
    _check(load_library().chiml_write_flux_incd(str(path).encode(), _dp(f), f.size, fc.shape[0], _dp(reim),
                                                1 if front_back else 0, _dp(ir)))


def write_freq_table(path, freqs, values, nincd):
    """The reference's frequency-detector table (parallelDetectorFREQ_Base::toFile
    with incident fields) from values[grids, nfreq]; the nincd incident
    fields are zero."""
    f = _f64(freqs)
    v = np.asarray(values, dtype=np.complex128).reshape(-1, f.size)
    reim = _f64(np.stack([v.real, v.imag], axis=-1))
    _check(load_library().chiml_write_freq_table(str(path).encode(), _dp(f), f.size, v.shape[0], _dp(reim), int(nincd)))


def write_freq_maps(path, freqs, sz, values, nincd):
    """The reference's frequency-detector maps (toMap with incident fields),
    one file per frequency, from values[grids, nfreq, box points]."""
    f = _f64(freqs)
    npt = int(sz[0]) * int(sz[1]) * int(sz[2])
    v = np.asarray(values, dtype=np.complex128).reshape(-1, f.size, npt)
    reim = _f64(np.stack([v.real, v.imag], axis=-1))
    s3 = (ctypes.c_int * 3)(*[int(x) for x in sz])
    _check(load_library().chiml_write_freq_maps(str(path).encode(), _dp(f), f.size, v.shape[0], s3, _dp(reim), int(nincd)))


def write_pops(path, pops, interval, dt):
    """The reference's population file (QEPopDtc::toFile)."""
    p = np.asarray(pops, dtype=np.complex128)
    reim = _f64(np.stack([p.real, p.imag], axis=-1))
    _check(load_library().chiml_write_pops(str(path).encode(), _dp(reim), p.size, int(interval), float(dt)))


class Detector:
    """A point/box time-domain detector (parallelDetectorBase)."""

    def __init__(self, sim, index):
        self._sim = sim
        self.index = index

    def series(self):
        """(times, values[nsamples, npts]) -- what the TXT detector writes."""
        L = load_library()
        ns = ctypes.c_size_t()
        npts = ctypes.c_int()
        _check(L.chiml_sim_detector(self._sim._h, self.index, None, None, 0, ctypes.byref(ns), ctypes.byref(npts)))
        t = np.zeros(ns.value)
        v = np.zeros(ns.value * npts.value)
        _check(L.chiml_sim_detector(self._sim._h, self.index, _dp(t), _dp(v), ns.value, ctypes.byref(ns), ctypes.byref(npts)))
        return t, v.reshape(ns.value, npts.value)


class FreqDetector:
    """A frequency detector (parallelDetectorFREQ_Base): running DFTs of the
    fields over a box of points, read out per grid and frequency."""

    def __init__(self, sim, index):
        self._sim = sim
        self.index = index

    def info(self):
        """(printed frequencies, grids, sampled points per grid, samples taken)."""
        L = load_library()
        nf, ng, npts, nt = (ctypes.c_int() for _ in range(4))
        _check(L.chiml_sim_freq_info(self._sim._h, self.index, ctypes.byref(nf), ctypes.byref(ng), ctypes.byref(npts),
                                     ctypes.byref(nt), None))
        fr = np.zeros(nf.value)
        _check(L.chiml_sim_freq_info(self._sim._h, self.index, ctypes.byref(nf), ctypes.byref(ng), ctypes.byref(npts),
                                     ctypes.byref(nt), _dp(fr)))
        return fr, ng.value, npts.value, nt.value

    def fields(self):
        """Accumulated DFT, shape (grids, points, nfreq); points x fastest, then z, then y."""
        fr, ng, npts, _ = self.info()
        buf = np.zeros(max(2 * ng * npts * fr.size, 2))
        if ng * npts * fr.size > 0:
            _check(load_library().chiml_sim_freq_fields(self._sim._h, self.index, _dp(buf), buf.size))
        n = ng * npts * fr.size
        return (buf[0:2 * n:2] + 1j * buf[1:2 * n:2]).reshape(ng, npts, fr.size)

    def incident(self):
        """Incident values of the table, (nincd, nfreq) complex (zero without
        a TFSF surface)."""
        L = load_library()
        ni = ctypes.c_int()
        _check(L.chiml_sim_freq_incd(self._sim._h, self.index, None, 0, ctypes.byref(ni)))
        nf = self.info()[0].size
        buf = np.zeros(2 * ni.value * nf)
        _check(L.chiml_sim_freq_incd(self._sim._h, self.index, _dp(buf), buf.size, ctypes.byref(ni)))
        return (buf[0::2] + 1j * buf[1::2]).reshape(ni.value, nf)

    def values(self):
        """The table's values, shape (grids, nfreq)."""
        fr, ng, _, _ = self.info()
        buf = np.zeros(2 * ng * fr.size)
        _check(load_library().chiml_sim_freq_values(self._sim._h, self.index, _dp(buf), buf.size))
        return (buf[0::2] + 1j * buf[1::2]).reshape(ng, fr.size)


class Flux:
    """A DFT flux box (parallelFluxDTC)."""

    def __init__(self, sim, index):
        self._sim = sim
        self.index = index

    def getFlux(self):
        """Total complex flux per frequency (sum over faces, Simpson's rule)."""
        L = load_library()
        nf = ctypes.c_int()
        _check(L.chiml_sim_flux(self._sim._h, self.index, None, 0, ctypes.byref(nf)))
        buf = np.zeros(2 * nf.value)
        _check(L.chiml_sim_flux(self._sim._h, self.index, _dp(buf), buf.size, ctypes.byref(nf)))
        return buf[0::2] + 1j * buf[1::2]

    def getFluxIncd(self):
        """Incident flux per frequency from the TFSF incident records (the
        incident columns of the table; zero without a TFSF surface)."""
        L = load_library()
        nf = ctypes.c_int()
        _check(L.chiml_sim_flux_incd(self._sim._h, self.index, None, 0, ctypes.byref(nf)))
        buf = np.zeros(2 * nf.value)
        _check(L.chiml_sim_flux_incd(self._sim._h, self.index, _dp(buf), buf.size, ctypes.byref(nf)))
        return buf[0::2] + 1j * buf[1::2]


class QE:
    """A Maxwell-Liouville emitter group (parallelQEBase)."""

    def __init__(self, sim, index):
        self._sim = sim
        self.index = index

    def population(self, j):
        """Population detector j series (QEPopDtc): complex per output step."""
        L = load_library()
        ns = ctypes.c_size_t()
        _check(L.chiml_sim_pops(self._sim._h, self.index, j, None, 0, ctypes.byref(ns)))
        buf = np.zeros(2 * ns.value)
        _check(L.chiml_sim_pops(self._sim._h, self.index, j, _dp(buf), buf.size, ctypes.byref(ns)))
        return buf[0::2] + 1j * buf[1::2]

    def density(self, member, emitter, nlevel):
        L = load_library()
        buf = np.zeros(2 * nlevel * nlevel)
        nl = ctypes.c_int()
        _check(L.chiml_sim_density(self._sim._h, self.index, member, emitter, _dp(buf), ctypes.byref(nl)))
        return (buf[0::2] + 1j * buf[1::2]).reshape(nlevel, nlevel, order="F")


class FDTDField:
    """Device-resident simulation built from a chiML JSON deck.

    Mirrors parallelFDTDFieldReal: ``step()`` advances one step
    (parallelFDTDField.hpp:1228), ``dtcArr()/fluxArr()/qeArr()`` give the
    detectors, flux boxes and emitters (:1766-1781), ``dt`` is the time step.
    """

    def __init__(self, deck_json=None, _handle=None):
        L = load_library()
        if _handle is None:
            h = ctypes.c_void_p()
            _check(L.chiml_sim_create(deck_json.encode(), ctypes.byref(h)))
        else:
            h = _handle
        self._h = h
        n3 = (ctypes.c_int * 3)()
        dt = ctypes.c_double()
        ns = ctypes.c_int()
        prec = ctypes.c_int()
        nd = ctypes.c_int()
        nf = ctypes.c_int()
        nq = ctypes.c_int()
        _check(L.chiml_sim_info(h, n3, ctypes.byref(dt), ctypes.byref(ns), ctypes.byref(prec), ctypes.byref(nd),
                                ctypes.byref(nf), ctypes.byref(nq)))
        self.shape = (n3[2], n3[1], n3[0])
        self.dt = dt.value
        self.nsteps = ns.value
        self.precision = prec.value
        self._ndet, self._nflux, self._nqe = nd.value, nf.value, nq.value

    def close(self):
        if getattr(self, "_h", None):
            load_library().chiml_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        _check(load_library().chiml_sim_step(self._h, int(n)))

    def set_graph(self, steps_per_graph):
        """Replay captured HIP graphs of this many steps (0: plain launches)."""
        _check(load_library().chiml_sim_set_graph(self._h, int(steps_per_graph)))

    def idle(self):
        v = ctypes.c_int()
        _check(load_library().chiml_sim_query(self._h, ctypes.byref(v)))
        return bool(v.value)

    def sync(self):
        _check(load_library().chiml_sim_sync(self._h))

    def run(self):
        """The reference's main loop: ceil(tLim/dt) steps (src/main.cpp:54-57)."""
        self.step(self.nsteps)
        self.sync()

    def toFile(self, prefix=""):
        """Writes every output file the deck names (TXT/BIN detectors, flux
        tables, population files) as the reference's run leaves them;
        prefix goes in front of relative paths."""
        _check(load_library().chiml_sim_write_outputs(self._h, str(prefix).encode()))

    def field(self, comp):
        if isinstance(comp, str):
            comp = FIELD_NAMES[comp]
        out = np.zeros(int(np.prod(self.shape)))
        _check(load_library().chiml_sim_field(self._h, int(comp), _dp(out)))
        return out.reshape(self.shape)

    def E(self, c):
        return self.field(c)

    def incd(self, comp):
        """Incident records of component comp (0..5 or 'Ex'..'Hz') from the TFSF
        surfaces: two complex values per step, the point and its offset
        (FF.ExIncd() .. FF.HzIncd(), src/main.cpp:70)."""
        if isinstance(comp, str):
            comp = FIELD_NAMES[comp]
        L = load_library()
        n = ctypes.c_size_t()
        _check(L.chiml_sim_incd(self._h, int(comp), None, 0, ctypes.byref(n)))
        buf = np.zeros(2 * n.value)
        _check(L.chiml_sim_incd(self._h, int(comp), _dp(buf), buf.size, ctypes.byref(n)))
        return buf[0::2] + 1j * buf[1::2]

    def H(self, c):
        return self.field(3 + c)

    def dtcArr(self):
        return [Detector(self, i) for i in range(self._ndet)]

    def dtcFreqArr(self):
        n = ctypes.c_int()
        _check(load_library().chiml_sim_num_freq_detectors(self._h, ctypes.byref(n)))
        return [FreqDetector(self, i) for i in range(n.value)]

    def fluxArr(self):
        return [Flux(self, i) for i in range(self._nflux)]

    def qeArr(self):
        return [QE(self, i) for i in range(self._nqe)]

    def cost(self):
        """(cells, algorithmic bytes/step of update_h, update_e, emitter state)."""
        vals = [ctypes.c_double() for _ in range(4)]
        _check(load_library().chiml_sim_cost(self._h, *[ctypes.byref(v) for v in vals]))
        return tuple(v.value for v in vals)

    def kernel_bytes(self, kernel):
        """Algorithmic HBM bytes per launch of a kernel (0 if unknown)."""
        b = ctypes.c_double()
        _check(load_library().chiml_sim_kernel_bytes(self._h, kernel.encode(), ctypes.byref(b)))
        return b.value

    def profile(self, enable=True):
        _check(load_library().chiml_sim_profile(self._h, 1 if enable else 0))

    def profile_only(self, kernel):
        """Times only the launches of one kernel (two HIP events per launch)."""
        _check(load_library().chiml_sim_profile_only(self._h, kernel.encode()))

    def profile_read(self, kernel):
        ms = ctypes.c_double()
        n = ctypes.c_int()
        _check(load_library().chiml_sim_profile_read(self._h, kernel.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def footprint(self):
        b = ctypes.c_size_t()
        _check(load_library().chiml_sim_footprint(self._h, ctypes.byref(b)))
        return b.value


def step_batch(sims, nsteps, noalloc=False):
    """Advances every simulation of `sims` by nsteps with shared launches
    (chiml_sim_step_batch; noalloc: chiml_sim_step_batch_noalloc, at most 64
    simulations, no device allocation).  Asynchronous, like FDTDField.step."""
    L = load_library()
    arr = (ctypes.c_void_p * len(sims))(*[sim._h for sim in sims])
    fn = L.chiml_sim_step_batch_noalloc if noalloc else L.chiml_sim_step_batch
    _check(fn(arr, len(sims), int(nsteps)))


def estimate_bytes(deck_json):
    b = ctypes.c_size_t()
    _check(load_library().chiml_sim_estimate_bytes(deck_json.encode(), ctypes.byref(b)))
    return b.value


class Server:
    """Sweep server (include/chiml/chiml_server.h): continuous batching of
    device-resident simulations admitted by HBM footprint."""

    QUEUED, BUILDING, RUNNING, DONE, FAILED = range(5)

    def __init__(self, hbm_budget, max_active=16, chunk_steps=64, graph_steps=0, batch=False):
        L = load_library()
        h = ctypes.c_void_p()
        _check(L.chiml_server_create(int(hbm_budget), int(max_active), int(chunk_steps), int(graph_steps),
                                     ctypes.byref(h)))
        self._h = h
        if batch:
            _check(L.chiml_server_set_batch(h, 1))

    def submit(self, deck_json):
        i = ctypes.c_longlong()
        _check(load_library().chiml_server_submit(self._h, deck_json.encode(), ctypes.byref(i)))
        return i.value

    def poll(self, rid):
        """(state, admitted_s, finished_s) -- seconds since submission, -1 if not yet."""
        st = ctypes.c_int()
        a = ctypes.c_double()
        f = ctypes.c_double()
        _check(load_library().chiml_server_poll(self._h, rid, ctypes.byref(st), ctypes.byref(a), ctypes.byref(f)))
        return st.value, a.value, f.value

    def take(self, rid):
        """The finished simulation as an FDTDField (the caller closes it)."""
        h = ctypes.c_void_p()
        _check(load_library().chiml_server_take(self._h, rid, ctypes.byref(h)))
        return FDTDField(_handle=h)

    def error(self, rid):
        buf = ctypes.create_string_buffer(1024)
        _check(load_library().chiml_server_error(self._h, rid, buf, 1024))
        return buf.value.decode()

    def drain(self):
        _check(load_library().chiml_server_drain(self._h))

    def stats(self):
        cu = ctypes.c_double()
        held = ctypes.c_size_t()
        act = ctypes.c_int()
        q = ctypes.c_int()
        peak = ctypes.c_int()
        _check(load_library().chiml_server_stats(self._h, ctypes.byref(cu), ctypes.byref(held), ctypes.byref(act),
                                                 ctypes.byref(q), ctypes.byref(peak)))
        return {"cell_updates": cu.value, "bytes_held": held.value, "active": act.value, "queued": q.value,
                "peak_active": peak.value}

    def close(self):
        if getattr(self, "_h", None):
            load_library().chiml_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
