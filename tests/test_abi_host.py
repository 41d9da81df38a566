"""C ABI and host-logic checks that need no GPU.

* libchiml_amd.so loads and exports every function include/chiml/*.h declares;
* deck parsing (the reference's JSON schema, src/INPUTS/parallelInputs.cpp)
  rejects what the time-stepping path does not support with the reference's
  own messages where it has one;
* the admission estimate is monotone in grid size and precision.
"""
import ctypes
import json
import os
import re

import pytest

import chiml_amd
from chiml_amd import decks

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include", "chiml")

DECL = re.compile(r"^\s*(?:const\s+)?(?:int|void|char|double|size_t)\s*\*?\s*(chiml_\w+)\s*\(", re.M)


def declared_functions():
    names = []
    for fn in sorted(os.listdir(INCLUDE)):
        if fn.endswith(".h"):
            with open(os.path.join(INCLUDE, fn)) as f:
                names += DECL.findall(f.read())
    return names


def test_headers_declare_the_abi():
    names = declared_functions()
    for required in ("chiml_dev_create", "chiml_dev_step", "chiml_dev_sync", "chiml_dev_read_field",
                     "chiml_dev_read_points", "chiml_dev_read_dft", "chiml_dev_read_pops", "chiml_dev_last_error",
                     "chiml_dev_destroy", "chiml_dev_step_batch", "chiml_sim_create", "chiml_sim_step"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = chiml_amd.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def _estimate(deck):
    return chiml_amd.estimate_bytes(json.dumps(deck))


def _base(n3=(32, 32, 1), precision="fp32"):
    return decks.empty_deck(n3, 0.05, 10, 4, precision=precision)


def test_estimate_monotone():
    a = _estimate(_base((32, 32, 32)))
    b = _estimate(_base((48, 48, 48)))
    c = _estimate(_base((48, 48, 48), "fp64"))
    assert 0 < a < b < c


@pytest.mark.parametrize("mutate,message", [
    (lambda d: d["CompCell"].__setitem__("PBC", True), "periodic"),
    (lambda d: d["CompCell"].__setitem__("cplxFields", True), "complex fields"),
    (lambda d: d["TFSF"].append({"size": [1, 1, 0], "loc": [0, 0, 0], "theta": 200.0}), "theta value for the TFSF"),
    (lambda d: d["TFSF"].append({"size": [0.5, 0.5, 0], "loc": [0, 0, 0], "m": [0, 0, 0]}), "direction of propgation"),
    (lambda d: d["TFSF"].append({"size": [0.5, 0.5, 0], "loc": [9.0, 0, 0], "m": [0, 1, 0]}), "outside the FDTD cell"),
    (lambda d: d["GPU"].__setitem__("precision", "bf16"), "precision"),
    (lambda d: d["CompCell"].__setitem__("dt", 1.0), "stable time step"),
    (lambda d: d["PML"].__setitem__("thickness", [5.0, 5.0, 0.0]), "PML size is larger"),
    (lambda d: d["CompCell"].__setitem__("pol", "Hz"), "TMz"),
    (lambda d: d["PML"].__setitem__("thickness", [0.2, 0.0, 0.0]), "CPML on some axes only"),
    (lambda d: d["SourceList"].append({"pol": "Ez", "size": [0, 0, 0], "loc": [0, 0, 0],
                                       "PulseList": [{"profile": "sinc", "fcen": 1.0}]}), "Pulse profile undefined"),
    (lambda d: d["SourceList"].append({"pol": "Ez", "size": [0, 0, 0], "loc": [50.0, 0, 0],
                                       "PulseList": [decks.gaussian_pulse(1.0, 1.0, 3.0, 1.0)]}), "outside the FDTD Cell"),
    (lambda d: d["DetectorList"].append({"type": "Ez", "dtc_class": "cout", "fname": "x", "size": [0, 0, 0],
                                         "loc": [50.0, 0, 0]}), "detector is outside"),
    (lambda d: d["ObjectList"].append({"shape": "sphere", "material": "custom", "eps": 2.0, "pols": [],
                                       "tellegen": 0.1, "radius": 0.2, "loc": [0, 0, 0], "Basis_Set": []}),
     "chiral"),
    (lambda d: d["ObjectList"].append({"shape": "sphere", "material": "unobtainium", "eps": 2.0, "pols": [],
                                       "radius": 0.2, "loc": [0, 0, 0], "Basis_Set": []}), "not supported"),
    (lambda d: d["FluxList"].append({"name": "f", "size": [0.4, 0.4, 0], "loc": [0, 0, 0], "fcen": 1.0,
                                     "fwidth": 0.5, "nfreq": 3, "save": True}), "save/load"),
    (lambda d: d["FluxList"].append({"name": "f", "size": [0.4, 0.4, 0], "loc": [0, 0, 0], "fcen": 1.0,
                                     "fwidth": 0.5, "nfreq": 3, "load": True}), "without a valid path"),
])
def test_deck_errors(mutate, message):
    deck = _base()
    mutate(deck)
    with pytest.raises(chiml_amd.ChimlError, match=message):
        _estimate(deck)


def test_magnetic_material_in_the_cpml_is_rejected():
    """mu != 1 on an H point inside the CPML takes the reference's B-field
    path, which is not built: the host rejects it before any device call."""
    deck = _base()
    deck["ObjectList"].append({"shape": "block", "material": "custom", "eps": 1.0, "mu": 2.0, "pols": [],
                               "size": [1.5, 0.4, 0.0], "loc": [0, 0, 0], "Basis_Set": []})
    with pytest.raises(chiml_amd.ChimlError, match="inside the PML"):
        chiml_amd.FDTDField(json.dumps(deck))


def _oblique(pol, size, loc, phi=30.0):
    return {"pol": pol, "size": size, "loc": loc, "phi": phi, "PulseList": [decks.gaussian_pulse(1.0, 1.0, 3.0, 1.0)]}


@pytest.mark.parametrize("n3,src,message", [
    ((32, 32, 1), _oblique("Hx", [1.4, 0.1, 0], [0, 0.7, 0]), "outside the grid's physical points"),
    ((32, 32, 1), _oblique("Ex", [0.2, 0.2, 0], [0, 0, 0]), "does not exist in 2D TMz"),
])
def test_source_table_errors(n3, src, message):
    """Errors raised while the host builds the source tables, before any
    device call: oblique points past the grid's physical points and
    components a 2D TMz grid lacks."""
    deck = _base(n3)
    deck["SourceList"].append(src)
    with pytest.raises(chiml_amd.ChimlError, match=message):
        chiml_amd.FDTDField(json.dumps(deck))


def test_error_message_is_per_thread_and_cleared_by_success():
    deck = _base()
    deck["CompCell"]["PBC"] = True
    with pytest.raises(chiml_amd.ChimlError):
        _estimate(deck)
    assert "periodic" in chiml_amd.load_library().chiml_sim_last_error().decode()
    assert _estimate(_base()) > 0


def test_oracle_is_test_infrastructure_only():
    """Nothing in the product package loads the oracle (build.py only compiles it)."""
    pkg = os.path.join(ROOT, "chiml_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".cpp", ".hpp", ".hip", ".h")) and fn != "build.py":
                with open(os.path.join(dirpath, fn), errors="replace") as f:
                    text = f.read().lower()
                assert "oracle" not in text, os.path.join(dirpath, fn)
    with open(os.path.join(ROOT, "oracle", "oracle.cpp")) as f:
        assert "test infrastructure only" in f.read(400)


def _build_c_consumer(out_dir):
    """Compiles tests/c/abi_consumer.c, a C99 program that includes every
    public header and builds a chiml_problem_t by hand, against the library."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    lib_dir = os.path.dirname(os.path.abspath(chiml_amd.load_library()._name))
    exe = os.path.join(str(out_dir), "abi_consumer")
    cmd = [gcc, "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "c", "abi_consumer.c"), "-L", lib_dir, "-lchiml_amd",
           f"-Wl,-rpath,{lib_dir}", "-lm", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_c_consumer_compiles_strict_c99_and_links(tmp_path):
    import subprocess
    exe = _build_c_consumer(tmp_path)
    out = subprocess.run([exe, "--link-only"], check=True, capture_output=True, text=True).stdout
    assert "linked 16 entry points" in out
