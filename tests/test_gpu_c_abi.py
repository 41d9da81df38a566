"""The C ABI driven from plain C: tests/c/abi_consumer.c builds a
chiml_problem_t by hand and checks the first two Yee steps of a point source
against their closed form (Ez = delta, then delta + dt^2 Lap_h delta)."""
import subprocess

import pytest

from tests.test_abi_host import _build_c_consumer

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [64, 32])
def test_c_consumer_known_answer(precision, tmp_path):
    exe = _build_c_consumer(tmp_path)
    print(f"abi_consumer {precision}", flush=True)
    res = subprocess.run([exe, str(precision)], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0
    assert res.stdout.strip().endswith("OK")
