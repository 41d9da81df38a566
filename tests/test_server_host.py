"""Sweep-server host logic that runs without a GPU: deck rejection at submit,
clean failure reporting, shutdown with work still queued."""
import json

import pytest

import chiml_amd
from chiml_amd import decks
from tests.conftest import gpu_available


def _deck(n=32):
    return decks.config1_2d(n=n, steps=20, pml=6, cyl_radius_cells=4, src_cell=(10, n // 2), ndet=1,
                            precision="fp32")


def test_submit_rejects_bad_deck():
    s = chiml_amd.Server(1 << 30)
    try:
        d = json.loads(_deck())
        d["CompCell"]["PBC"] = True
        with pytest.raises(chiml_amd.ChimlError, match="periodic"):
            s.submit(json.dumps(d))
    finally:
        s.close()


def test_invalid_server_arguments():
    with pytest.raises(chiml_amd.ChimlError):
        chiml_amd.Server(0)
    with pytest.raises(chiml_amd.ChimlError):
        chiml_amd.Server(1 << 30, max_active=0)


@pytest.mark.skipif(gpu_available(), reason="checks the no-device failure path")
def test_requests_fail_cleanly_without_a_device():
    s = chiml_amd.Server(1 << 30, max_active=2)
    try:
        ids = [s.submit(_deck()) for _ in range(3)]
        s.drain()
        for i in ids:
            st, adm, fin = s.poll(i)
            assert st == chiml_amd.Server.FAILED
            assert adm >= 0 and fin >= adm
            assert s.error(i)
        with pytest.raises(chiml_amd.ChimlError):
            s.take(ids[0])
        assert s.stats()["bytes_held"] == 0
    finally:
        s.close()


def test_unknown_request_id():
    s = chiml_amd.Server(1 << 30)
    try:
        with pytest.raises(chiml_amd.ChimlError, match="unknown"):
            s.poll(12345)
    finally:
        s.close()


def test_destroy_with_queued_work_returns():
    # budget smaller than one request: the first is admitted alone, the rest wait
    s = chiml_amd.Server(1, max_active=1)
    for _ in range(4):
        s.submit(_deck())
    s.close()
