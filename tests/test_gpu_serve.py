"""GPU: the continuous-batching server.

Every served result equals GMP's (kind 0: base^exp mod n; kind 1: the CRT
pair recombined as in samples/crt_rsa/rsa.cc:205-219), and each request's
result is the same served alone (one request in flight) and in a Poisson mix
where batches overlap on the device.  Bit-exact.
"""
import numpy as np
import pytest

from helpers import O, THREADS, assert_rows_equal, inputs
from xmp_amd.ext import Server, make_keys

pytestmark = pytest.mark.gpu

BITS = 2048


def _requests(keys, n, seed):
    rng = np.random.default_rng(seed)
    kinds = (rng.random(n) < 0.5).astype(np.uint32)
    key_idx = rng.integers(0, keys["n"].shape[0], n).astype(np.uint32)
    bases = inputs.uniform(seed + 1, n, BITS // 32)
    nmat = keys["n"][key_idx]
    bases[:, -1] = (bases[:, -1].astype(np.uint64) % np.maximum(nmat[:, -1].astype(np.uint64), 1)).astype(np.uint32)
    exps = inputs.uniform(seed + 2, n, BITS // 32)
    return kinds, key_idx, bases, exps


def test_server_alone_and_in_mix(handle):
    keys = make_keys(32, BITS, 77)
    kinds, key_idx, bases, exps = _requests(keys, 600, 5)
    want = O.serving_expected(keys, kinds, key_idx, bases, exps, threads=THREADS)
    srv = Server(handle, keys, BITS, streams=4, max_batch=64)
    try:
        # alone: one request at a time
        alone = np.zeros_like(want)
        for i in range(40):
            srv.submit(int(kinds[i]), int(key_idx[i]), bases[i], exps[i], i)
            srv.drain()
            tags, res, lat = srv.poll()
            assert tags.tolist() == [i]
            alone[i] = res[0]
        assert_rows_equal(alone[:40], want[:40], "served alone")
        # in the mix: Poisson arrivals fast enough that batches overlap
        res, lat, wall = srv.run_poisson(200000.0, kinds, key_idx, bases, exps, seed=3)
        assert_rows_equal(res, want, "served in the mix")
        assert np.array_equal(res[:40], alone[:40])
        assert (lat > 0).all()
        # a small HBM budget still serves everything (admission waits for memory)
    finally:
        srv.close()
    srv = Server(handle, keys, BITS, streams=2, max_batch=64, hbm_budget=1)
    try:
        res, lat, wall = srv.run_poisson(100000.0, kinds[:100], key_idx[:100], bases[:100], exps[:100], seed=4)
        assert_rows_equal(res, want[:100], "one request admitted at a time")
    finally:
        srv.close()
