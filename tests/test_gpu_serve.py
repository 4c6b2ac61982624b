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


@pytest.mark.parametrize("bits", [1024, 2048, 3072, 4096])
def test_server_crt_recombination(handle, bits):
    """Up to 4096-bit keys the kind-1 recombination (mp*cp + mq*cq) mod n is
    one Montgomery reduction against the resident cp*R mod n and cq*R mod n.
    Exact against the oracle with the keys' CRT coefficients (keys 0-2) and
    with arbitrary ones, for which the result is still that linear
    combination mod n: all ones (above n), zero and n - 1, and random words."""
    nkeys, nreq = 6, 64
    keys = make_keys(nkeys, bits, 900 + bits)
    n32 = bits // 32
    keys["cp"][3] = 0xFFFFFFFF
    keys["cq"][3] = 0xFFFFFFFF
    keys["cp"][4] = 0
    keys["cq"][4] = keys["n"][4]
    keys["cq"][4, 0] ^= 1                      # n - 1 (n is odd)
    keys["cp"][5] = inputs.uniform(bits + 7, 1, n32)[0]
    keys["cq"][5] = inputs.uniform(bits + 8, 1, n32)[0]
    rng = np.random.default_rng(bits + 3)
    kinds = (rng.random(nreq) < 0.8).astype(np.uint32)
    key_idx = (np.arange(nreq) % nkeys).astype(np.uint32)
    bases = inputs.uniform(bits + 5, nreq, n32)
    nmat = keys["n"][key_idx]
    bases[:, -1] = (bases[:, -1].astype(np.uint64) % nmat[:, -1].astype(np.uint64)).astype(np.uint32)
    exps = inputs.uniform(bits + 6, nreq, n32)
    want = O.serving_expected(keys, kinds, key_idx, bases, exps, threads=THREADS)
    srv = Server(handle, keys, bits, streams=2, max_batch=16)
    try:
        res, lat, wall = srv.run_poisson(1e6, kinds, key_idx, bases, exps, seed=11)
    finally:
        srv.close()
    assert_rows_equal(res, want, "%d-bit CRT recombination" % bits)


@pytest.mark.parametrize("bits,nkeys,nreq", [(8192, 4, 16), (16384, 2, 6)])
def test_server_wide_keys(handle, bits, nkeys, nreq):
    """Keys above 4096 bits keep the multiply, add and division chain: its
    final division by n (256 limbs at 8192 bits) and, at 16384 bits, also the
    division by p and q (256 limbs) go past the register division and use
    the slot's division workspace."""
    keys = make_keys(nkeys, bits, 500 + bits)
    rng = np.random.default_rng(bits)
    kinds = np.array([i % 2 for i in range(nreq)], np.uint32)
    key_idx = rng.integers(0, nkeys, nreq).astype(np.uint32)
    bases = inputs.uniform(bits + 1, nreq, bits // 32)
    nmat = keys["n"][key_idx]
    bases[:, -1] = (bases[:, -1].astype(np.uint64) % nmat[:, -1].astype(np.uint64)).astype(np.uint32)
    exps = inputs.uniform(bits + 2, nreq, bits // 32)
    want = O.serving_expected(keys, kinds, key_idx, bases, exps, threads=THREADS)
    srv = Server(handle, keys, bits, streams=2, max_batch=8)
    try:
        res, lat, wall = srv.run_poisson(1e6, kinds, key_idx, bases, exps, seed=9)
    finally:
        srv.close()
    assert_rows_equal(res, want, "served %d-bit requests" % bits)


def test_server_failed_batch_is_reported(handle, monkeypatch):
    """A batch whose launch fails part-way (forced from the host after its
    powm requests were enqueued) is not published: the server waits for what
    was enqueued, reports xmpErrorCuda from Drain/RunPoisson, admits nothing
    more, and shuts down cleanly."""
    import xmp_amd
    keys = make_keys(8, BITS, 78)
    kinds, key_idx, bases, exps = _requests(keys, 200, 6)
    monkeypatch.setenv("XMP_SERVER_FAIL_BATCH", "2")
    srv = Server(handle, keys, BITS, streams=2, max_batch=16)
    monkeypatch.delenv("XMP_SERVER_FAIL_BATCH")
    try:
        with pytest.raises(xmp_amd.XmpError) as err:
            srv.run_poisson(1e6, kinds, key_idx, bases, exps, seed=5)
        assert err.value.code == xmp_amd.CUDA_ERROR
    finally:
        srv.close()
    # a server created afterwards (no failure forced) serves normally
    want = O.serving_expected(keys, kinds[:50], key_idx[:50], bases[:50], exps[:50], threads=THREADS)
    srv = Server(handle, keys, BITS, streams=2, max_batch=16)
    try:
        res, lat, wall = srv.run_poisson(1e6, kinds[:50], key_idx[:50], bases[:50], exps[:50], seed=5)
    finally:
        srv.close()
    assert_rows_equal(res, want, "served after a failed server")


def test_server_full_scale(handle):
    """The served configuration the benchmark times: 1,024 resident 2048-bit
    keys on 4 slots of 8,192 requests.  20,000 Poisson requests (half single
    2048-bit powm, half CRT pairs) at 400 K req/s; GMP / Python int agree on
    2,048 of them spread over the run, and 128 of those served alone (one in
    flight) give the same bits as in the mix."""
    keys = make_keys(1024, BITS, 99)
    n = 20000
    kinds, key_idx, bases, exps = _requests(keys, n, 11)
    srv = Server(handle, keys, BITS, streams=4, max_batch=8192)
    try:
        res, lat, wall = srv.run_poisson(400000.0, kinds, key_idx, bases, exps, seed=12)
        assert (lat > 0).all()
        rows = np.unique(np.linspace(0, n - 1, num=2048).astype(int))
        want = O.serving_expected(keys, kinds[rows], key_idx[rows], bases[rows], exps[rows], threads=THREADS)
        assert_rows_equal(res[rows], want, "served in the mix")
        alone_rows = rows[:: rows.size // 128][:128]
        for i in alone_rows:
            srv.submit(int(kinds[i]), int(key_idx[i]), bases[i], exps[i], int(i))
            srv.drain()
            tags, got, _ = srv.poll()
            assert tags.tolist() == [i]
            assert np.array_equal(got[0], res[i]), "request %d: alone differs from the mix" % i
    finally:
        srv.close()


def test_server_hbm_budget_binds_under_load(handle):
    """The HBM budget: by default half the device memory free at creation
    (it does not bind at these sizes); a budget of 100 requests' bytes shrinks
    two slots of 256 to 50 each, and under arrivals far above capacity the
    scheduler fills it but never exceeds it.  Every result is still exact."""
    keys = make_keys(16, BITS, 79)
    kinds, key_idx, bases, exps = _requests(keys, 2000, 13)
    want = O.serving_expected(keys, kinds, key_idx, bases, exps, threads=THREADS)
    srv = Server(handle, keys, BITS, streams=2, max_batch=256)
    try:
        st = srv.stats()
        assert st["hbm_budget"] > 0 and st["slot_capacity"] == 256 and st["max_inflight"] == 512
        bpr = st["bytes_per_request"]
    finally:
        srv.close()
    srv = Server(handle, keys, BITS, streams=2, max_batch=256, hbm_budget=100 * bpr)
    try:
        st = srv.stats()
        assert st["slot_capacity"] == 50 and 90 <= st["max_inflight"] <= 100, st
        res, lat, wall = srv.run_poisson(1e6, kinds, key_idx, bases, exps, seed=14)
        st = srv.stats()
        assert st["peak_inflight"] == st["max_inflight"], st
        assert st["peak_batch"] <= st["slot_capacity"], st
        assert st["batches"] >= 2000 // 50
    finally:
        srv.close()
    assert_rows_equal(res, want, "served within a 100-request budget")


def test_server_reserved_cus(handle):
    """Short kernels on CU-masked streams (reserved_cus): the same results
    alone and in a Poisson mix."""
    keys = make_keys(16, BITS, 80)
    kinds, key_idx, bases, exps = _requests(keys, 400, 15)
    want = O.serving_expected(keys, kinds, key_idx, bases, exps, threads=THREADS)
    srv = Server(handle, keys, BITS, streams=2, max_batch=128, reserved_cus=8)
    try:
        for i in range(8):
            srv.submit(int(kinds[i]), int(key_idx[i]), bases[i], exps[i], i)
            srv.drain()
            tags, res1, _ = srv.poll()
            assert tags.tolist() == [i]
            assert np.array_equal(res1[0], want[i])
        res, lat, wall = srv.run_poisson(300000.0, kinds, key_idx, bases, exps, seed=16)
    finally:
        srv.close()
    assert_rows_equal(res, want, "served with reserved CUs")
