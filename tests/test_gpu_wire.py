"""GPU tests of the page-query endpoints end to end: JSON request -> request
loop on the device -> JSON pages, against direct scans and the oracle.

Follows the reference's endpoint tests (crates/morphogen-server/tests/
network_api.rs:1666-1866): single queries return 3 pages, batch results come
back in request order and equal the single-query answers; plus two-party
reconstruction of the target page, epoch ids after updates
(api/mod.rs:557-622 consistency), and status codes that need the device.
"""
import json
import os
import urllib.request

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
THREADS = max(1, min(16, len(os.sched_getaffinity(0))))

N, W, D = 1 << 12, 4096, 12  # 4096-byte pages, as the reference's page matrix


@pytest.fixture(scope="module")
def mg():
    import morphogenesis_amd
    return morphogenesis_amd


@pytest.fixture(scope="module")
def setup(mg):
    scanner = mg.GpuScanner(0)
    rows = oracle.fill_rows(71, 0, N, W)
    m = mg.GpuPageMatrix.new(scanner, rows)
    srv = mg.GpuServer(scanner, m, segments=2)
    api = mg.WireApi(srv)
    yield scanner, m, srv, api, rows
    srv.close()
    m.close()
    scanner.close()


def keys_for(targets, party, seed):
    return [oracle.keygen_bytes(D, t, seed=seed + i)[party] for i, t in enumerate(targets)]


def post(api, path, body):
    st, out, err = api.handle("POST", path, "application/json", json.dumps(body).encode())
    return st, (json.loads(out) if st == 200 else err)


def pages_of(doc):
    return np.stack([np.frombuffer(bytes.fromhex(p[2:]), dtype=np.uint8) for p in doc["pages"]])


def test_single_query_pages_equal_scan_and_oracle(mg, setup):
    scanner, m, srv, api, rows = setup
    keys = keys_for([5, 4095, 1234], 0, 10)
    st, doc = post(api, mg.WIRE_PATH, {"keys": ["0x" + k.hex() for k in keys]})
    assert st == 200, doc
    assert doc["epoch_id"] == 0 and len(doc["pages"]) == 3
    assert all(len(p) == 2 + 2 * W for p in doc["pages"])
    got = pages_of(doc)
    assert np.array_equal(got, scanner.scan_keys(m, keys))
    assert np.array_equal(got, oracle.scan_tiled(keys, rows, THREADS))


def test_two_parties_reconstruct_the_pages(mg, setup):
    _, _, _, api, rows = setup
    targets = [0, 77, 4095]
    docs = []
    for party in (0, 1):
        keys = keys_for(targets, party, 20)
        st, doc = post(api, mg.WIRE_PATH, {"keys": ["0x" + k.hex() for k in keys]})
        assert st == 200
        docs.append(pages_of(doc))
    assert np.array_equal(docs[0] ^ docs[1], rows[targets])


def test_batch_results_in_request_order_equal_singles(mg, setup):
    scanner, m, _, api, rows = setup
    queries = [keys_for([42, 11, 3000], 0, 30), keys_for([11, 42, 7], 1, 40)]
    st, doc = post(api, mg.WIRE_BATCH_PATH,
                   {"queries": [{"keys": ["0x" + k.hex() for k in q]} for q in queries]})
    assert st == 200
    assert len(doc["results"]) == 2
    for q, res in zip(queries, doc["results"]):
        st1, single = post(api, mg.WIRE_PATH, {"keys": ["0x" + k.hex() for k in q]})
        assert st1 == 200
        assert res["pages"] == single["pages"]
        assert np.array_equal(pages_of(res), oracle.scan_tiled(q, rows, THREADS))


def test_full_batch_of_32(mg, setup):
    scanner, m, _, api, _ = setup
    rng = np.random.default_rng(3)
    queries = [keys_for([int(x) for x in rng.integers(0, N, 3)], q % 2, 100 * q)
               for q in range(32)]
    body = {"queries": [{"keys": ["0x" + k.hex() for k in q]} for q in queries]}
    assert len(json.dumps(body)) <= mg.WIRE_MAX_BODY
    st, doc = post(api, mg.WIRE_BATCH_PATH, body)
    assert st == 200
    flat = [k for q in queries for k in q]
    want = scanner.scan_keys(m, flat).reshape(32, 3, W)
    for q in range(32):
        assert np.array_equal(pages_of(doc["results"][q]), want[q]), q


def test_epoch_id_and_pages_follow_updates(mg, setup):
    _, _, srv, api, rows = setup
    idx = np.array([100, 200, 100], dtype=np.uint64)
    diffs = np.random.default_rng(4).integers(0, 256, size=(3, W), dtype=np.uint8)
    srv.update(idx, diffs, 7)
    state = rows.copy()
    for i, dl in zip(idx, diffs):
        state[int(i)] ^= dl
    keys = keys_for([100, 200, 300], 0, 50)
    st, doc = post(api, mg.WIRE_PATH, {"keys": ["0x" + k.hex() for k in keys]})
    assert st == 200 and doc["epoch_id"] == 7
    assert np.array_equal(pages_of(doc), oracle.scan_tiled(keys, state, THREADS))
    # The same diffs again cancel (XOR), so later tests see the base rows.
    srv.update(idx, diffs, 8)
    st, doc = post(api, mg.WIRE_PATH, {"keys": ["0x" + k.hex() for k in keys]})
    assert st == 200 and doc["epoch_id"] == 8
    assert np.array_equal(pages_of(doc), oracle.scan_tiled(keys, rows, THREADS))


def test_statuses_that_need_the_loop(mg, setup):
    scanner, m, _, api, _ = setup
    # A key of another domain parses, but the loop serves d = 12 only.
    k11 = oracle.keygen_bytes(11, 1, seed=1)[0]
    assert post(api, mg.WIRE_PATH, {"keys": ["0x" + k11.hex()] * 3})[0] == 400
    # More keys than the loop's slots can ever hold.
    small = mg.GpuServer(scanner, m, segments=1, max_active_keys=6)
    try:
        api6 = mg.WireApi(small)
        q = ["0x" + k.hex() for k in keys_for([1, 2, 3], 0, 60)]
        assert post(api6, mg.WIRE_PATH, {"keys": q})[0] == 200
        assert post(api6, mg.WIRE_BATCH_PATH, {"queries": [{"keys": q}] * 3})[0] == 503
    finally:
        small.close()


def test_client_flow_through_the_endpoint(mg, setup):
    # generate_gpu_page_query -> one request per server -> aggregate -> rows
    # (crates/morphogen-client/src/gpu.rs:17-49, page.rs:67-162). Both
    # parties' requests go to the same loop here; each sees only its own keys.
    _, _, _, api, rows = setup
    for acct in (b"0xd8dA6BF26964aF9D7eEd9e03E53415D37aA96045", b"test_account"):
        q = mg.generate_gpu_page_query(acct, N, rng_seed=17)
        answers = []
        for keys in (q.keys_a, q.keys_b):
            st, out, _ = api.handle("POST", mg.WIRE_PATH, "application/json",
                                    mg.page_query_body(keys))
            assert st == 200
            answers.append(mg.parse_page_response(out))
        epoch, pages = mg.aggregate_gpu_page_responses(*answers[0], *answers[1])
        got = mg.extract_rows_from_gpu_pages(pages, q.page_addresses)
        for i, a in enumerate(q.page_addresses):
            assert pages[i] == rows[a.page_index].tobytes()
            assert got[i] == rows[a.page_index][256 * a.row_offset:256 * (a.row_offset + 1)].tobytes()


def test_http_round_trip(mg, setup):
    from morphogenesis_amd.http_api import serve
    scanner, m, _, api, _ = setup
    httpd, th = serve(api)
    try:
        keys = keys_for([9, 99, 999], 1, 70)
        body = json.dumps({"keys": ["0x" + k.hex() for k in keys]}).encode()
        req = urllib.request.Request(f"http://127.0.0.1:{httpd.server_address[1]}{mg.WIRE_PATH}",
                                     data=body, method="POST",
                                     headers={"Content-Type": "application/json"})
        with urllib.request.urlopen(req, timeout=60) as r:
            assert r.status == 200
            assert r.headers["Content-Type"] == "application/json"
            doc = json.loads(r.read())
        assert np.array_equal(pages_of(doc), scanner.scan_keys(m, keys))
    finally:
        httpd.shutdown()
        httpd.server_close()
        th.join(timeout=10)
