"""CPU tests of the GPU page-query wire API (mg_wire_parse / mg_wire_format /
mg_wire_handle without a device, and the HTTP transport).

The cases follow the reference's own endpoint tests:
crates/morphogen-server/tests/network_api.rs:1666-1866 (page length 8194 hex
chars for a 4096-byte page, wrong key count -> 400, empty batch -> 400,
batch results in request order), src/network/api/tests.rs:208-269 (domain
bits above the kernel limit -> 400, the limit itself accepted, batch request
and response shapes), :1361 (more than 32 queries -> 400); and the extractor
behaviour of axum's Json (415 / 413 / 400 / 422) with serde_hex.rs:71-84.
"""
import json
import urllib.error
import urllib.request

import numpy as np
import pytest

import oracle

mg = pytest.importorskip("morphogenesis_amd")


def key_hex(d=8, target=42, seed=1, party=0, prefix="0x"):
    return prefix + oracle.keygen_bytes(d, target, seed=seed)[party].hex()


def single_body(keys):
    return json.dumps({"keys": keys})


def batch_body(queries):
    return json.dumps({"queries": [{"keys": q} for q in queries]})


# ------------------------------------------------------------- format -----

def test_single_response_shape_for_4096_byte_pages():
    rng = np.random.default_rng(0)
    pages = rng.integers(0, 256, size=(3, 4096), dtype=np.uint8)
    out = mg.wire_format(42, pages, batch=False)
    doc = json.loads(out)
    assert list(doc) == ["epoch_id", "pages"]
    assert doc["epoch_id"] == 42
    assert len(doc["pages"]) == 3
    for i in range(3):
        assert len(doc["pages"][i]) == 8194  # "0x" + 2 x 4096 (network_api.rs:1714)
        assert doc["pages"][i] == "0x" + pages[i].tobytes().hex()


def test_exact_bytes_of_small_responses():
    pages = np.arange(3 * 16, dtype=np.uint8).reshape(1, 3, 16)
    hexes = ['"0x' + bytes(range(16 * i, 16 * i + 16)).hex() + '"' for i in range(3)]
    single = '{"epoch_id":7,"pages":[' + ",".join(hexes) + "]}"
    assert mg.wire_format(7, pages[0], batch=False) == single.encode()
    batch = '{"epoch_id":18446744073709551615,"results":[{"pages":[' + ",".join(hexes) + "]}]}"
    assert mg.wire_format(2**64 - 1, pages, batch=True) == batch.encode()


def test_batch_response_keeps_query_order():
    rng = np.random.default_rng(1)
    pages = rng.integers(0, 256, size=(5, 3, 32), dtype=np.uint8)
    doc = json.loads(mg.wire_format(3, pages, batch=True))
    assert list(doc) == ["epoch_id", "results"]
    assert len(doc["results"]) == 5
    for q in range(5):
        assert doc["results"][q]["pages"] == ["0x" + pages[q, k].tobytes().hex() for k in range(3)]


# -------------------------------------------------------------- parse -----

def test_valid_single_request_yields_the_key_bytes():
    keys = [key_hex(seed=s) for s in range(3)]
    status, parsed, nq = mg.wire_parse(single_body(keys))
    assert (status, nq) == (200, 1)
    assert parsed == [bytes.fromhex(k[2:]) for k in keys]


def test_hex_prefix_optional_and_case_insensitive():
    k = key_hex(prefix="")
    status, parsed, _ = mg.wire_parse(single_body([k, "0x" + k.upper(), k.upper()]))
    assert status == 200
    assert parsed[0] == parsed[1] == parsed[2] == bytes.fromhex(k)


def test_unknown_fields_escapes_and_sequence_form():
    k = key_hex()
    escaped = "\\u0030" + k[1:]  # "0" is "0"
    body = '{"extra": [1, {"a": null}], "keys": ["%s", "%s", "%s"], "z": -1.5e3}' % (
        escaped, k, k)
    status, parsed, _ = mg.wire_parse(body)
    assert status == 200 and parsed[0] == parsed[1]
    # serde's derived structs also read a one-element sequence.
    status, parsed2, _ = mg.wire_parse(json.dumps([[k, k, k]]))
    assert status == 200 and parsed2 == parsed


def test_valid_batch_request():
    queries = [[key_hex(seed=10 * q + i) for i in range(3)] for q in range(4)]
    status, parsed, nq = mg.wire_parse(batch_body(queries), path=mg.WIRE_BATCH_PATH)
    assert (status, nq) == (200, 4)
    assert parsed == [bytes.fromhex(k[2:]) for q in queries for k in q]


def test_batch_of_32_accepted_and_33_rejected():
    k = key_hex()
    assert mg.wire_parse(batch_body([[k] * 3] * 32), path=mg.WIRE_BATCH_PATH)[0] == 200
    assert mg.wire_parse(batch_body([[k] * 3] * 33), path=mg.WIRE_BATCH_PATH)[0] == 400


@pytest.mark.parametrize("n_keys", [0, 1, 2, 4])
def test_wrong_key_count_is_bad_request(n_keys):
    k = key_hex()
    assert mg.wire_parse(single_body([k] * n_keys))[0] == 400
    body = batch_body([[k] * 3, [k] * n_keys])
    assert mg.wire_parse(body, path=mg.WIRE_BATCH_PATH)[0] == 400


def test_empty_batch_is_bad_request():
    assert mg.wire_parse('{"queries":[]}', path=mg.WIRE_BATCH_PATH)[0] == 400


def test_domain_bits_limit():
    # The kernel's key record holds 32 correction words: d = 32 parses, d = 33 does not.
    k32 = oracle.keygen_bytes(32, 5, seed=3)[0]
    assert mg.wire_parse(single_body(["0x" + k32.hex()] * 3))[0] == 200
    d = 33
    k33 = bytearray(19 + 18 * d + 16)
    k33[18] = d
    assert mg.wire_parse(single_body(["0x" + bytes(k33).hex()] * 3))[0] == 400


def test_key_that_does_not_parse_is_bad_request():
    k = key_hex()
    assert mg.wire_parse(single_body([k, k, k[:-2]]))[0] == 400  # one byte short
    assert mg.wire_parse(single_body([k, k, "0x"]))[0] == 400  # empty key


@pytest.mark.parametrize("body", [
    '{"keys": ["0xzz", "0x00", "0x00"]}',          # bad hex digit
    '{"keys": ["0x123", "0x00", "0x00"]}',          # odd number of digits
    '{"key": []}',                                   # missing field
    '{"keys": "0x00"}',                              # not a sequence
    '{"keys": [1, 2, 3]}',                           # not strings
    '{"keys": null}',
    '{"keys": [], "keys": []}',                      # duplicate field
    '[[], []]',                                      # sequence form, wrong length
    '"keys"',
    '17',
])
def test_wrong_shape_is_unprocessable(body):
    assert mg.wire_parse(body)[0] == 422


@pytest.mark.parametrize("body", [
    "", "{", '{"keys": [}', '{"keys": []} x', '{"keys": ["\\ud800"]}', "{'keys': []}",
    '{"keys": [01]}', '{"keys": ["a\nb"]}', b'{"keys": ["\xff"]}',
])
def test_malformed_json_is_bad_request(body):
    assert mg.wire_parse(body)[0] == 400


def test_batch_shape_errors():
    k = key_hex()
    assert mg.wire_parse('{"queries": {}}', path=mg.WIRE_BATCH_PATH)[0] == 422
    assert mg.wire_parse('{"queries": [{"keys": ["0xq"]}]}', path=mg.WIRE_BATCH_PATH)[0] == 422
    # Deserialization runs before the size check: 33 queries with bad hex -> 422.
    bad = [[k] * 3] * 32 + [["0xq"] * 3]
    assert mg.wire_parse(batch_body(bad), path=mg.WIRE_BATCH_PATH)[0] == 422


def test_content_type_checks():
    body = single_body([key_hex()] * 3)
    assert mg.wire_parse(body, content_type=None)[0] == 415
    assert mg.wire_parse(body, content_type="text/plain")[0] == 415
    assert mg.wire_parse(body, content_type="application/jsonx")[0] == 415
    assert mg.wire_parse(body, content_type="application/json; charset=utf-8")[0] == 200
    assert mg.wire_parse(body, content_type="Application/JSON")[0] == 200
    assert mg.wire_parse(body, content_type="application/vnd.api+json")[0] == 200


def test_body_limit():
    k = key_hex(d=20)
    body = batch_body([[k] * 3] * 32)  # 32 x 3 x 1002 hex digits > 64 KiB
    assert len(body) > mg.WIRE_MAX_BODY
    assert mg.wire_parse(body, path=mg.WIRE_BATCH_PATH)[0] == 413
    assert mg.wire_parse(body, path=mg.WIRE_BATCH_PATH, max_body_bytes=1 << 20)[0] == 200
    # 415 is decided before the body is read.
    assert mg.wire_parse(body, path=mg.WIRE_BATCH_PATH, content_type="text/plain")[0] == 415


def test_routing():
    body = single_body([key_hex()] * 3)
    assert mg.wire_parse(body, path="/query/page")[0] == 404
    assert mg.wire_parse(body, path="/query/page/gpu/")[0] == 404
    assert mg.wire_parse(body, method="GET")[0] == 405
    assert mg.wire_parse(body, method="GET", path=mg.WIRE_BATCH_PATH)[0] == 405


# ------------------------------------------------- handler, no device -----

def test_handler_without_a_loop():
    api = mg.WireApi(None)
    ok = single_body([key_hex()] * 3)
    assert api.handle("POST", mg.WIRE_PATH, "application/json", ok.encode())[:2] == (503, b"")
    st, out, err = api.handle("POST", mg.WIRE_PATH, "application/json", b'{"keys": []}')
    assert (st, out) == (400, b"") and "expected 3" in err
    assert api.handle("DELETE", mg.WIRE_PATH, "application/json", b"")[0] == 405


def test_http_transport_on_localhost():
    from morphogenesis_amd.http_api import serve
    httpd, th = serve(mg.WireApi(None))
    port = httpd.server_address[1]
    try:
        def post(path, body, ctype="application/json", method="POST"):
            req = urllib.request.Request(f"http://127.0.0.1:{port}{path}", data=body,
                                         method=method)
            if ctype:
                req.add_header("Content-Type", ctype)
            try:
                with urllib.request.urlopen(req, timeout=10) as r:
                    return r.status
            except urllib.error.HTTPError as e:
                return e.code

        ok = single_body([key_hex()] * 3).encode()
        assert post(mg.WIRE_PATH, ok) == 503  # valid, but no GPU loop behind it
        assert post(mg.WIRE_PATH, b'{"keys": ["0x00"]}') == 400
        assert post(mg.WIRE_PATH, ok, ctype=None) == 415
        assert post(mg.WIRE_BATCH_PATH, b"x" * (mg.WIRE_MAX_BODY + 10)) == 413
        assert post("/nope", ok) == 404
        assert post(mg.WIRE_PATH, None, method="GET") == 405
    finally:
        httpd.shutdown()
        httpd.server_close()
        th.join(timeout=10)
