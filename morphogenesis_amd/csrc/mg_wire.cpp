// The reference's GPU page-query wire API over the request loop:
//   POST /query/page/gpu        {"keys":[k0,k1,k2]}       -> {"epoch_id":E,"pages":[p0,p1,p2]}
//   POST /query/page/gpu/batch  {"queries":[{"keys":[..]},..]}
//                               -> {"epoch_id":E,"results":[{"pages":[..]},..]}
// (crates/morphogen-server/src/network/api/mod.rs:544-842 handlers, :886-896
// routes and body limit; dto.rs:126-156 shapes; serde_hex.rs:58-85 hex).
//
// Status codes follow the reference's extractor and handler order:
//   404 unknown path, 405 wrong method (router);
//   415 content type not JSON, 413 body over the limit, 400 malformed JSON,
//   422 JSON of the wrong shape or bad hex (axum's Json extractor, serde);
//   400 batch of 0 or more than 32 queries, a query without exactly 3 keys,
//   a key that does not parse or exceeds the kernel's domain (the handlers,
//   parse_gpu_query_keys at api/mod.rs:243-269);
//   503 the request can never fit the loop's key slots or its wait timed out,
//   500 device error.
// One difference: keys are accepted up to d = 32 (the kernel's key record
// holds 32 correction words; the reference's MAX_DOMAIN_BITS = 25 comes from
// its 25-slot device key, kernel.rs:45).
#include <ctype.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <utility>
#include <vector>

#include "mg_internal.hpp"

namespace {

// ---------------------------------------------------------------- JSON ----
// A small DOM: the request bodies are at most a few hundred KiB.
struct JVal {
  enum Kind { kNull, kBool, kNum, kStr, kArr, kObj } kind = kNull;
  std::string str;
  std::vector<JVal> arr;
  std::vector<std::pair<std::string, JVal>> obj;
};

constexpr int kMaxDepth = 128;  // serde_json's recursion limit

struct Reader {
  const char* p;
  const char* end;
  std::string err;

  void ws() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
  }
  bool fail(const char* what) {
    if (err.empty()) err = what;
    return false;
  }
  bool lit(const char* s, JVal& v, JVal::Kind k) {
    const size_t n = strlen(s);
    if ((size_t)(end - p) < n || memcmp(p, s, n) != 0) return fail("invalid literal");
    p += n;
    v.kind = k;
    return true;
  }
  static int hexval(char c) {
    if (c >= '0' && c <= '9') return c - '0';
    if (c >= 'a' && c <= 'f') return c - 'a' + 10;
    if (c >= 'A' && c <= 'F') return c - 'A' + 10;
    return -1;
  }
  bool hex4(uint32_t& u) {
    if (end - p < 4) return fail("truncated \\u escape");
    u = 0;
    for (int i = 0; i < 4; i++) {
      const int h = hexval(p[i]);
      if (h < 0) return fail("invalid \\u escape");
      u = u << 4 | (uint32_t)h;
    }
    p += 4;
    return true;
  }
  static void put_utf8(std::string& s, uint32_t cp) {
    if (cp < 0x80) {
      s += (char)cp;
    } else if (cp < 0x800) {
      s += (char)(0xC0 | cp >> 6);
      s += (char)(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      s += (char)(0xE0 | cp >> 12);
      s += (char)(0x80 | ((cp >> 6) & 0x3F));
      s += (char)(0x80 | (cp & 0x3F));
    } else {
      s += (char)(0xF0 | cp >> 18);
      s += (char)(0x80 | ((cp >> 12) & 0x3F));
      s += (char)(0x80 | ((cp >> 6) & 0x3F));
      s += (char)(0x80 | (cp & 0x3F));
    }
  }
  // Length of a valid UTF-8 sequence starting at q, or 0.
  size_t utf8_len(const unsigned char* q, size_t avail) {
    const unsigned c = q[0];
    size_t n;
    uint32_t cp;
    if (c < 0x80) return 1;
    if ((c & 0xE0) == 0xC0) { n = 2; cp = c & 0x1F; }
    else if ((c & 0xF0) == 0xE0) { n = 3; cp = c & 0x0F; }
    else if ((c & 0xF8) == 0xF0) { n = 4; cp = c & 0x07; }
    else return 0;
    if (avail < n) return 0;
    for (size_t i = 1; i < n; i++) {
      if ((q[i] & 0xC0) != 0x80) return 0;
      cp = cp << 6 | (q[i] & 0x3F);
    }
    if ((n == 2 && cp < 0x80) || (n == 3 && cp < 0x800) || (n == 4 && cp < 0x10000) ||
        cp > 0x10FFFF || (cp >= 0xD800 && cp <= 0xDFFF))
      return 0;
    return n;
  }
  bool string(std::string& s) {
    ++p;  // opening quote
    while (true) {
      if (p >= end) return fail("EOF while parsing a string");
      const unsigned char c = (unsigned char)*p;
      if (c == '"') { ++p; return true; }
      if (c < 0x20) return fail("control character in string");
      if (c == '\\') {
        if (++p >= end) return fail("EOF in escape");
        const char e = *p++;
        switch (e) {
          case '"': s += '"'; break;
          case '\\': s += '\\'; break;
          case '/': s += '/'; break;
          case 'b': s += '\b'; break;
          case 'f': s += '\f'; break;
          case 'n': s += '\n'; break;
          case 'r': s += '\r'; break;
          case 't': s += '\t'; break;
          case 'u': {
            uint32_t u;
            if (!hex4(u)) return false;
            if (u >= 0xDC00 && u <= 0xDFFF) return fail("lone trailing surrogate");
            if (u >= 0xD800 && u <= 0xDBFF) {
              if (end - p < 2 || p[0] != '\\' || p[1] != 'u') return fail("lone leading surrogate");
              p += 2;
              uint32_t lo;
              if (!hex4(lo)) return false;
              if (lo < 0xDC00 || lo > 0xDFFF) return fail("invalid surrogate pair");
              u = 0x10000 + ((u - 0xD800) << 10) + (lo - 0xDC00);
            }
            put_utf8(s, u);
            break;
          }
          default:
            return fail("invalid escape");
        }
        continue;
      }
      const size_t n = utf8_len((const unsigned char*)p, (size_t)(end - p));
      if (!n) return fail("invalid UTF-8 in string");
      s.append(p, n);
      p += n;
    }
  }
  bool number(JVal& v) {
    const char* q = p;
    if (q < end && *q == '-') ++q;
    if (q >= end || !isdigit((unsigned char)*q)) return fail("invalid number");
    if (*q == '0') ++q;
    else while (q < end && isdigit((unsigned char)*q)) ++q;
    if (q < end && *q == '.') {
      ++q;
      if (q >= end || !isdigit((unsigned char)*q)) return fail("invalid number");
      while (q < end && isdigit((unsigned char)*q)) ++q;
    }
    if (q < end && (*q == 'e' || *q == 'E')) {
      ++q;
      if (q < end && (*q == '+' || *q == '-')) ++q;
      if (q >= end || !isdigit((unsigned char)*q)) return fail("invalid number");
      while (q < end && isdigit((unsigned char)*q)) ++q;
    }
    v.kind = JVal::kNum;
    v.str.assign(p, q);
    p = q;
    return true;
  }
  bool value(JVal& v, int depth) {
    ws();
    if (p >= end) return fail("EOF while parsing a value");
    switch (*p) {
      case 'n': return lit("null", v, JVal::kNull);
      case 't': v.str = "true"; return lit("true", v, JVal::kBool);
      case 'f': v.str = "false"; return lit("false", v, JVal::kBool);
      case '"': v.kind = JVal::kStr; return string(v.str);
      case '[': {
        if (depth >= kMaxDepth) return fail("recursion limit exceeded");
        ++p;
        v.kind = JVal::kArr;
        ws();
        if (p < end && *p == ']') { ++p; return true; }
        while (true) {
          v.arr.emplace_back();
          if (!value(v.arr.back(), depth + 1)) return false;
          ws();
          if (p >= end) return fail("EOF while parsing a list");
          if (*p == ',') { ++p; continue; }
          if (*p == ']') { ++p; return true; }
          return fail("expected , or ]");
        }
      }
      case '{': {
        if (depth >= kMaxDepth) return fail("recursion limit exceeded");
        ++p;
        v.kind = JVal::kObj;
        ws();
        if (p < end && *p == '}') { ++p; return true; }
        while (true) {
          ws();
          if (p >= end || *p != '"') return fail("key must be a string");
          std::string k;
          if (!string(k)) return false;
          ws();
          if (p >= end || *p != ':') return fail("expected :");
          ++p;
          v.obj.emplace_back(std::move(k), JVal());
          if (!value(v.obj.back().second, depth + 1)) return false;
          ws();
          if (p >= end) return fail("EOF while parsing an object");
          if (*p == ',') { ++p; continue; }
          if (*p == '}') { ++p; return true; }
          return fail("expected , or }");
        }
      }
      default:
        if (*p == '-' || isdigit((unsigned char)*p)) return number(v);
        return fail("expected value");
    }
  }
};

// Parses a whole body; syntax errors (including trailing characters) -> false.
bool parse_json(const char* body, size_t len, JVal& out, std::string& err) {
  Reader r{body, body + len, {}};
  if (!r.value(out, 0)) {
    err = r.err;
    return false;
  }
  r.ws();
  if (r.p != r.end) {
    err = "trailing characters";
    return false;
  }
  return true;
}

// -------------------------------------------------- schema (serde rules) --
// serde_hex.rs:71-84: an optional "0x" prefix, then an even number of hex
// digits of either case.
bool decode_hex(const std::string& s, std::vector<uint8_t>& out, std::string& err) {
  size_t i = (s.size() >= 2 && s[0] == '0' && s[1] == 'x') ? 2 : 0;
  if ((s.size() - i) % 2) {
    err = "odd number of hex digits";
    return false;
  }
  out.clear();
  out.reserve((s.size() - i) / 2);
  for (; i < s.size(); i += 2) {
    const int hi = Reader::hexval(s[i]), lo = Reader::hexval(s[i + 1]);
    if (hi < 0 || lo < 0) {
      err = "invalid hex character";
      return false;
    }
    out.push_back((uint8_t)(hi << 4 | lo));
  }
  return true;
}

// A struct with one field, as serde's derive reads it: a map holding the
// field once (other fields are ignored) or a one-element sequence.
const JVal* single_field(const JVal& v, const char* name, std::string& err) {
  if (v.kind == JVal::kArr) {
    if (v.arr.size() != 1) {
      err = "invalid length " + std::to_string(v.arr.size()) + ", expected struct with 1 element";
      return nullptr;
    }
    return &v.arr[0];
  }
  if (v.kind != JVal::kObj) {
    err = std::string("invalid type: expected struct with field `") + name + "`";
    return nullptr;
  }
  const JVal* found = nullptr;
  for (const auto& kv : v.obj) {
    if (kv.first != name) continue;
    if (found) {
      err = std::string("duplicate field `") + name + "`";
      return nullptr;
    }
    found = &kv.second;
  }
  if (!found) err = std::string("missing field `") + name + "`";
  return found;
}

// GpuPageQueryRequest {keys: Vec<hex>} (dto.rs:126-130).
bool query_keys(const JVal& v, std::vector<std::vector<uint8_t>>& keys, std::string& err) {
  const JVal* k = single_field(v, "keys", err);
  if (!k) return false;
  if (k->kind != JVal::kArr) {
    err = "invalid type: expected a sequence of hex strings";
    return false;
  }
  keys.clear();
  for (const JVal& e : k->arr) {
    if (e.kind != JVal::kStr) {
      err = "invalid type: expected a hex string";
      return false;
    }
    keys.emplace_back();
    if (!decode_hex(e.str, keys.back(), err)) return false;
  }
  return true;
}

bool ieq_prefix(const std::string& s, size_t at, const char* lit) {
  const size_t n = strlen(lit);
  if (s.size() < at + n) return false;
  for (size_t i = 0; i < n; i++)
    if (tolower((unsigned char)s[at + i]) != lit[i]) return false;
  return true;
}

// axum's Json extractor: `application/json`, `application/<x>+json`, with
// any parameters.
bool json_content_type(const char* ct) {
  if (!ct) return false;
  std::string s(ct);
  const size_t semi = s.find(';');
  if (semi != std::string::npos) s.resize(semi);
  while (!s.empty() && isspace((unsigned char)s.back())) s.pop_back();
  size_t b = 0;
  while (b < s.size() && isspace((unsigned char)s[b])) ++b;
  if (!ieq_prefix(s, b, "application/")) return false;
  std::string sub = s.substr(b + 12);
  for (auto& c : sub) c = (char)tolower((unsigned char)c);
  if (sub == "json") return true;
  return sub.size() > 5 && sub.compare(sub.size() - 5, 5, "+json") == 0;
}

uint32_t opt_max_batch(const mg_wire_opts* o) {
  return o && o->max_batch ? o->max_batch : MG_WIRE_MAX_BATCH;
}
uint64_t opt_max_body(const mg_wire_opts* o) {
  return o && o->max_body_bytes ? o->max_body_bytes : MG_WIRE_MAX_BODY;
}

struct Route {
  int status = 200;
  bool batch = false;
};

Route route(const char* method, const char* path) {
  Route r;
  if (!path) { r.status = 404; return r; }
  if (!strcmp(path, "/query/page/gpu")) r.batch = false;
  else if (!strcmp(path, "/query/page/gpu/batch")) r.batch = true;
  else { r.status = 404; return r; }
  if (!method || strcmp(method, "POST") != 0) r.status = 405;
  return r;
}

// Extractor and handler validation; fills flat keys on success.
int parse_request(const Route& rt, const mg_wire_opts* opts, const char* content_type,
                  const char* body, size_t body_len,
                  std::vector<std::vector<uint8_t>>& keys, uint32_t* n_queries, int* status) {
  *n_queries = 0;
  if (rt.status != 200) {
    *status = rt.status;
    return mg::fail(MG_ERR_INVALID_ARG, rt.status == 404 ? "no such route" : "method not allowed");
  }
  if (!json_content_type(content_type)) {
    *status = 415;
    return mg::fail(MG_ERR_INVALID_ARG, "expected request with Content-Type: application/json");
  }
  if (body_len > opt_max_body(opts)) {
    *status = 413;
    return mg::fail(MG_ERR_INVALID_ARG, "body of %zu bytes exceeds the %llu-byte limit", body_len,
                    (unsigned long long)opt_max_body(opts));
  }
  JVal doc;
  std::string err;
  if (!parse_json(body ? body : "", body ? body_len : 0, doc, err)) {
    *status = 400;
    return mg::fail(MG_ERR_INVALID_ARG, "failed to parse the request body as JSON: %s",
                    err.c_str());
  }
  std::vector<std::vector<std::vector<uint8_t>>> queries;
  if (!rt.batch) {
    queries.emplace_back();
    if (!query_keys(doc, queries.back(), err)) {
      *status = 422;
      return mg::fail(MG_ERR_INVALID_ARG, "failed to deserialize the JSON body: %s", err.c_str());
    }
  } else {
    const JVal* q = single_field(doc, "queries", err);
    if (q && q->kind != JVal::kArr) {
      err = "invalid type: expected a sequence of queries";
      q = nullptr;
    }
    if (q) {
      for (const JVal& e : q->arr) {
        queries.emplace_back();
        if (!query_keys(e, queries.back(), err)) {
          q = nullptr;
          break;
        }
      }
    }
    if (!q) {
      *status = 422;
      return mg::fail(MG_ERR_INVALID_ARG, "failed to deserialize the JSON body: %s", err.c_str());
    }
    if (queries.empty() || queries.size() > opt_max_batch(opts)) {
      *status = 400;
      return mg::fail(MG_ERR_INVALID_ARG, "a batch holds 1 to %u queries, not %zu",
                      opt_max_batch(opts), queries.size());
    }
  }
  keys.clear();
  for (size_t qi = 0; qi < queries.size(); qi++) {
    if (queries[qi].size() != MG_WIRE_KEYS_PER_QUERY) {
      *status = 400;
      return mg::fail(MG_ERR_INVALID_ARG, "query %zu has %zu keys, expected 3", qi,
                      queries[qi].size());
    }
    for (auto& k : queries[qi]) {
      mg::DevKey dk;
      uint32_t party;
      const mg::ParseStatus ps = mg::parse_wire_key(k.data(), k.size(), &dk, &party);
      if (ps != mg::kParseOk) {
        *status = 400;
        return mg::fail(ps == mg::kParseDomain ? MG_ERR_DOMAIN : MG_ERR_KEY_FORMAT,
                        "query %zu: key of %zu bytes does not parse", qi, k.size());
      }
      keys.push_back(std::move(k));
    }
  }
  *n_queries = (uint32_t)queries.size();
  *status = 200;
  return MG_OK;
}

const char kHex[] = "0123456789abcdef";

size_t format_response(uint64_t epoch, const uint8_t* pages, uint32_t n_queries,
                       uint32_t row_bytes, bool batch, char* out) {
  char* o = out;
  o += sprintf(o, "{\"epoch_id\":%llu,", (unsigned long long)epoch);
  auto page_list = [&](const uint8_t* p3) {
    memcpy(o, "\"pages\":[", 9);
    o += 9;
    for (int k = 0; k < MG_WIRE_KEYS_PER_QUERY; k++) {
      if (k) *o++ = ',';
      *o++ = '"';
      *o++ = '0';
      *o++ = 'x';
      const uint8_t* p = p3 + (size_t)k * row_bytes;
      for (uint32_t i = 0; i < row_bytes; i++) {
        *o++ = kHex[p[i] >> 4];
        *o++ = kHex[p[i] & 15];
      }
      *o++ = '"';
    }
    *o++ = ']';
  };
  if (!batch) {
    page_list(pages);
  } else {
    memcpy(o, "\"results\":[", 11);
    o += 11;
    for (uint32_t q = 0; q < n_queries; q++) {
      if (q) *o++ = ',';
      *o++ = '{';
      page_list(pages + (size_t)q * MG_WIRE_KEYS_PER_QUERY * row_bytes);
      *o++ = '}';
    }
    *o++ = ']';
  }
  *o++ = '}';
  *o = 0;
  return (size_t)(o - out);
}

}  // namespace

extern "C" size_t mg_wire_response_capacity(uint32_t n_queries, uint32_t row_bytes) {
  // {"epoch_id":<20 digits>,"results":[ ... ]} + NUL; per query
  // {"pages":[ three "0x<2W>" strings ]}.
  const size_t per_page = 2 * (size_t)row_bytes + 5;
  const size_t per_query = 12 + MG_WIRE_KEYS_PER_QUERY * per_page;
  return 48 + (size_t)(n_queries ? n_queries : 1) * per_query;
}

extern "C" int mg_wire_parse(const mg_wire_opts* opts, const char* method, const char* path,
                             const char* content_type, const char* body, size_t body_len,
                             uint8_t* keys, size_t keys_cap, size_t* key_lens,
                             uint32_t key_lens_cap, uint32_t* n_queries, int* http_status) {
  if (!n_queries || !http_status || !key_lens || (!keys && keys_cap))
    return mg::fail(MG_ERR_INVALID_ARG, "null argument");
  std::vector<std::vector<uint8_t>> flat;
  const Route rt = route(method, path);
  int rc = parse_request(rt, opts, content_type, body, body_len, flat, n_queries, http_status);
  if (rc) return rc;
  if (flat.size() > key_lens_cap)
    return mg::fail(MG_ERR_INVALID_ARG, "%zu keys exceed key_lens_cap %u", flat.size(),
                    key_lens_cap);
  size_t total = 0;
  for (const auto& k : flat) total += k.size();
  if (total > keys_cap)
    return mg::fail(MG_ERR_INVALID_ARG, "%zu key bytes exceed keys_cap %zu", total, keys_cap);
  size_t off = 0;
  for (size_t i = 0; i < flat.size(); i++) {
    memcpy(keys + off, flat[i].data(), flat[i].size());
    key_lens[i] = flat[i].size();
    off += flat[i].size();
  }
  return MG_OK;
}

extern "C" int mg_wire_format(uint64_t epoch, const uint8_t* pages, uint32_t n_queries,
                              uint32_t row_bytes, int batch, char* out, size_t out_cap,
                              size_t* out_len) {
  if (!pages || !out || !out_len) return mg::fail(MG_ERR_INVALID_ARG, "null argument");
  if (!n_queries || (!batch && n_queries != 1))
    return mg::fail(MG_ERR_INVALID_ARG, "a single response holds one query, a batch at least one");
  const size_t need = mg_wire_response_capacity(n_queries, row_bytes);
  if (out_cap < need) {
    *out_len = need;
    return mg::fail(MG_ERR_INVALID_ARG, "response needs %zu bytes, buffer holds %zu", need,
                    out_cap);
  }
  *out_len = format_response(epoch, pages, n_queries, row_bytes, batch != 0, out);
  return MG_OK;
}

extern "C" int mg_wire_handle(mg_server* s, const mg_wire_opts* opts, const char* method,
                              const char* path, const char* content_type, const char* body,
                              size_t body_len, char* out, size_t out_cap, size_t* out_len,
                              int* http_status) {
  if (!out || !out_len || !http_status) return mg::fail(MG_ERR_INVALID_ARG, "null argument");
  *out_len = 0;
  if (out_cap) out[0] = 0;
  std::vector<std::vector<uint8_t>> flat;
  uint32_t n_queries = 0;
  const Route rt = route(method, path);
  if (parse_request(rt, opts, content_type, body, body_len, flat, &n_queries, http_status))
    return MG_OK;  // an error response; mg_last_error() says why
  if (!s) {
    *http_status = 503;
    mg::fail(MG_ERR_NO_DEVICE, "no GPU request loop behind this API");
    return MG_OK;
  }
  const uint32_t w = mg::server_row_bytes(s);
  const size_t need = mg_wire_response_capacity(n_queries, w);
  if (out_cap < need) {
    *out_len = need;
    return mg::fail(MG_ERR_INVALID_ARG, "response needs %zu bytes, buffer holds %zu", need,
                    out_cap);
  }
  // All keys of a request (a batch included) form one loop request, so they
  // share one admission epoch, as the reference's batch shares one snapshot.
  std::vector<uint8_t> keys;
  std::vector<size_t> lens;
  for (const auto& k : flat) {
    keys.insert(keys.end(), k.begin(), k.end());
    lens.push_back(k.size());
  }
  uint64_t ticket = 0;
  int rc = mg_server_submit(s, keys.data(), lens.data(), (uint32_t)lens.size(), &ticket);
  if (rc) {
    *http_status = (rc == MG_ERR_DOMAIN || rc == MG_ERR_KEY_FORMAT) ? 400
                   : rc == MG_ERR_BUSY                               ? 503
                                                                     : 500;
    return MG_OK;
  }
  std::vector<uint8_t> pages(lens.size() * (size_t)w);
  mg_request_info info;
  rc = mg_server_wait(s, ticket, pages.data(), opts ? opts->timeout_ms : 0, &info);
  if (rc) {
    if (rc == MG_ERR_BUSY) (void)mg_server_abandon(s, ticket);  // timed out: nobody collects it
    *http_status = rc == MG_ERR_BUSY ? 503 : 500;
    return MG_OK;
  }
  *out_len = format_response(info.epoch, pages.data(), n_queries, w, rt.batch, out);
  *http_status = 200;
  return MG_OK;
}
