// ChaCha DPF key: wire parsing (host) and the device record the kernels read.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace mg {

constexpr int kMaxDomainBits = 32;

// Device key: one 16-byte-aligned record per DPF key, 688 bytes. Correction
// word control bytes are packed (left in bits 0-7, right in bits 8-15) so one
// scalar load serves a level. Not the reference's DpfKeyGpu layout
// (kernel.rs:109-154, 25 levels): this one holds 32 levels.
struct alignas(16) DevKey {
  uint32_t root_seed[4];
  uint32_t root_t;       // control byte, 0..255
  uint32_t domain_bits;  // 0..32
  uint32_t pad[2];
  uint32_t final_cw[4];
  uint32_t cw_seed[kMaxDomainBits][4];
  uint32_t cw_t[kMaxDomainBits];  // t_cw_left | t_cw_right << 8
};
static_assert(sizeof(DevKey) == 688, "DevKey layout");

enum ParseStatus { kParseOk = 0, kParseFormat = 1, kParseDomain = 2 };

// ChaChaKey::from_bytes (dpf.rs:277-316): root_seed[16] | root_t | party |
// d | d x (seed_cw[16] | t_cw_left | t_cw_right) | final_cw[16], exact
// length 19 + 18d + 16; seeds are 4 little-endian u32 words
// (chacha_prg.rs:27-36). d > 32 is refused (ChaChaParams::new, dpf.rs:65-78).
inline ParseStatus parse_wire_key(const uint8_t* b, size_t len, DevKey* out,
                                  uint32_t* party) {
  if (len < 19 + 16) return kParseFormat;
  const uint32_t d = b[18];
  if (len != 19 + 18 * (size_t)d + 16) return kParseFormat;
  if (d > (uint32_t)kMaxDomainBits) return kParseDomain;
  auto le = [](const uint8_t* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
           ((uint32_t)p[3] << 24);
  };
  DevKey k{};
  for (int i = 0; i < 4; i++) k.root_seed[i] = le(b + 4 * i);
  k.root_t = b[16];
  *party = b[17];
  k.domain_bits = d;
  size_t off = 19;
  for (uint32_t l = 0; l < d; l++) {
    for (int i = 0; i < 4; i++) k.cw_seed[l][i] = le(b + off + 4 * i);
    k.cw_t[l] = (uint32_t)b[off + 16] | ((uint32_t)b[off + 17] << 8);
    off += 18;
  }
  for (int i = 0; i < 4; i++) k.final_cw[i] = le(b + off + 4 * i);
  *out = k;
  return kParseOk;
}

}  // namespace mg
