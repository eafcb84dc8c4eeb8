"""Writes chacha20_openssl.json: ChaCha20 keystream blocks from OpenSSL.

OpenSSL's `chacha20` cipher is an implementation of RFC 8439 independent of
this project. Its 16-byte IV is the 32-bit block counter (little-endian)
followed by the 96-bit nonce. Encrypting zeros yields the raw keystream,
i.e. serialised ChaCha20 blocks. The oracle's block function is checked
against these at 20 rounds; the 8-round PRG runs the same code.

Cases with key = seed || seed pin the PRG expand wiring (chacha_prg.rs:153-178):
block(counter 0) bytes 0..31 are the two child seeds, block(counter 1) bytes
0 and 4 carry the control bits.

Usage: python tests/golden/gen_chacha_openssl.py   (needs the openssl CLI)
"""
import json
import os
import random
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def keystream(key: bytes, counter: int, nonce: bytes, n_bytes: int) -> bytes:
    iv = counter.to_bytes(4, "little") + nonce
    out = subprocess.run(
        ["openssl", "enc", "-chacha20", "-K", key.hex(), "-iv", iv.hex()],
        input=bytes(n_bytes), capture_output=True, check=True).stdout
    assert len(out) == n_bytes
    return out


def main():
    rng = random.Random(20261203)
    cases = []
    keys = [bytes(32), bytes(range(32)), bytes([0xFF] * 32)]
    keys += [bytes(rng.getrandbits(8) for _ in range(32)) for _ in range(3)]
    for key in keys:
        for counter, nonce in [(0, bytes(12)), (1, bytes(12)),
                               (7, bytes(range(12))),
                               (0x12345678, bytes([0, 0, 0, 9, 0, 0, 0, 0x4a, 0, 0, 0, 0]))]:
            ks = keystream(key, counter, nonce, 128)
            cases.append({"key": key.hex(), "counter": counter,
                          "nonce": nonce.hex(), "keystream": ks.hex()})
    seeds = [bytes(16), bytes(range(16)),
             bytes.fromhex("78563412f0debc9a98badcfe10325476")]
    seeds += [bytes(rng.getrandbits(8) for _ in range(16)) for _ in range(5)]
    expand = []
    for seed in seeds:
        ks = keystream(seed + seed, 0, bytes(12), 128)
        expand.append({"seed": seed.hex(), "keystream": ks.hex()})
    ver = subprocess.run(["openssl", "version"], capture_output=True,
                         text=True, check=True).stdout.strip()
    with open(os.path.join(HERE, "chacha20_openssl.json"), "w") as f:
        json.dump({"source": ver, "block_cases": cases,
                   "expand_cases": expand}, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
