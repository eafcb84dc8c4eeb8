"""The PRG microbenchmark computes the real node chain: its checksum equals the
oracle's for a small launch, so the timed loop is the PRG and not something
the compiler thinned out."""
import pytest

import oracle

pytestmark = pytest.mark.gpu

M32 = 0xFFFFFFFF


def prg_checksum_reference(blocks, iters):
    """The kernel's checksum from the oracle's PRG expand. Lane gid starts at
    seed (gid * 0x9E3779B9, gid ^ 0x85EBCA6B, gid + 0xC2B2AE35, ~gid), t =
    gid & 1, and follows left children; the key's correction words are
    cw_seed[l][w] = 0x01234567 * (4l + w + 1), both control bits set, applied
    at level i & 15 when t == 1. The checksum XORs every right child and each
    lane's last node."""
    total = 0
    for gid in range(blocks * 256):
        s = [(gid * 0x9E3779B9) & M32, gid ^ 0x85EBCA6B, (gid + 0xC2B2AE35) & M32, ~gid & M32]
        t = gid & 1
        acc = 0
        for i in range(iters):
            lvl = i & 15
            left, right, lt, rt = oracle.prg_expand(s)
            left = [int(x) for x in left]
            right = [int(x) for x in right]
            if t == 1:
                cw = [(0x01234567 * (lvl * 4 + w + 1)) & M32 for w in range(4)]
                left = [a ^ c for a, c in zip(left, cw)]
                right = [a ^ c for a, c in zip(right, cw)]
                lt ^= 1
                rt ^= 1
            acc ^= right[0] ^ right[1] ^ right[2] ^ right[3] ^ rt
            s, t = left, lt
        total ^= acc ^ s[0] ^ s[1] ^ s[2] ^ s[3] ^ t
    return total


def test_prg_microbench_matches_oracle():
    from morphogenesis_amd import microbench
    for iters in (1, 5, 17):
        want = prg_checksum_reference(1, iters)
        ms, chk = microbench.prg(1, iters)
        assert ms > 0
        assert chk == want, iters
        # Every expansion variant.
        for ordered in range(len(microbench.PRG_VARIANTS)):
            assert microbench.prg_variant(ordered, 1, iters)[1] == want, (iters, ordered)


def test_valu_microbench_matches_its_definition():
    from morphogenesis_amd import microbench
    # x -> (x + y) ^ y per chain and iteration, y = gid * 0x9E3779B9.
    iters = 3
    ms, chk = microbench.valu(1, iters)
    want = 0
    for gid in range(256):
        y = (gid * 0x9E3779B9) & M32
        acc = 0
        for j in range(8):
            x = gid + j
            for _ in range(iters):
                x = ((x + y) & M32) ^ y
            acc ^= x
        want ^= acc
    assert chk == want


def test_rotate16_forms_compute_the_quarter_rounds():
    """The packed-add and SDWA rotate-by-16 quarter-round forms compute the same
    words as the alignbit form."""
    from morphogenesis_amd import microbench
    names = microbench.MIX_FORMS
    base = names.index("8 quarter-rounds, op-major order")
    for name in ("8 quarter-rounds, op-major, rotate 16 as v_pk_add_u16 op_sel",
                 "8 quarter-rounds, op-major, xor + rotate 16 as two SDWA xors"):
        for iters in (2, 6):
            assert (microbench.mix(names.index(name), 1, iters)[1]
                    == microbench.mix(base, 1, iters)[1]), (name, iters)
