"""PRG and VALU microbenchmarks: the VALU roofline of the scan kernels.

`prg` times chains of ChaCha8 GGM node expansions in registers (the scan
kernels' node routine, correction words included); `valu` times independent
chains of single-instruction int32 adds and xors, whose rate over 256 CUs x 64
lanes is the issue clock. Both run through libmg_microbench.so
(csrc/mg_microbench.hip); importing this module without it raises ImportError.

CLI: python -m morphogenesis_amd.microbench   (one JSON line)
"""
import argparse
import ctypes
import json
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmg_microbench.so")
if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built (make -C morphogenesis_amd)")
_lib = ctypes.CDLL(LIB_PATH)
for _name in ("mgb_prg", "mgb_valu"):
    _f = getattr(_lib, _name)
    _f.restype = ctypes.c_int
    _f.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint32)]
_lib.mgb_prg2.restype = ctypes.c_int
_lib.mgb_prg2.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint32)]
_lib.mgb_op.restype = ctypes.c_int
_lib.mgb_op.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint32)]

_lib.mgb_mix.restype = ctypes.c_int
_lib.mgb_mix.argtypes = _lib.mgb_op.argtypes
# Dependency patterns of mgb_mix: (chains interleaved, instruction mix).
MIX_FORMS = ["1 chain add", "2 chains add", "4 chains add", "8 chains add", "16 chains add",
             "4 chains add/xor/alignbit", "8 chains add/xor/alignbit",
             "16 chains add/xor/alignbit", "4 chains add/xor", "8 chains add/xor",
             "8 chains add, operand in the same VGPR bank",
             "8 chains add, operand in another VGPR bank",
             "8 chains add/xor/alignbit, 720-instruction loop",
             "8 chains add/xor, 720-instruction loop",
             "8 chains x += z, z ^= x (both operands fresh)",
             "8 chains x += z, z ^= x, z = rotl(z) (both operands fresh)",
             "8 quarter-rounds, op-major order", "8 quarter-rounds, one QR at a time",
             "16 quarter-rounds, op-major order",
             "8 quarter-rounds, two groups skewed by one stage",
             "8 quarter-rounds, op-major, rotate 16 as v_pk_add_u16 op_sel",
             "8 quarter-rounds, op-major, xor + rotate 16 as two SDWA xors"]

# Instruction forms of mgb_op (csrc/mg_microbench.hip, op_once).
OP_FORMS = ["v_add_u32 (VOP2)", "v_xor_b32 (VOP2)", "v_alignbit_b32 x,x,x,25",
            "v_perm_b32 x,x,x,sel", "v_lshlrev_b32 7 (VOP2)", "v_bitop3_b32 0x78",
            "v_add3_u32", "v_xad_u32", "v_lshl_or_b32", "v_alignbyte_b32 x,x,x,2",
            "v_add_u32_e64 (VOP3)", "v_lshrrev_b32 25 (VOP2)", "v_or_b32 (VOP2)",
            "v_alignbit_b32 x,y,25", "v_pk_add_u16 op_sel swap (rotate 16)",
            "v_pk_add_u16", "v_xor_b32_sdwa WORD_1 -> WORD_0, pad",
            "v_xor_b32_sdwa WORD_0 -> WORD_1, preserve"]

# Node routines of mgb_prg2 (csrc/mg_prg_variants.hpp, bench_expand_node).
PRG_VARIANTS = ["compiler_order", "ordered", "round_statements_unspaced",
                "rounds_2_to_8_spaced_add2_xor2_rot6", "rounds_2_to_8_spaced_add2_xor2_rot10",
                "rounds_2_to_8_spaced_add2_xor2_rot14"]
# The kernels' lane-DFS / eval routine (kPrgSpaced) and DFS-wave routine (kPrgOrdered).
KERNEL_VARIANT = "rounds_2_to_8_spaced_add2_xor2_rot10"
DFS_WAVE_VARIANT = "ordered"

# int32 VALU ops per GGM node expansion, the constant every roofline here uses:
# two ChaCha8 blocks after dead-code elimination plus the correction word.
OPS_PER_NODE = 700
VALU_OPS_PER_ITER = 16  # valu_bench: 8 v_add_u32 + 8 v_xor_b32 per lane


def _run(fn, device, blocks, iters):
    ms = ctypes.c_double()
    chk = ctypes.c_uint32()
    rc = fn(device, blocks, iters, ctypes.byref(ms), ctypes.byref(chk))
    if rc != 0:
        raise RuntimeError(f"HIP error {rc}")
    return ms.value, chk.value


def prg(blocks, iters, device=0):
    """(milliseconds, checksum) of blocks x 256 lanes x iters node expansions."""
    return _run(_lib.mgb_prg, device, blocks, iters)


def prg_variant(ordered, blocks, iters, device=0):
    """prg() with expansion `ordered`, an index into PRG_VARIANTS."""
    ms = ctypes.c_double()
    chk = ctypes.c_uint32()
    rc = _lib.mgb_prg2(device, ordered, blocks, iters, ctypes.byref(ms), ctypes.byref(chk))
    if rc != 0:
        raise RuntimeError(f"HIP error {rc}")
    return ms.value, chk.value


def valu(blocks, iters, device=0):
    """(milliseconds, checksum) of blocks x 256 lanes x iters x 16 int32 ops."""
    return _run(_lib.mgb_valu, device, blocks, iters)


def op(form, blocks, iters, device=0):
    """(milliseconds, checksum) of blocks x 256 lanes x iters x 16 instructions
    of form `form` (an index into OP_FORMS)."""
    ms = ctypes.c_double()
    chk = ctypes.c_uint32()
    rc = _lib.mgb_op(device, form, blocks, iters, ctypes.byref(ms), ctypes.byref(chk))
    if rc != 0:
        raise RuntimeError(f"HIP error {rc}")
    return ms.value, chk.value


def mix(form, blocks, iters, device=0):
    """(milliseconds, checksum) of blocks x 256 lanes x iters x 48 instructions
    in dependency pattern `form` (an index into MIX_FORMS)."""
    ms = ctypes.c_double()
    chk = ctypes.c_uint32()
    rc = _lib.mgb_mix(device, form, blocks, iters, ctypes.byref(ms), ctypes.byref(chk))
    if rc != 0:
        raise RuntimeError(f"HIP error {rc}")
    return ms.value, chk.value


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--blocks", type=int, default=0, help="0 = 8 blocks of 256 threads per CU")
    ap.add_argument("--prg-iters", type=int, default=4096)
    ap.add_argument("--valu-iters", type=int, default=200_000)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--forms", action="store_true",
                    help="also time each instruction form of OP_FORMS")
    args = ap.parse_args()
    from . import GpuScanner
    s = GpuScanner(args.device)
    cus = s.info()["compute_units"]
    s.close()
    blocks = args.blocks or 8 * cus
    lanes = blocks * 256
    prg_ms = min(prg(blocks, args.prg_iters, args.device)[0] for _ in range(args.reps + 1))
    variants = {}
    for v in range(len(PRG_VARIANTS)):
        v_ms = min(prg_variant(v, blocks, args.prg_iters, args.device)[0]
                   for _ in range(args.reps + 1))
        variants[PRG_VARIANTS[v]] = {
            "ms": v_ms, "tops_at_700_ops_per_node":
                lanes * args.prg_iters / (v_ms * 1e-3) * OPS_PER_NODE / 1e12}
    valu_ms = min(valu(blocks, args.valu_iters, args.device)[0] for _ in range(args.reps + 1))
    nodes_s = lanes * args.prg_iters / (prg_ms * 1e-3)
    valu_ops_s = lanes * args.valu_iters * VALU_OPS_PER_ITER / (valu_ms * 1e-3)
    peak = cus * 64 * 2.4e9
    forms = {}
    if args.forms:
        for k, name in enumerate(OP_FORMS):
            ms_k = min(op(k, blocks, args.valu_iters, args.device)[0] for _ in range(args.reps))
            per_simd_ns = lanes / 64 * args.valu_iters * 16 / (cus * 4) / (ms_k * 1e6)
            forms[name] = {"ms": ms_k, "wave_instr_per_simd_per_ns": per_simd_ns,
                           "cycles_per_instr_at_2p4ghz": 2.4 / per_simd_ns}
        it = args.valu_iters // 45 * 15
        for k, name in enumerate(MIX_FORMS):
            ms_k = min(mix(k, blocks, it, args.device)[0] for _ in range(args.reps))
            per_simd_ns = lanes / 64 * it * 48 / (cus * 4) / (ms_k * 1e6)
            forms[name] = {"ms": ms_k, "wave_instr_per_simd_per_ns": per_simd_ns,
                           "cycles_per_instr_at_2p4ghz": 2.4 / per_simd_ns}
    print(json.dumps({
        "compute_units": cus, "blocks": blocks, "threads_per_block": 256,
        "prg": {"iters": args.prg_iters, "ms": prg_ms, "nodes_per_s": nodes_s,
                "nodes_per_s_per_cu": nodes_s / cus,
                "tops_at_700_ops_per_node": nodes_s * OPS_PER_NODE / 1e12,
                "frac_of_2p4ghz_peak": nodes_s * OPS_PER_NODE / peak},
        "valu": {"iters": args.valu_iters, "ms": valu_ms, "tops": valu_ops_s / 1e12,
                 "frac_of_2p4ghz_peak": valu_ops_s / peak,
                 "issue_clock_ghz": valu_ops_s / (cus * 64) / 1e9},
        "peak_tops_2p4ghz": peak / 1e12,
        "prg_variants": variants,
        "forms": forms or None,
    }))


if __name__ == "__main__":
    main()
