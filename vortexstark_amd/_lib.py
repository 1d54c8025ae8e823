"""ctypes binding of libvortex_hip.so (the C-ABI declared in include/vortex_hip.h).

The library is built in-tree by vortexstark_amd/csrc/Makefile (see
__graft_entry__.build()). Loading fails loudly when it is missing: there is no
CPU fallback for any entry point.
"""
import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvortex_hip.so")
HEADER = os.path.join(HERE, "..", "include", "vortex_hip.h")

u32, u64, sz, ip = C.c_uint32, C.c_uint64, C.c_size_t, C.c_int
vp = C.c_void_p
u32p = C.POINTER(C.c_uint32)
u8p = C.POINTER(C.c_uint8)

# name -> (restype, argtypes). Device pointers and streams are passed as ints.
SIGNATURES = {
    "vx_last_error": (C.c_char_p, []),
    "vx_device_sync": (ip, []),
    "vx_malloc": (ip, [C.POINTER(vp), sz]),
    "vx_free": (ip, [vp]),
    "vx_memcpy": (ip, [vp, vp, sz, ip]),
    "vx_memcpy_async": (ip, [vp, vp, sz, ip, vp]),
    "vx_memset": (ip, [vp, ip, sz]),
    "vx_stream_create": (ip, [C.POINTER(vp)]),
    "vx_stream_synchronize": (ip, [vp]),
    "vx_stream_destroy": (ip, [vp]),
    "vx_circle_ntt_evaluate": (ip, [vp, vp, vp, vp, vp, u32, u32, vp]),
    "vx_circle_ntt_interpolate": (ip, [vp, vp, vp, vp, vp, u32, u32, vp]),
    "vx_circle_ntt_evaluate_batch": (ip, [vp, vp, vp, vp, vp, u32, u32, u32, vp]),
    "vx_circle_ntt_interpolate_batch": (ip, [vp, vp, vp, vp, vp, u32, u32, u32, vp]),
    "vx_circle_ntt_evaluate_batch_host": (ip, [vp, vp, vp, vp, vp, u32, u32, u32, vp]),
    "vx_circle_ntt_interpolate_batch_host": (ip, [vp, vp, vp, vp, vp, u32, u32, u32, vp]),
    "vx_circle_ntt_layer": (ip, [vp, vp, u32, u32, ip, vp]),
    "vx_circle_ntt_evaluate_lde": (ip, [vp, vp, u32, vp, vp, vp, u32, u32, vp]),
    "vx_bit_reverse_m31": (ip, [vp, u32, vp]),
    "vx_bit_reverse_qm31": (ip, [vp, vp, u32, u32, vp]),
    "vx_compute_coset_points": (ip, [u32, u32, u32, u32, vp, vp, u32, vp]),
    "vx_extract_and_squash": (ip, [vp, vp, vp, u32, vp]),
    "vx_squash_x": (ip, [vp, vp, u32, vp]),
    "vx_batch_inverse_m31": (ip, [vp, vp, u32, vp]),
    "vx_compute_fold_twiddle_sources": (ip, [u32, u32, u32, u32, vp, u32, u32, ip, vp]),
    "vx_fibonacci_trace": (ip, [u32, u32, vp, u32, vp]),
    "vx_fibonacci_quotient": (ip, [vp, vp, vp, vp, vp, vp, u32, vp]),
    "vx_fibonacci_quotient_chunk": (ip, [vp, vp, vp, vp, vp, vp, u32, u32, u32, vp]),
    "vx_zero_pad": (ip, [vp, vp, u32, u32, vp]),
    "vx_fold_line_soa": (ip, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, u32, vp]),
    "vx_fold_circle_into_line_soa": (ip, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, u32, vp]),
    "vx_merkle_hash_leaves": (ip, [vp, vp, u32, u32, vp]),
    "vx_merkle_hash_leaves_host": (ip, [vp, vp, u32, u32, vp]),
    "vx_merkle_hash_leaves_single": (ip, [vp, vp, u32, vp]),
    "vx_merkle_hash_leaves_quad": (ip, [vp, vp, vp, vp, vp, u32, vp]),
    "vx_merkle_hash_nodes": (ip, [vp, vp, u32, vp]),
    "vx_merkle_reduce_to_root": (ip, [vp, vp, u32, vp]),
    "vx_merkle_commit_small_soa4": (ip, [vp, vp, vp, vp, vp, u32, vp]),
    "vx_merkle_hash_leaves_merge_soa4": (ip, [vp, vp, vp, vp, vp, u32, vp]),
    "vx_merkle_tiled_soa4": (ip, [vp, vp, vp, vp, vp, u32, vp]),
    "vx_merkle_tiled_generic": (ip, [vp, vp, u32, u32, vp]),
    "vx_merkle_tiled_generic_host": (ip, [vp, vp, u32, u32, vp]),
    "vx_grind_pow": (ip, [vp, vp, u32, u64, u32, vp]),
    "vx_grind_pow_min": (ip, [vp, u32, vp, vp]),
    "vx_proof_size": (sz, [u32, u32, u32, u32]),
    "vx_prover_create": (ip, [u32, u32, u32, u32, vp, C.POINTER(vp)]),
    "vx_prover_destroy": (ip, [vp]),
    "vx_prover_prove": (ip, [vp, u32, u32, vp, sz, C.POINTER(sz)]),
    "vx_prover_enqueue": (ip, [vp, u32, u32, ip]),
    "vx_prover_wait": (ip, [vp, vp, sz, C.POINTER(sz)]),
    "vx_prover_stage_times": (ip, [vp, C.POINTER(C.c_float)]),
    "vx_prover_device_bytes": (u64, [vp]),
    "vx_prover_buffer": (ip, [vp, ip, C.POINTER(vp), C.POINTER(u64)]),
    "vx_prove": (ip, [u32, u32, u32, u32, u32, u32, vp, sz, C.POINTER(sz)]),
    "vx_verify": (ip, [C.c_char_p, sz, u32, u32, C.c_char_p, sz]),
    "vx_bytecode_constraint_eval": (ip, [vp, u32, vp, vp, u32, u32, u32, vp, vp, u32, vp, vp, vp, vp, u32, vp]),
    "vx_warp_bytecode_constraint_eval": (ip, [vp, u32, vp, vp, u32, u32, u32, vp, vp, u32, vp, vp, vp, vp, u32,
                                              vp]),
    "vx_bytecode_slots": (ip, [vp, u32, u32, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]),
    "vx_bytecode_set_jit": (ip, [ip]),
    "vx_bytecode_jit_source": (ip, [vp, u32, u32, u32, u32, vp, sz, C.POINTER(sz)]),
    "vx_bytecode_jit_compile": (ip, [vp, u32, u32, u32, u32, C.POINTER(sz)]),
    "vx_export_stwo_json": (ip, [C.c_char_p, sz, u32, vp, sz, C.POINTER(sz)]),
    "vx_stwo_hash_witness": (ip, [vp, sz, vp, u32, vp, sz, C.POINTER(sz)]),
    "vx_stwo_fri_witness": (ip, [vp, sz, vp, vp, sz, C.POINTER(sz)]),
    "vx_stwo_reconstruct_root": (ip, [u32, vp, sz, vp, u32, vp, sz, vp, C.c_char_p, sz]),
    "vx_stwo_last_layer_coeffs": (ip, [vp, u32, vp]),
    "vx_prover_set_graphs": (ip, [vp, ip]),
    "vx_tuner_configure": (ip, [ip, C.c_char_p]),
    "vx_merkle_levels_words": (u64, [u32, u32]),
    "vx_merkle_commit_tiled": (ip, [vp, u32, u32, u32, vp, vp, vp]),
    "vx_quotient_commit": (ip, [vp, vp, u32, u32, vp, vp, vp]),
    "vx_circle_fold_commit": (ip, [vp, vp, vp, vp, vp, vp, vp, vp, u32, u32, vp, vp, vp]),
    "vx_fold_line_commit": (ip, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, u32, u32, vp, vp, vp]),
    "vx_tuner_entries": (ip, [C.POINTER(C.c_uint32), C.c_uint32, C.POINTER(C.c_uint32)]),
    "vx_tuner_ntt_entries": (ip, [C.POINTER(C.c_uint32), C.c_uint32, C.POINTER(C.c_uint32)]),
    "vx_d2_prover_create": (ip, [u32, u32, u32, u64, u32, u32, u32, vp, C.POINTER(vp)]),
    "vx_d2_proof_size": (sz, [u32, u32, u32, u32]),
    "vx_m31_fill_random": (ip, [vp, u32, u64, u32, vp]),
    "vx_batch_prover_create": (ip, [vp, u32, u32, u32, u32, vp, C.POINTER(vp)]),
    "vx_batch_proof_size": (sz, [vp, u32, u32, u32]),
    "vx_circle_ntt_lde_batch": (ip, [vp, vp, vp, vp, u32, u32, vp]),
    "vx_stwo_twiddles": (ip, [u32, vp, vp, vp]),
    "vx_eval_at_point": (ip, [vp, vp, vp, u32, vp, vp, vp]),
    "vx_eval_at_point_batch": (ip, [vp, vp, u32, vp, u32, vp, vp]),
    "vx_barycentric_eval": (ip, [vp, vp, u32, vp, u32, vp]),
    "vx_grind_pow_stwo": (ip, [vp, u32, vp, vp]),
    "vx_gkr_fix_first_variable_base": (ip, [vp, vp, vp, u32, vp]),
    "vx_gkr_fix_first_variable_secure": (ip, [vp, vp, vp, u32, vp]),
    "vx_gkr_gen_eq_evals_init": (ip, [vp, vp, vp]),
    "vx_gkr_gen_eq_evals_pass": (ip, [vp, vp, u32, vp]),
    "vx_gkr_gen_eq_evals": (ip, [vp, vp, u32, vp, vp]),
    "vx_gkr_next_layer_grand_product": (ip, [vp, vp, u32, vp]),
    "vx_gkr_next_layer_logup_generic": (ip, [vp, vp, vp, vp, u32, vp]),
    "vx_gkr_next_layer_logup_mult": (ip, [vp, vp, vp, vp, u32, vp]),
    "vx_gkr_next_layer_logup_singles": (ip, [vp, vp, vp, u32, vp]),
    "vx_gkr_sum_poly_grand_product": (ip, [vp, vp, vp, u32, vp, vp]),
    "vx_gkr_sum_poly_logup_generic": (ip, [vp, vp, vp, vp, vp, u32, vp, vp]),
    "vx_gkr_sum_poly_logup_mult": (ip, [vp, vp, vp, vp, vp, u32, vp, vp]),
    "vx_gkr_sum_poly_logup_singles": (ip, [vp, vp, vp, vp, u32, vp, vp]),
    "vx_grind_pow_stwo_m31": (ip, [vp, u32, vp, vp]),
    "vx_grind_pow_poseidon": (ip, [vp, vp, u32, u64, u32, vp]),
    "vx_grind_pow_stwo_poseidon": (ip, [vp, u32, vp, vp]),
    "vx_build_leaves_poseidon252": (ip, [vp, vp, u32, u32, vp, u32, vp]),
    "vx_build_next_layer_poseidon252": (ip, [vp, vp, u32, vp]),
    "vx_permute_hc_to_canonic_brt": (ip, [vp, vp, u32, u32, vp]),
    "vx_oods_sample": (ip, [vp, vp, vp, u32, u32, vp, vp, vp]),
    "vx_permute_canonic_brt_to_hc_natural": (ip, [vp, vp, u32, u32, vp]),
    "vx_grind_pow_m31_output": (ip, [vp, vp, u32, u64, u32, vp]),
    "vx_m31_add": (ip, [vp, vp, vp, u32, vp]),
    "vx_m31_mul": (ip, [vp, vp, vp, u32, vp]),
    "vx_accumulate_lift": (ip, [vp, vp, u32, u32, vp]),
    "vx_lift_and_accumulate": (ip, [vp, vp, u32, vp]),
    "vx_gather_u32": (ip, [vp, vp, vp, u32, vp]),
    "vx_gather_u32_checked": (ip, [vp, vp, vp, u32, u32, vp]),
    "vx_gather_u256": (ip, [vp, vp, vp, u32, vp]),
    "vx_gather_u256_checked": (ip, [vp, vp, vp, u32, u32, vp]),
    "vx_pack_leaves": (ip, [vp, vp, u32, vp]),
    "vx_accumulate_numerators": (ip, [vp, vp, vp, vp, u32, u32, vp, vp, vp, vp, vp]),
    "vx_accumulate_numerators_dual": (ip, [vp, vp, vp, vp, vp, vp, vp, u32, u32] + [vp] * 8 + [vp]),
    "vx_compute_quotients_combine": (ip, [vp] * 8 + [u32, vp, vp, u32, u32, vp, vp, vp, vp, vp]),
    "vx_stwo_ntt_evaluate": (ip, [vp, vp, u32, vp]),
    "vx_stwo_ntt_interpolate": (ip, [vp, vp, u32, vp]),
    "vx_stwo_ntt_batch_evaluate": (ip, [vp, vp, u32, u32, vp]),
    "vx_stwo_ntt_batch_interpolate": (ip, [vp, vp, u32, u32, vp]),
    "vx_stwo_ntt_evaluate_async": (ip, [vp, vp, u32, u32, u64, vp]),
    "vx_stwo_ntt_evaluate_lde": (ip, [vp, vp, u32, vp, u32, vp]),
    "vx_stwo_ntt_evaluate_lde_batch": (ip, [vp, vp, u32, u32, vp, u32, vp]),
    "vx_stwo_ntt_interpolate_async": (ip, [vp, vp, u32, u32, u64, vp]),
    "vx_bit_reverse_m31_batch": (ip, [vp, vp, vp, u32, vp]),
    "vx_build_leaves_lifted": (ip, [vp, vp, u32, u32, vp, vp]),
    "vx_merkle_root_lifted": (ip, [vp, vp, u32, vp, vp]),
    "vx_merkle_root_lifted_noalloc": (ip, [vp, vp, u32, vp, vp]),
    "vx_cairo_quotient": (ip, [vp] + [vp] * 32 + [vp] * 4 + [vp] * 4 + [u32, u32, vp]),
    "vx_cairo_quotient_slab": (ip, [vp] + [vp] * 32 + [vp] * 4 + [vp] * 4 + [u32, u32, u32, vp]),
    "vx_cairo_quotient_chunk": (ip, [vp] + [vp] * 12 + [vp] * 4 + [vp] * 4 + [u32, u32, u32, u32, vp]),
    "vx_cairo_quotient_commit": (ip, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, u32, u32, u32, vp, vp, vp]),
    "vx_compute_vanishing_inv": (ip, [u32, u32, u32, u32, vp, u32, u32, vp]),
    "vx_compute_vanishing": (ip, [u32, u32, u32, u32, vp, u32, u32, vp]),
    "vx_axpy_m31": (ip, [u32, vp, vp, u32, vp]),
    "vx_batch_create": (ip, [vp, u32, u32, ip, vp, C.POINTER(vp)]),
    "vx_batch_destroy": (ip, [vp]),
    "vx_batch_column": (ip, [vp, u32, ip, C.POINTER(vp), C.POINTER(u64)]),
    "vx_batch_enqueue": (ip, [vp, ip]),
    "vx_batch_wait": (ip, [vp, vp]),
    "vx_batch_stage_times": (ip, [vp, C.POINTER(C.c_float)]),
    "vx_batch_device_bytes": (u64, [vp]),
    "vx_server_create": (ip, [u32, u64, ip, C.POINTER(vp)]),
    "vx_server_destroy": (ip, [vp]),
    "vx_server_prepare": (ip, [vp, vp]),
    "vx_server_submit": (ip, [vp, vp, vp, sz, C.POINTER(u64)]),
    "vx_server_wait": (ip, [vp, u64, C.POINTER(sz), C.POINTER(C.c_double)]),
    "vx_server_done": (ip, [vp, u64]),
    "vx_server_resident_bytes": (u64, [vp]),
    "vx_server_stats": (ip, [vp, C.POINTER(u64)]),
    "vx_server_run_trace": (ip, [vp, vp, C.POINTER(C.c_double), u32, C.POINTER(vp), C.POINTER(sz),
                                 C.POINTER(sz), C.POINTER(C.c_double)]),
}


class Request(C.Structure):
    """vx_request (include/vortex_hip.h)."""
    _fields_ = [("air", u32), ("log_n", u32), ("blowup_bits", u32), ("n_queries", u32), ("pow_bits", u32),
                ("a", u32), ("b", u32), ("n_cols", u32), ("n_cons", u32), ("reserved", u32), ("seed", u64),
                ("air_seed", u64)]

_lib = None


class VortexError(RuntimeError):
    """Non-zero status from a vx_* entry point."""

    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed with status {code}: {msg}")
        self.code = code


def header_symbols():
    """Function names declared in include/vortex_hip.h."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vx_[a-z0-9_]+)\s*\(", text)))


def load():
    """Load libvortex_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(the HIP path has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(fn, rc):
    if rc != 0:
        raise VortexError(fn, rc, load().vx_last_error().decode(errors="replace"))


def call(fn, *args):
    """Call a status-returning entry point and raise on failure."""
    rc = getattr(load(), fn)(*args)
    check(fn, rc)
    return rc
