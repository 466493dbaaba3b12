"""ctypes loader for the CPU oracle (oracle/librefcpu.so) -- the checker.

Only tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
The library is built by __graft_entry__.build(); this loader builds it only
when no built copy exists.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "librefcpu.so")
P = 2**64 - 2**32 + 1

u64p = ctypes.POINTER(ctypes.c_uint64)
u8p = ctypes.POINTER(ctypes.c_uint8)
sz = ctypes.c_size_t


class FriParams(ctypes.Structure):
    _fields_ = [("schedule", ctypes.POINTER(ctypes.c_size_t)), ("n_layers", sz), ("r", sz), ("s0", sz),
                ("seed_z", ctypes.c_uint64), ("coeff_commit_final", ctypes.c_int), ("d_final", sz),
                ("stir", ctypes.c_int), ("public_inputs_hash", ctypes.c_char_p)]


def fri_params(schedule, r=54, s0=54, seed_z=0xDEEFBAAD, coeff_commit_final=True, d_final=1, stir=True, pi_hash=None):
    arr = (ctypes.c_size_t * len(schedule))(*schedule)
    p = FriParams(arr, len(schedule), r, s0, seed_z, int(coeff_commit_final), d_final, int(stir), pi_hash)
    p._keep = arr
    return p


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = ctypes.CDLL(LIB)
        L.ref_gl_mul.restype = ctypes.c_uint64
        L.ref_gl_mul.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.ref_gl_inv.restype = ctypes.c_uint64
        L.ref_gl_inv.argtypes = [ctypes.c_uint64]
        L.ref_gl_root_of_unity.restype = ctypes.c_uint64
        L.ref_gl_root_of_unity.argtypes = [sz]
        L.ref_set_threads.restype = None
        L.ref_set_threads.argtypes = [ctypes.c_int]
        L.ref_get_threads.restype = ctypes.c_int
        L.ref_stdrng_first_u64.restype = ctypes.c_uint64
        L.ref_stdrng_first_u64.argtypes = [ctypes.c_uint64]
        L.ref_fp6_mul.argtypes = [u64p, u64p, u64p]
        L.ref_fp6_inv.argtypes = [u64p, u64p]
        L.ref_ntt.argtypes = [u64p, sz, ctypes.c_int]
        L.ref_sha3_256.argtypes = [ctypes.c_char_p, sz, ctypes.c_char_p]
        L.ref_keccak_f1600.argtypes = [u64p]
        L.ref_chacha_block.argtypes = [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_uint32)]
        L.ref_lde_columns.argtypes = [u64p, sz, sz, sz, ctypes.c_uint64, u64p]
        L.ref_build_hash_rollup_trace.argtypes = [sz, u64p, sz, u64p]
        L.ref_merge_hash_rollup.argtypes = [u64p, sz, sz, ctypes.c_int, u64p]
        L.ref_merge_hash_rollup_lam.argtypes = [u64p, sz, sz, ctypes.c_int, u64p, u64p]
        L.ref_merge_nsec3_lam.argtypes = [u64p, sz, sz, ctypes.c_int, u64p, u64p]
        L.ref_merge_sha256_lam.argtypes = [u64p, sz, sz, sz, ctypes.c_int, u64p, u64p]
        L.ref_merge_sha512_lam.argtypes = [u64p, sz, sz, sz, ctypes.c_int, u64p, u64p]
        L.ref_merkle_root.argtypes = [u64p, sz, sz, sz, sz, ctypes.c_uint64, ctypes.c_char_p]
        L.ref_deep_fri_prove.argtypes = [u64p, sz, ctypes.POINTER(FriParams), ctypes.POINTER(u8p),
                                         ctypes.POINTER(sz)]
        L.ref_fri_layers.argtypes = [u64p, sz, ctypes.POINTER(FriParams), u64p, ctypes.c_char_p, ctypes.c_char_p]
        L.ref_free.argtypes = [ctypes.c_void_p]
        L.ref_zone_leaf_hashes.argtypes = [ctypes.c_uint64, sz, ctypes.c_char_p, ctypes.c_char_p]
        L.ref_zone_record_canonical.restype = sz
        L.ref_zone_record_canonical.argtypes = [ctypes.c_uint64, ctypes.c_char_p]
        L.ref_dns_merkle_root.argtypes = [ctypes.c_char_p, sz, ctypes.c_char_p]
        L.ref_shard_fs_binding.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint32, ctypes.c_char_p,
                                           ctypes.c_char_p]
        L.ref_prove_inner_shard.argtypes = [ctypes.c_char_p, ctypes.c_char_p, sz, ctypes.c_char_p, ctypes.c_int,
                                            ctypes.POINTER(u8p), ctypes.POINTER(sz), ctypes.c_char_p,
                                            ctypes.c_char_p, ctypes.c_char_p]
        L.ref_prove_outer_rollup.argtypes = [ctypes.c_char_p, sz, ctypes.c_char_p, ctypes.c_int,
                                             ctypes.POINTER(u8p), ctypes.POINTER(sz), ctypes.c_char_p]
        L.ref_prover_core.argtypes = [u64p, sz, sz, sz, ctypes.c_uint64, ctypes.POINTER(FriParams),
                                      ctypes.c_char_p, u64p, ctypes.POINTER(u8p), ctypes.POINTER(sz)]
        L.ref_sha256_n_blocks.restype = sz
        L.ref_sha256_n_blocks.argtypes = [sz]
        L.ref_sha256_trace_height.restype = sz
        L.ref_sha256_trace_height.argtypes = [sz]
        L.ref_build_sha256_trace.restype = ctypes.c_int
        L.ref_build_sha256_trace.argtypes = [ctypes.c_char_p, sz, u64p, ctypes.c_char_p]
        L.ref_sha256_constraints.argtypes = [u64p, u64p, sz, sz, u64p]
        L.ref_merge_sha256.argtypes = [u64p, sz, sz, sz, ctypes.c_int, u64p]
        L.ref_prove_ds_ksk_binding.argtypes = [ctypes.c_char_p, sz, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                               ctypes.POINTER(u8p), ctypes.POINTER(sz), ctypes.c_char_p,
                                               ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(sz),
                                               ctypes.POINTER(sz)]
        L.ref_sha512_n_blocks.restype = sz
        L.ref_sha512_n_blocks.argtypes = [sz]
        L.ref_sha512_trace_height.restype = sz
        L.ref_sha512_trace_height.argtypes = [sz]
        L.ref_build_sha512_trace.restype = ctypes.c_int
        L.ref_build_sha512_trace.argtypes = [ctypes.c_char_p, sz, u64p, ctypes.c_char_p]
        L.ref_sha512_constraints.argtypes = [u64p, u64p, sz, sz, u64p]
        L.ref_merge_sha512.argtypes = [u64p, sz, sz, sz, ctypes.c_int, u64p]
        L.ref_build_nsec3_trace.argtypes = [sz, u64p, sz, u64p]
        L.ref_merge_nsec3.argtypes = [u64p, sz, sz, ctypes.c_int, u64p]
        L.ref_nsec3_constraints.restype = None
        L.ref_nsec3_constraints.argtypes = [u64p, u64p, u64p]
        L.ref_prove_nsec3_completeness.restype = ctypes.c_int
        L.ref_prove_nsec3_completeness.argtypes = [ctypes.c_char_p, sz, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                                   ctypes.POINTER(u8p), ctypes.POINTER(sz), ctypes.c_char_p,
                                                   ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(sz)]
        L.ref_rsa_exp_width.restype = sz
        L.ref_rsa_exp_constraints_per_record.restype = sz
        L.ref_fill_rsa_exp.restype = ctypes.c_int
        L.ref_fill_rsa_exp.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, sz, sz, u64p]
        L.ref_rsa_exp_constraints.argtypes = [u64p, u64p, sz, sz, u64p]
        L.ref_merge_rsa.argtypes = [u64p, sz, sz, sz, u64p, u64p]
        L.ref_rsa_pkcs1_em_sha256.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        L.ref_prove_rsa_exp.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                        ctypes.POINTER(u8p), ctypes.POINTER(sz), ctypes.c_char_p, ctypes.c_char_p]
        L.ref_p256_width.restype = sz
        L.ref_p256_n_constraints.restype = sz
        L.ref_p256_columns.argtypes = [ctypes.POINTER(sz)]
        L.ref_p256_gadgets.restype = ctypes.c_int
        L.ref_p256_gadgets.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(sz)]
        L.ref_fill_p256.restype = ctypes.c_int
        L.ref_fill_p256.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, sz, ctypes.c_char_p, sz, sz,
                                    u64p]
        L.ref_p256_constraints.argtypes = [u64p, u64p, sz, sz, u64p]
        L.ref_merge_p256.argtypes = [u64p, sz, sz, u64p, u64p]
        L.ref_p256_ecdsa_bits.restype = ctypes.c_int
        L.ref_p256_ecdsa_bits.argtypes = [ctypes.c_char_p] * 7
        L.ref_p256_ecdsa_trace.restype = ctypes.c_int
        L.ref_p256_ecdsa_trace.argtypes = [ctypes.c_char_p] * 5 + [sz, u64p, ctypes.c_char_p]
        L.ref_prove_p256_ecdsa.restype = ctypes.c_int
        L.ref_prove_p256_ecdsa.argtypes = [ctypes.c_char_p] * 5 + [ctypes.POINTER(u8p), ctypes.POINTER(sz),
                                                                    ctypes.c_char_p, ctypes.c_char_p]
        vp = ctypes.c_void_p
        L.ref_ed25519_ctx_new.restype = vp
        L.ref_ed25519_ctx_new.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, sz, sz]
        L.ref_ed25519_ctx_free.argtypes = [vp]
        for f in ("width", "height", "n_constraints"):
            getattr(L, "ref_ed25519_" + f).restype = sz
            getattr(L, "ref_ed25519_" + f).argtypes = [vp]
        L.ref_ed25519_v16_constraints.restype = sz
        L.ref_ed25519_v16_constraints.argtypes = [sz]
        L.ref_ed25519_rows.argtypes = [vp, ctypes.POINTER(sz)]
        L.ref_ed25519_columns.argtypes = [vp, ctypes.POINTER(sz)]
        L.ref_ed25519_public.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]
        L.ref_ed25519_points.argtypes = [vp, ctypes.c_char_p]
        L.ref_ed25519_fill.restype = ctypes.c_int
        L.ref_ed25519_fill.argtypes = [vp, u64p]
        L.ref_ed25519_constraints.argtypes = [vp, u64p, u64p, sz, u64p]
        L.ref_merge_ed25519.restype = ctypes.c_int
        L.ref_merge_ed25519.argtypes = [vp, u64p, sz, u64p, u64p]
        L.ref_ed25519_verify.restype = ctypes.c_int
        L.ref_ed25519_verify.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, sz]
        L.ref_sha512.argtypes = [ctypes.c_char_p, sz, ctypes.c_char_p]
        L.ref_prove_zsk_ksk_binding_v2.restype = ctypes.c_int
        L.ref_prove_zsk_ksk_binding_v2.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, sz,
                                                   ctypes.c_char_p, ctypes.c_char_p, sz, sz, ctypes.POINTER(u8p),
                                                   ctypes.POINTER(sz), ctypes.c_char_p, ctypes.c_char_p]
        L.ref_set_ldt.argtypes = [ctypes.c_int]
        L.ref_set_ldt.restype = None
        L.ref_deep_fri_verify.restype = ctypes.c_int
        L.ref_deep_fri_verify.argtypes = [ctypes.POINTER(FriParams), ctypes.c_char_p, sz]
        L.ref_merkle_verify_opening.restype = ctypes.c_int
        L.ref_merkle_verify_opening.argtypes = [sz, sz, ctypes.c_uint64, ctypes.c_char_p, ctypes.c_char_p, sz]
        L.ref_merkle_open.restype = ctypes.c_int
        L.ref_merkle_open.argtypes = [u64p, sz, sz, sz, sz, ctypes.c_uint64, sz, ctypes.POINTER(u8p),
                                      ctypes.POINTER(sz)]
        _lib = L
    return _lib


def ptr(a):
    assert a.dtype == np.uint64 and a.flags.c_contiguous
    return a.ctypes.data_as(u64p)


def take_bytes(pp, pl):
    b = ctypes.string_at(pp, pl.value)
    lib().ref_free(pp)
    return b


def ntt(a, inverse=False):
    a = np.ascontiguousarray(a, dtype=np.uint64).copy()
    lib().ref_ntt(ptr(a), a.size, int(inverse))
    return a


def lde_columns(cols, blowup, shift=1):
    cols = np.ascontiguousarray(cols, dtype=np.uint64)
    w, n = cols.shape
    out = np.empty((w, n * blowup), dtype=np.uint64)
    rc = lib().ref_lde_columns(ptr(cols), w, n, blowup, shift, ptr(out))
    if rc:
        raise ValueError("lde_columns rc=%d" % rc)
    return out


def hash_rollup_trace(n_trace, leaves):
    leaves = np.ascontiguousarray(leaves, dtype=np.uint64)
    t = np.empty((4, n_trace), dtype=np.uint64)
    lib().ref_build_hash_rollup_trace(n_trace, ptr(leaves), leaves.size, ptr(t))
    return t


def _lam(coeffs):
    if coeffs is None:
        return None, None
    c = np.ascontiguousarray(coeffs, dtype=np.uint64)
    return c, ptr(c)


def merge_hash_rollup(lde, n_trace, blowup, phi_mode=1, coeffs=None):
    lde = np.ascontiguousarray(lde, dtype=np.uint64)
    out = np.empty(n_trace * blowup, dtype=np.uint64)
    keep, cp = _lam(coeffs)
    lib().ref_merge_hash_rollup_lam(ptr(lde), n_trace, blowup, phi_mode, cp, ptr(out))
    return out


def deep_fri_prove(f0, params):
    f0 = np.ascontiguousarray(f0, dtype=np.uint64)
    pp, pl = u8p(), sz()
    rc = lib().ref_deep_fri_prove(ptr(f0), f0.size, ctypes.byref(params), ctypes.byref(pp), ctypes.byref(pl))
    if rc:
        raise ValueError("deep_fri_prove rc=%d" % rc)
    return take_bytes(pp, pl)


class ldt_fri:
    """with O.ldt_fri(): the swarm-dns provers of the oracle run LdtMode::Fri
    (classic DEEP-FRI, arity-2 folds, prover.rs:53-67)"""

    def __enter__(self):
        lib().ref_set_ldt(1)

    def __exit__(self, *a):
        lib().ref_set_ldt(0)


def deep_fri_verify(params, proof):
    """1 accept, 0 reject, -1 bytes that do not deserialize (fri.rs:2136-2710)"""
    return lib().ref_deep_fri_verify(ctypes.byref(params), bytes(proof), len(proof))


def merkle_root(vals, arity, label, depth=None):
    vals = np.ascontiguousarray(vals, dtype=np.uint64)
    n, vpl = vals.shape
    if depth is None:
        depth = merkle_depth(n, arity)
    out = ctypes.create_string_buffer(32)
    lib().ref_merkle_root(ptr(vals), n, vpl, arity, depth, label, out)
    return out.raw


def merkle_open(vals, arity, label, index, depth=None):
    vals = np.ascontiguousarray(vals, dtype=np.uint64)
    n, vpl = vals.shape
    if depth is None:
        depth = merkle_depth(n, arity)
    pp, pl = u8p(), sz()
    rc = lib().ref_merkle_open(ptr(vals), n, vpl, arity, depth, label, index, ctypes.byref(pp), ctypes.byref(pl))
    if rc:
        raise ValueError("merkle_open rc=%d" % rc)
    return take_bytes(pp, pl)


def merkle_verify_opening(arity, depth, label, root, opening):
    return lib().ref_merkle_verify_opening(arity, depth, label, root, bytes(opening), len(opening))


def merkle_depth(leaves, arity):
    """fri.rs:950-959"""
    d, cur = 1, leaves
    while cur > arity:
        cur = (cur + arity - 1) // arity
        d += 1
    return d


def pick_arity(n, m):
    """pick_arity_for_layer(n, m).max(2) (fri.rs:1094-1103)"""
    for a in (128, 64, 32, 16, 8, 4):
        if m >= a and n % a == 0:
            return a
    return 2


def zone_leaf_hashes(first, count, salt):
    out = ctypes.create_string_buffer(32 * count)
    lib().ref_zone_leaf_hashes(first, count, salt, out)
    return out.raw


def dns_merkle_root(leaves):
    out = ctypes.create_string_buffer(32)
    lib().ref_dns_merkle_root(leaves, len(leaves) // 32, out)
    return out.raw


def shard_fs_binding(pk_hash, job_id, shard_id, nonce):
    out = ctypes.create_string_buffer(32)
    lib().ref_shard_fs_binding(pk_hash, job_id, shard_id, nonce, out)
    return out.raw


def prove_inner_shard(salt, leaf_hashes, fs_binding, phi_mode=1):
    pp, pl = u8p(), sz()
    pi, mr, rf = (ctypes.create_string_buffer(32) for _ in range(3))
    rc = lib().ref_prove_inner_shard(salt, leaf_hashes, len(leaf_hashes) // 32, fs_binding, phi_mode,
                                     ctypes.byref(pp), ctypes.byref(pl), pi, mr, rf)
    if rc:
        raise ValueError("prove_inner_shard rc=%d" % rc)
    return take_bytes(pp, pl), pi.raw, mr.raw, rf.raw


def prove_outer_rollup(pi_hashes, pk_hash, phi_mode=1):
    pp, pl = u8p(), sz()
    rf = ctypes.create_string_buffer(32)
    rc = lib().ref_prove_outer_rollup(pi_hashes, len(pi_hashes) // 32, pk_hash, phi_mode, ctypes.byref(pp),
                                      ctypes.byref(pl), rf)
    if rc:
        raise ValueError("prove_outer_rollup rc=%d" % rc)
    return take_bytes(pp, pl), rf.raw


SHA256_W, SHA256_K = 756, 766


def sha256_trace(msg):
    """(trace 756 x h, n_blocks, digest) of build_sha256_trace_multi."""
    h = lib().ref_sha256_trace_height(len(msg))
    tr = np.zeros((SHA256_W, h), dtype=np.uint64)
    dg = ctypes.create_string_buffer(32)
    nb = lib().ref_build_sha256_trace(bytes(msg), len(msg), ptr(tr), dg)
    return tr, nb, dg.raw


def sha256_constraints(cur, nxt, row, n_blocks):
    cur = np.ascontiguousarray(cur, dtype=np.uint64)
    nxt = np.ascontiguousarray(nxt, dtype=np.uint64)
    out = np.zeros(SHA256_K, dtype=np.uint64)
    lib().ref_sha256_constraints(ptr(cur), ptr(nxt), row, n_blocks, ptr(out))
    return out


def merge_sha256(lde, n_trace, blowup, n_blocks, phi_mode=1, coeffs=None):
    lde = np.ascontiguousarray(lde, dtype=np.uint64)
    out = np.zeros(n_trace * blowup, dtype=np.uint64)
    keep, cp = _lam(coeffs)
    lib().ref_merge_sha256_lam(ptr(lde), n_trace, blowup, n_blocks, phi_mode, cp, ptr(out))
    return out


def prove_ds_ksk_binding(dnskey, parent_ds_hash, fs_binding, phi_mode=1):
    """(proof, pi_hash, asserted_digest, root_f0, n_blocks, n_trace)"""
    pp, pl = u8p(), sz()
    pi, dg, rf = (ctypes.create_string_buffer(32) for _ in range(3))
    nb, nt = sz(), sz()
    rc = lib().ref_prove_ds_ksk_binding(bytes(dnskey), len(dnskey), bytes(parent_ds_hash), bytes(fs_binding),
                                        phi_mode, ctypes.byref(pp), ctypes.byref(pl), pi, dg, rf,
                                        ctypes.byref(nb), ctypes.byref(nt))
    assert rc == 0
    return take_bytes(pp, pl), pi.raw, dg.raw, rf.raw, nb.value, nt.value


SHA512_W, SHA512_K = 1510, 1526


def sha512_trace(msg):
    """(trace 1510 x h, n_blocks, digest) of build_sha512_trace_multi."""
    h = lib().ref_sha512_trace_height(len(msg))
    tr = np.zeros((SHA512_W, h), dtype=np.uint64)
    dg = ctypes.create_string_buffer(64)
    nb = lib().ref_build_sha512_trace(bytes(msg), len(msg), ptr(tr), dg)
    return tr, nb, dg.raw


def sha512_constraints(cur, nxt, row, n_blocks):
    cur = np.ascontiguousarray(cur, dtype=np.uint64)
    nxt = np.ascontiguousarray(nxt, dtype=np.uint64)
    out = np.zeros(SHA512_K, dtype=np.uint64)
    lib().ref_sha512_constraints(ptr(cur), ptr(nxt), row, n_blocks, ptr(out))
    return out


def merge_sha512(lde, n_trace, blowup, n_blocks, phi_mode=1, coeffs=None):
    lde = np.ascontiguousarray(lde, dtype=np.uint64)
    out = np.zeros(n_trace * blowup, dtype=np.uint64)
    keep, cp = _lam(coeffs)
    lib().ref_merge_sha512_lam(ptr(lde), n_trace, blowup, n_blocks, phi_mode, cp, ptr(out))
    return out


RSA_W, RSA_K = 11041, 11280


def _be256(x):
    return int(x).to_bytes(256, "big")


def rsa_exp_trace(records, n_trace=32):
    """(w x n_trace) trace of fill_rsa_stacked / fill_rsa_exp_multirow for
    records [(n, s, em), ...] given as Python ints; raises on a bad input."""
    tr = np.zeros((RSA_W * len(records), n_trace), dtype=np.uint64)
    for r, (n, s, em) in enumerate(records):
        rc = lib().ref_fill_rsa_exp(_be256(n), _be256(s), _be256(em), n_trace, RSA_W * r, ptr(tr))
        if rc:
            raise ValueError("ref_fill_rsa_exp: %d" % rc)
    return tr


def rsa_exp_constraints(cur, nxt, row, n_trace):
    cur = np.ascontiguousarray(cur, dtype=np.uint64)
    nxt = np.ascontiguousarray(nxt, dtype=np.uint64)
    out = np.zeros(RSA_K, dtype=np.uint64)
    lib().ref_rsa_exp_constraints(ptr(cur), ptr(nxt), row, n_trace, ptr(out))
    return out


def merge_rsa(lde, n_trace, blowup, n_records, coeffs=None):
    lde = np.ascontiguousarray(lde, dtype=np.uint64)
    out = np.zeros(n_trace * blowup, dtype=np.uint64)
    cp = None
    if coeffs is not None:
        coeffs = np.ascontiguousarray(coeffs, dtype=np.uint64)
        cp = ptr(coeffs)
    lib().ref_merge_rsa(ptr(lde), n_trace, blowup, n_records, cp, ptr(out))
    return out


def rsa_pkcs1_em_sha256(digest):
    em = ctypes.create_string_buffer(256)
    lib().ref_rsa_pkcs1_em_sha256(bytes(digest), em)
    return em.raw


def prove_rsa_exp(n, s, em, pis):
    """(proof, root_f0, pi_hash) of the per-record RSA-2048 prover"""
    pp, pl = u8p(), sz()
    rf, pi = ctypes.create_string_buffer(32), ctypes.create_string_buffer(32)
    rc = lib().ref_prove_rsa_exp(_be256(n), _be256(s), _be256(em), bytes(pis), ctypes.byref(pp), ctypes.byref(pl),
                                 rf, pi)
    assert rc == 0, rc
    return take_bytes(pp, pl), rf.raw, pi.raw


def nsec3_trace(n_trace, chain):
    """build_nsec3_chain_trace; chain = [(owner limbs[4], next limbs[4]), ...] as u64"""
    ch = np.array([list(o) + list(x) for o, x in chain], dtype=np.uint64).reshape(-1) if chain else \
        np.zeros(8, dtype=np.uint64)
    tr = np.zeros((8, n_trace), dtype=np.uint64)
    lib().ref_build_nsec3_trace(n_trace, ptr(ch), len(chain), ptr(tr))
    return tr


def nsec3_constraints(cur, nxt):
    """the 4 chain-link constraints of one row"""
    c = np.ascontiguousarray(cur, dtype=np.uint64)
    x = np.ascontiguousarray(nxt, dtype=np.uint64)
    out = np.zeros(4, dtype=np.uint64)
    lib().ref_nsec3_constraints(ptr(c), ptr(x), ptr(out))
    return out


def merge_nsec3(lde, n_trace, blowup, phi_mode=1, coeffs=None):
    lde = np.ascontiguousarray(lde, dtype=np.uint64)
    out = np.zeros(n_trace * blowup, dtype=np.uint64)
    keep, cp = _lam(coeffs)
    lib().ref_merge_nsec3_lam(ptr(lde), n_trace, blowup, phi_mode, cp, ptr(out))
    return out


def prove_nsec3_completeness(records, salt, fs_binding, phi_mode=1):
    """records = [(owner 32 B, next 32 B)]; (proof, pi_hash, chain_root, root_f0, n_trace) or rc < 0"""
    blob = b"".join(bytes(o) + bytes(x) for o, x in records)
    pp, pl, nt = u8p(), sz(), sz()
    pi, cr, rf = (ctypes.create_string_buffer(32) for _ in range(3))
    rc = lib().ref_prove_nsec3_completeness(blob, len(records), bytes(salt), bytes(fs_binding), phi_mode,
                                            ctypes.byref(pp), ctypes.byref(pl), pi, cr, rf, ctypes.byref(nt))
    if rc:
        return rc
    return take_bytes(pp, pl), pi.raw, cr.raw, rf.raw, nt.value


def _be32(x):
    return int(x).to_bytes(32, "big")


def p256_width():
    return lib().ref_p256_width()


def p256_n_constraints():
    return lib().ref_p256_n_constraints()


def p256_columns():
    """dict of the column bases the tests name (see ref_p256_columns)"""
    a = (sz * 26)()
    lib().ref_p256_columns(a)
    v = list(a)
    return {"bit": v[0:2], "sel": [v[2:5], v[5:8]], "rproj": v[8:14], "acc": [v[14:17], v[17:20]],
            "base": [v[20:23], v[23:26]]}


def p256_gadgets(ch):
    """[(kind, (c, cb, q/dl, qb/db, carries/pos, neg)), ...] of chain ch in constraint order;
    kind 0 mul, 1 add, 2 sub, 3 freeze, 4 select"""
    kinds = (ctypes.c_int * 256)()
    cols = (sz * (6 * 256))()
    n = lib().ref_p256_gadgets(ch, kinds, cols)
    return [(kinds[i], tuple(cols[6 * i:6 * i + 6])) for i in range(n)]


def p256_trace(points, rproj, a_bits, b_bits, n_trace):
    """fill_ecdsa_double_multirow: points = 12 ints (a_initial xyz, a_base xyz,
    b_initial xyz, b_base xyz), rproj = 6 ints, bits lists of 0/1"""
    tr = np.zeros((p256_width(), n_trace), dtype=np.uint64)
    rc = lib().ref_fill_p256(b"".join(_be32(v) for v in points), b"".join(_be32(v) for v in rproj),
                             bytes(a_bits), len(a_bits), bytes(b_bits), len(b_bits), n_trace, ptr(tr))
    if rc:
        raise ValueError("ref_fill_p256: %d" % rc)
    return tr


def p256_constraints(cur, nxt, row, n_trace):
    cur = np.ascontiguousarray(cur, dtype=np.uint64)
    nxt = np.ascontiguousarray(nxt, dtype=np.uint64)
    out = np.zeros(p256_n_constraints(), dtype=np.uint64)
    lib().ref_p256_constraints(ptr(cur), ptr(nxt), row, n_trace, ptr(out))
    return out


def merge_p256(lde, n_trace, blowup, coeffs=None):
    lde = np.ascontiguousarray(lde, dtype=np.uint64)
    out = np.zeros(n_trace * blowup, dtype=np.uint64)
    cp = None
    if coeffs is not None:
        coeffs = np.ascontiguousarray(coeffs, dtype=np.uint64)
        cp = ptr(coeffs)
    lib().ref_merge_p256(ptr(lde), n_trace, blowup, cp, ptr(out))
    return out


def p256_ecdsa_bits(qx, qy, r, s, digest):
    """(u1 bits, u2 bits) MSB first, or the negative code for a bad record"""
    u1, u2 = ctypes.create_string_buffer(256), ctypes.create_string_buffer(256)
    rc = lib().ref_p256_ecdsa_bits(_be32(qx), _be32(qy), _be32(r), _be32(s), bytes(digest), u1, u2)
    if rc:
        return rc
    return list(u1.raw), list(u2.raw)


def p256_ecdsa_trace(qx, qy, r, s, digest):
    """(trace, r_proj ints) of the per-record ECDSA prover's 256-row trace"""
    tr = np.zeros((p256_width(), 256), dtype=np.uint64)
    rp = ctypes.create_string_buffer(6 * 32)
    rc = lib().ref_p256_ecdsa_trace(_be32(qx), _be32(qy), _be32(r), _be32(s), bytes(digest), 256, ptr(tr), rp)
    if rc:
        raise ValueError("ref_p256_ecdsa_trace: %d" % rc)
    return tr, [int.from_bytes(rp.raw[32 * i:32 * i + 32], "big") for i in range(6)]


def prove_p256_ecdsa(qx, qy, r, s, digest):
    """(proof, root_f0, pi_hash) of the per-record ECDSA-P256 prover"""
    pp, pl = u8p(), sz()
    rf, pi = ctypes.create_string_buffer(32), ctypes.create_string_buffer(32)
    rc = lib().ref_prove_p256_ecdsa(_be32(qx), _be32(qy), _be32(r), _be32(s), bytes(digest), ctypes.byref(pp),
                                    ctypes.byref(pl), rf, pi)
    assert rc == 0, rc
    return take_bytes(pp, pl), rf.raw, pi.raw


ED_ROWS = ("k_hash", "reduce", "dec_r", "dec_a", "sb_first", "ka_first", "residual", "dbl1", "dbl2", "dbl3",
           "result")
ED_COLS = ("digest_bits", "reduce_r_bits", "dec_x", "acc_x", "cadd_out_x", "sblk_sb", "sblk_ka", "r_thread",
           "res_x", "res_y", "res_z", "res_t", "th_sb_x", "th_ka_x", "th_r_x", "th_r_t", "sub_x_c", "sub_yz_c",
           "dbl_in_x", "pa_r2_x3", "dec_sign", "dbl_sqa_carry_bits", "cadd_bit", "dec_x_bits", "reduce_q",
           "reduce_carry_bits")


class Ed25519:
    """One signature's v16 AIR in the oracle (layout, trace, evaluator, merge)."""

    def __init__(self, pk, sig, msg, k=256):
        self.pk, self.sig, self.msg, self.k = bytes(pk), bytes(sig), bytes(msg), k
        h = lib().ref_ed25519_ctx_new(self.pk, self.sig, self.msg, len(self.msg), k)
        if not h:
            raise ValueError("R or A does not decompress")
        self.h = h
        L = lib()
        self.width, self.height = L.ref_ed25519_width(h), L.ref_ed25519_height(h)
        self.n_constraints = L.ref_ed25519_n_constraints(h)
        a = (sz * 11)()
        L.ref_ed25519_rows(h, a)
        self.rows = dict(zip(ED_ROWS, list(a)))
        b = (sz * 26)()
        L.ref_ed25519_columns(h, b)
        self.cols = dict(zip(ED_COLS, list(b)))

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_ed25519_ctx_free(self.h)
            self.h = None

    def public(self):
        dg, kc = ctypes.create_string_buffer(64), ctypes.create_string_buffer(32)
        sb, kb = ctypes.create_string_buffer(self.k), ctypes.create_string_buffer(self.k)
        lib().ref_ed25519_public(self.h, dg, kc, sb, kb)
        return dg.raw, kc.raw, list(sb.raw), list(kb.raw)

    def points(self):
        """{R, A, sB, kA, residual_2, 8 residual_2}: (x, y, z, t) ints"""
        buf = ctypes.create_string_buffer(6 * 128)
        lib().ref_ed25519_points(self.h, buf)
        names = ("R", "A", "sB", "kA", "res2", "res2x8")
        out = {}
        for i, n in enumerate(names):
            out[n] = tuple(int.from_bytes(buf.raw[128 * i + 32 * j:128 * i + 32 * j + 32], "little") for j in range(4))
        return out

    def trace(self):
        tr = np.zeros((self.width, self.height), dtype=np.uint64)
        assert lib().ref_ed25519_fill(self.h, ptr(tr)) == 0
        return tr

    def constraints(self, cur, nxt, row):
        cur = np.ascontiguousarray(cur, dtype=np.uint64)
        nxt = np.ascontiguousarray(nxt, dtype=np.uint64)
        out = np.zeros(self.n_constraints, dtype=np.uint64)
        lib().ref_ed25519_constraints(self.h, ptr(cur), ptr(nxt), row, ptr(out))
        return out

    def merge(self, lde, blowup, coeffs=None):
        lde = np.ascontiguousarray(lde, dtype=np.uint64)
        out = np.zeros(self.height * blowup, dtype=np.uint64)
        cp = None
        if coeffs is not None:
            coeffs = np.ascontiguousarray(coeffs, dtype=np.uint64)
            cp = ptr(coeffs)
        lib().ref_merge_ed25519(self.h, ptr(lde), blowup, cp, ptr(out))
        return out


def ed25519_v16_constraints(k):
    return lib().ref_ed25519_v16_constraints(k)


def ed25519_verify(pk, sig, msg):
    return bool(lib().ref_ed25519_verify(bytes(pk), bytes(sig), bytes(msg), len(msg)))


def sha512(msg):
    out = ctypes.create_string_buffer(64)
    lib().ref_sha512(bytes(msg), len(msg), out)
    return out.raw


def prove_zsk_ksk_binding_v2(pk, sig, msg, fs_binding, merkle_root, k=256, blowup=32):
    """(proof, root_f0, pi_hash) of prove_zsk_ksk_binding_v2, or the negative code"""
    pp, pl = u8p(), sz()
    rf, pi = ctypes.create_string_buffer(32), ctypes.create_string_buffer(32)
    rc = lib().ref_prove_zsk_ksk_binding_v2(bytes(pk), bytes(sig), bytes(msg), len(msg), bytes(fs_binding),
                                           bytes(merkle_root), k, blowup, ctypes.byref(pp), ctypes.byref(pl), rf, pi)
    if rc:
        return rc
    return take_bytes(pp, pl), rf.raw, pi.raw


class threads:
    """with O.threads(n): run the oracle's rayon-site loops on n threads"""

    def __init__(self, n):
        self.n = n

    def __enter__(self):
        self.old = lib().ref_get_threads()
        lib().ref_set_threads(self.n)
        return self

    def __exit__(self, *a):
        lib().ref_set_threads(self.old)
