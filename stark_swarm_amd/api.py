"""ctypes binding of libstark_gpu.so and the reference-named host API."""
import ctypes
import enum
import hashlib
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstark_gpu.so")

P = 2**64 - 2**32 + 1
BLOWUP = 32          # crates/swarm-dns/src/prover.rs:32
NUM_QUERIES = 54     # prover.rs:33
SEED_Z = 0xDEEFBAAD  # prover.rs:34

PHI_SEQUENTIAL, PHI_RAYON, PHI_ONE_THREAD = 0, 1, 2
_LDT_STIR, _LDT_FRI = 0, 1
_AIR_HASH_ROLLUP = 5
_AIR_NSEC3_CHAIN = 6

u64p = ctypes.POINTER(ctypes.c_uint64)
u8p = ctypes.POINTER(ctypes.c_uint8)
sz = ctypes.c_size_t


class StarkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("stark_gpu error %d: %s" % (code, msg))
        self.code = code


class _FriParams(ctypes.Structure):
    _fields_ = [("schedule", ctypes.POINTER(ctypes.c_size_t)), ("n_layers", sz), ("r", sz), ("s0", sz),
                ("seed_z", ctypes.c_uint64), ("coeff_commit_final", ctypes.c_int), ("d_final", sz),
                ("stir", ctypes.c_int), ("public_inputs_hash", ctypes.c_char_p)]


class _DnsRecord(ctypes.Structure):
    _fields_ = [("domain", ctypes.c_char_p), ("domain_len", sz), ("record_type", ctypes.c_uint16),
                ("ttl", ctypes.c_uint32), ("rdata", ctypes.c_char_p), ("rdata_len", sz)]


class _InnerOut(ctypes.Structure):
    _fields_ = [("pi_hash", ctypes.c_uint8 * 32), ("merkle_root", ctypes.c_uint8 * 32),
                ("root_f0", ctypes.c_uint8 * 32), ("record_count", sz), ("n_trace", sz),
                ("proof", u8p), ("proof_len", sz)]


class _ShaOut(ctypes.Structure):
    _fields_ = [("digest", ctypes.c_uint8 * 64), ("root_f0", ctypes.c_uint8 * 32), ("n_blocks", sz),
                ("n_trace", sz), ("proof", u8p), ("proof_len", sz)]


class _RsaOut(ctypes.Structure):
    _fields_ = [("pi_hash", ctypes.c_uint8 * 32), ("root_f0", ctypes.c_uint8 * 32), ("n_trace", sz),
                ("proof", u8p), ("proof_len", sz)]


class _P256Chains(ctypes.Structure):
    _fields_ = [("pts", (ctypes.c_uint8 * 32) * 12), ("rproj", (ctypes.c_uint8 * 32) * 6), ("a_bits", u8p),
                ("k_a", sz), ("b_bits", u8p), ("k_b", sz), ("bind_outputs", ctypes.c_int)]


class _P256Out(ctypes.Structure):
    _fields_ = [("pi_hash", ctypes.c_uint8 * 32), ("root_f0", ctypes.c_uint8 * 32), ("rproj", (ctypes.c_uint8 * 32) * 6),
                ("n_trace", sz), ("proof", u8p), ("proof_len", sz)]


class _Nsec3Out(ctypes.Structure):
    _fields_ = [("pi_hash", ctypes.c_uint8 * 32), ("chain_root", ctypes.c_uint8 * 32), ("root_f0", ctypes.c_uint8 * 32),
                ("record_count", sz), ("n_trace", sz), ("proof", u8p), ("proof_len", sz)]


class _ZskOut(ctypes.Structure):
    _fields_ = [("pi_hash", ctypes.c_uint8 * 32), ("root_f0", ctypes.c_uint8 * 32), ("n_trace", sz),
                ("proof", u8p), ("proof_len", sz)]


class _DsKskOut(ctypes.Structure):
    _fields_ = [("pi_hash", ctypes.c_uint8 * 32), ("asserted_digest", ctypes.c_uint8 * 32),
                ("root_f0", ctypes.c_uint8 * 32), ("n_blocks", sz), ("n_trace", sz),
                ("proof", u8p), ("proof_len", sz)]


_lib = None


def lib():
    """Load libstark_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libstark_gpu.so is not built; run __graft_entry__.build() "
                          "(make -C stark_swarm_amd)")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    sig = {
        "stark_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(vp)]),
        "stark_ctx_destroy": (None, [vp]),
        "stark_last_error": (ctypes.c_char_p, [vp]),
        "stark_free": (None, [vp]),
        "stark_lde_columns": (ctypes.c_int, [vp, u64p, sz, sz, sz, u64p]),
        "stark_merge": (ctypes.c_int, [vp, ctypes.c_int, u64p, sz, sz, sz, u64p, sz, ctypes.c_int, u64p]),
        "stark_deep_fri_prove": (ctypes.c_int, [vp, u64p, sz, ctypes.POINTER(_FriParams), ctypes.POINTER(u8p),
                                                ctypes.POINTER(sz)]),
        "stark_merkle_commit": (ctypes.c_int, [vp, u64p, sz, sz, sz, ctypes.c_uint64, ctypes.c_char_p]),
        "stark_prove_inner_shard": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.POINTER(_DnsRecord), sz,
                                                   ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                                   ctypes.POINTER(_InnerOut)]),
        "stark_prove_inner_shard_batch": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.POINTER(_DnsRecord),
                                                         ctypes.POINTER(sz), sz, ctypes.c_char_p, ctypes.c_int,
                                                         ctypes.c_int, ctypes.POINTER(_InnerOut)]),
        "stark_deep_fri_prove_batch": (ctypes.c_int, [vp, u64p, sz, sz, ctypes.POINTER(_FriParams), ctypes.c_char_p,
                                                      ctypes.POINTER(u8p), ctypes.POINTER(sz)]),
        "stark_prove_outer_rollup": (ctypes.c_int, [vp, ctypes.c_char_p, sz, ctypes.c_char_p, ctypes.c_int,
                                                    ctypes.c_int, ctypes.POINTER(u8p), ctypes.POINTER(sz),
                                                    ctypes.c_char_p]),
        "stark_dev_alloc": (ctypes.c_int, [vp, sz, ctypes.POINTER(vp)]),
        "stark_dev_free": (ctypes.c_int, [vp, vp]),
        "stark_memcpy_h2d": (ctypes.c_int, [vp, vp, vp, sz]),
        "stark_memcpy_d2h": (ctypes.c_int, [vp, vp, vp, sz]),
        "stark_synchronize": (ctypes.c_int, [vp]),
        "stark_ntt_dev": (ctypes.c_int, [vp, vp, vp, sz, sz, ctypes.c_int]),
        "stark_lde_columns_dev": (ctypes.c_int, [vp, vp, sz, sz, sz, ctypes.c_uint64, vp]),
        "stark_prover_core_dev": (ctypes.c_int, [vp, vp, sz, sz, sz, ctypes.c_uint64, ctypes.POINTER(_FriParams),
                                                 vp, ctypes.c_char_p, ctypes.POINTER(u8p), ctypes.POINTER(sz)]),
        "stark_profile_enable": (ctypes.c_int, [vp, ctypes.c_int]),
        "stark_profile_reset": (ctypes.c_int, [vp]),
        "stark_profile_get": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double)]),
        "stark_profile_names": (sz, [vp, ctypes.c_char_p, sz]),
        "stark_set_autotune": (ctypes.c_int, [vp, ctypes.c_int]),
        "stark_set_ntt_split": (ctypes.c_int, [vp, ctypes.c_uint]),
        "stark_set_ntt_tile": (ctypes.c_int, [vp, ctypes.c_uint]),
        "stark_merge_sha_dev": (ctypes.c_int, [vp, ctypes.c_int, vp, sz, sz, sz, u64p, sz, ctypes.c_int, vp]),
        "stark_merge_sha_from_trace_dev": (ctypes.c_int, [vp, ctypes.c_int, vp, sz, sz, sz, u64p, sz, ctypes.c_int,
                                                          vp]),
        "stark_autotune_report": (sz, [ctypes.c_char_p, sz]),
        "stark_host_sha3_256": (None, [ctypes.c_char_p, sz, ctypes.c_char_p]),
        "stark_host_stdrng_first_u64": (ctypes.c_uint64, [ctypes.c_uint64]),
        "stark_host_dns_leaf_hash": (None, [ctypes.POINTER(_DnsRecord), ctypes.c_char_p, ctypes.c_char_p]),
        "stark_merge_sha256": (ctypes.c_int, [vp, u64p, sz, sz, sz, u64p, sz, ctypes.c_int, u64p]),
        "stark_prove_ds_ksk_binding": (ctypes.c_int, [vp, ctypes.c_char_p, sz, ctypes.c_char_p, ctypes.c_char_p,
                                                      ctypes.c_int, ctypes.c_int, ctypes.POINTER(_DsKskOut)]),
        "stark_prove_ds_ksk_binding_batch": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.POINTER(sz),
                                                            ctypes.POINTER(sz), sz, ctypes.c_char_p,
                                                            ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                                            ctypes.POINTER(_DsKskOut)]),
        "stark_host_sha256_trace_height": (sz, [sz]),
        "stark_host_sha512_trace_height": (sz, [sz]),
        "stark_host_build_sha512_trace": (sz, [ctypes.c_char_p, sz, u64p, ctypes.c_char_p]),
        "stark_merge_sha512": (ctypes.c_int, [vp, u64p, sz, sz, sz, u64p, sz, ctypes.c_int, u64p]),
        "stark_prove_sha_batch": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.POINTER(sz),
                                                 ctypes.POINTER(sz), ctypes.POINTER(sz), sz, ctypes.c_char_p,
                                                 ctypes.c_int, ctypes.POINTER(_ShaOut)]),
        "stark_host_build_sha256_trace": (sz, [ctypes.c_char_p, sz, u64p, ctypes.c_char_p]),
        "stark_prove_nsec3_completeness": (ctypes.c_int, [vp, ctypes.c_char_p, sz, ctypes.c_char_p, ctypes.c_char_p,
                                                          ctypes.c_int, ctypes.c_int, ctypes.POINTER(_Nsec3Out)]),
        "stark_prove_nsec3_completeness_batch": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.POINTER(sz), sz,
                                                                ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                                                ctypes.c_int, ctypes.POINTER(_Nsec3Out)]),
        "stark_merge_rsa_exp": (ctypes.c_int, [vp, u64p, sz, sz, sz, u64p, sz, u64p]),
        "stark_prove_rsa_exp_batch": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, sz, sz,
                                                     sz, ctypes.c_char_p, ctypes.POINTER(_RsaOut)]),
        "stark_host_build_rsa_exp_trace": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, sz,
                                                          sz, u64p]),
        "stark_host_rsa_pkcs1_em_sha256": (None, [ctypes.c_char_p, ctypes.c_char_p]),
        "stark_merge_ecdsa_p256": (ctypes.c_int, [vp, u64p, sz, sz, u64p, sz, u64p]),
        "stark_prove_p256_chains_batch": (ctypes.c_int, [vp, ctypes.POINTER(_P256Chains), sz, sz, ctypes.c_char_p,
                                                         ctypes.POINTER(_P256Out)]),
        "stark_prove_ecdsa_p256_batch": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                                        ctypes.c_char_p, ctypes.c_char_p, sz,
                                                        ctypes.POINTER(_P256Out)]),
        "stark_host_build_ecdsa_p256_trace": (ctypes.c_int, [ctypes.POINTER(_P256Chains), sz, u64p]),
        "stark_ed25519_shape": (ctypes.c_int, [sz, sz, ctypes.POINTER(sz), ctypes.POINTER(sz), ctypes.POINTER(sz)]),
        "stark_prove_zsk_ksk_binding_v2_batch": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                                                ctypes.POINTER(sz), ctypes.POINTER(sz),
                                                                ctypes.c_char_p, ctypes.c_char_p, sz, sz, sz, sz,
                                                                ctypes.c_int, ctypes.POINTER(_ZskOut)]),
        "stark_merge_ed25519": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, sz, sz, u64p,
                                               sz, u64p, sz, u64p]),
        "stark_build_ed25519_trace": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, sz, sz,
                                                     u64p]),
        "stark_host_build_ed25519_trace": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, sz, sz,
                                                          u64p]),
        "stark_host_ed25519_verify": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, sz]),
        "stark_host_ed25519_sign": (None, [ctypes.c_char_p, ctypes.c_char_p, sz, ctypes.c_char_p, ctypes.c_char_p]),
        "stark_merkle_open": (ctypes.c_int, [vp, u64p, sz, sz, sz, ctypes.c_uint64, ctypes.POINTER(sz), sz,
                                             ctypes.c_char_p, ctypes.POINTER(u8p), ctypes.POINTER(sz),
                                             ctypes.POINTER(sz)]),
        "stark_merkle_verify_opening": (ctypes.c_int, [sz, sz, ctypes.c_uint64, ctypes.c_char_p, ctypes.c_char_p,
                                                       sz]),
        "stark_host_merkle_leaf_hash": (None, [sz, ctypes.c_uint64, sz, u64p, sz, ctypes.c_char_p]),
        "stark_deep_fri_verify": (ctypes.c_int, [ctypes.POINTER(_FriParams), ctypes.c_char_p, sz]),
        "stark_deep_fri_verify_batch": (ctypes.c_int, [ctypes.POINTER(_FriParams), ctypes.POINTER(ctypes.c_char_p),
                                                       ctypes.POINTER(sz), sz, ctypes.c_char_p,
                                                       ctypes.POINTER(ctypes.c_int)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _u64(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a, a.ctypes.data_as(u64p)


class DeviceBuffer:
    """Device memory owned by a Context (caching allocator)."""

    def __init__(self, ctx, nbytes):
        self.ctx, self.nbytes = ctx, nbytes
        p = ctypes.c_void_p()
        ctx._check(lib().stark_dev_alloc(ctx.h, nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        self.ctx._check(lib().stark_memcpy_h2d(self.ctx.h, self.ptr, arr.ctypes.data, arr.nbytes))
        return self

    def download(self, shape, dtype=np.uint64, offset=0):
        """`shape` values from byte `offset` of the buffer"""
        out = np.empty(shape, dtype=dtype)
        assert offset >= 0 and offset + out.nbytes <= self.nbytes
        self.ctx._check(lib().stark_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr + offset, out.nbytes))
        return out

    def free(self):
        if self.ptr is not None:
            lib().stark_dev_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One device context: a HIP stream, twiddle tables and an allocator."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        rc = lib().stark_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise StarkError(rc, "stark_ctx_create(device=%d) failed (no gfx950 device visible?)" % device)
        self.h = h.value

    def _check(self, rc):
        if rc != 0:
            raise StarkError(rc, lib().stark_last_error(self.h).decode(errors="replace"))

    def close(self):
        if self.h:
            lib().stark_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def synchronize(self):
        self._check(lib().stark_synchronize(self.h))

    # profiling (HIP events around each launch on the context stream)
    def profile(self, on=True):
        self._check(lib().stark_profile_enable(self.h, int(on)))

    def profile_reset(self):
        self._check(lib().stark_profile_reset(self.h))

    def profile_stats(self):
        n = lib().stark_profile_names(self.h, None, 0)
        buf = ctypes.create_string_buffer(n + 1)
        lib().stark_profile_names(self.h, buf, n + 1)
        out = {}
        for name in buf.value.decode().split("\n"):
            if not name:
                continue
            ms, la, by = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_double()
            lib().stark_profile_get(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(la), ctypes.byref(by))
            out[name] = {"ms": ms.value, "launches": la.value, "bytes": by.value}
        return out

    # launch-plan autotuner (the split of a transform's two passes)
    def autotune(self, on=True):
        """Time the admissible plans (pass split, LDS tile) on the first NTT /
        LDE of each shape and keep the fastest (process-wide); off: default
        plans."""
        self._check(lib().stark_set_autotune(self.h, int(on)))

    def ntt_split(self, log_first_pass):
        """Pin the first pass of every two-pass transform that admits it to
        2^log_first_pass points (0 = no pin)."""
        self._check(lib().stark_set_ntt_split(self.h, int(log_first_pass)))

    def ntt_tile(self, tile_log):
        """Pin the LDS tile of every tunable transform that admits it: 13
        (2^13 elements on 512 lanes) or 14 (2^14 on 1024 lanes, passes of 2^8
        points and more); 0 = no pin."""
        self._check(lib().stark_set_ntt_tile(self.h, int(tile_log)))


def set_autotune_default(on=True):
    """Autotune for every context created afterwards (the proving server's
    worker contexts included)."""
    rc = lib().stark_set_autotune(None, int(on))
    if rc != 0:
        raise StarkError(rc, "stark_set_autotune")


def autotune_report():
    """{(kind, log_points, log_blowup, log_cols): ((L1, tile_log) chosen,
    {(L1, tile_log): ms})} of the shapes tuned so far in this process."""
    n = lib().stark_autotune_report(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    lib().stark_autotune_report(buf, n + 1)
    out = {}
    for line in buf.value.decode().splitlines():
        head, times = line.split("|")
        shape, best = head.split("->")
        kind, L, LB, lc = shape.split()
        def plan(x):
            a, b = x.split(".")
            return int(a), int(b)
        ms = {plan(a): float(b) for a, b in (t.split(":") for t in times.split())}
        out[(kind, int(L), int(LB), int(lc))] = (plan(best.strip()), ms)
    return out


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context(int(os.environ.get("STARK_GPU_DEVICE", "0")))
    return _default


def _take(ptr, n):
    b = ctypes.string_at(ptr, n)
    lib().stark_free(ptr)
    return b


# --------------------------------------------------------------------- LDE / NTT
def lde_trace_columns(columns, trace_len, blowup, ctx=None):
    """lde_trace_columns (trace_import.rs:251-281): w columns of trace_len values ->
    w x (trace_len * blowup) evaluations on <w_n0>. Raises ValueError where the
    reference returns Err(String)."""
    if len(columns) == 0:
        raise ValueError("no columns provided")
    cols = np.asarray(columns, dtype=np.uint64) if not isinstance(columns, np.ndarray) else columns
    if cols.ndim != 2 or cols.shape[1] != trace_len:
        raise ValueError("expected %d rows per column, got shape %s" % (trace_len, cols.shape))
    if trace_len <= 0 or trace_len & (trace_len - 1):
        raise ValueError("trace_len %d must be a power of 2" % trace_len)
    if blowup < 2 or blowup & (blowup - 1):
        raise ValueError("blowup %d must be a power-of-2 >= 2" % blowup)
    ctx = ctx or default_context()
    cols, cp = _u64(cols)
    out = np.empty((cols.shape[0], trace_len * blowup), dtype=np.uint64)
    ctx._check(lib().stark_lde_columns(ctx.h, cp, cols.shape[0], trace_len, blowup, out.ctypes.data_as(u64p)))
    return out


def ntt(values, inverse=False, ctx=None):
    """Radix2EvaluationDomain::{fft, ifft} over each row of a (ncols, 2^k) array."""
    ctx = ctx or default_context()
    a = np.ascontiguousarray(np.atleast_2d(values), dtype=np.uint64)
    ncols, n = a.shape
    lg = n.bit_length() - 1
    if n != 1 << lg:
        raise ValueError("size must be a power of two")
    d = ctx.alloc(a.nbytes).upload(a)
    o = ctx.alloc(a.nbytes)
    ctx._check(lib().stark_ntt_dev(ctx.h, d.ptr, o.ptr, lg, ncols, int(inverse)))
    r = o.download(a.shape)
    d.free()
    o.free()
    return r.reshape(np.shape(values))


# --------------------------------------------------------------------- merge
class AirType(enum.Enum):
    HashRollup = _AIR_HASH_ROLLUP  # air_workloads.rs:83 (w = 4, 3 constraints, degree 2)
    Nsec3Chain = _AIR_NSEC3_CHAIN  # air_workloads.rs:117 (w = 8, 4 constraints, degree 1)

    def width(self):
        return 4 if self is AirType.HashRollup else 8

    def num_constraints(self):
        return 3 if self is AirType.HashRollup else 4

    def max_constraint_degree(self):
        return 2 if self is AirType.HashRollup else 1


@dataclass
class CompositionInfo:
    phi_degree_bound: int
    quotient_degree_bound: int
    rate: float
    num_constraints: int
    max_constraint_degree: int
    trace_width: int


def deep_ali_merge_general(trace_evals_on_lde, combination_coeffs, air, omega, n_trace, blowup,
                           phi_mode=PHI_RAYON, ctx=None):
    """deep_ali_merge_general (lib.rs:343-437). phi_mode selects the reference
    build's behaviour: PHI_RAYON (the swarm-dns default build on a multi-core
    host) leaves phi == 0 when n < 4096. Contract violations raise
    AssertionError as the reference panics."""
    air = AirType(air) if not isinstance(air, AirType) else air
    w, k = air.width(), air.num_constraints()
    n = n_trace * blowup
    lde = np.asarray(trace_evals_on_lde, dtype=np.uint64)
    assert lde.shape[0] == w, "trace width mismatch"
    assert len(combination_coeffs) == k, \
        "need one combination coefficient per constraint, got %d for %d constraints" % (len(combination_coeffs), k)
    assert lde.shape[1] == n, "trace column length mismatch"
    assert int(omega) == pow(7, (P - 1) // n, P), "omega must generate the n-point domain"
    ctx = ctx or default_context()
    lde, lp = _u64(lde)
    co, cp = _u64(np.asarray(combination_coeffs, dtype=np.uint64))
    out = np.empty(n, dtype=np.uint64)
    ctx._check(lib().stark_merge(ctx.h, air.value, lp, w, n_trace, blowup, cp, k, phi_mode, out.ctypes.data_as(u64p)))
    max_deg = air.max_constraint_degree()
    phi_b = max_deg * n_trace
    qb = phi_b - n_trace if phi_b > n_trace else 0
    return out, CompositionInfo(phi_b, qb, qb / n, k, max_deg, w)


SHA256_WIDTH, SHA256_CONSTRAINTS = 756, 766


def build_sha256_trace_multi(message):
    """build_sha256_trace_multi (sha256_air.rs:429-479), host side: returns
    (trace as a 756 x n_trace uint64 array, n_blocks)."""
    message = bytes(message)
    h = lib().stark_host_sha256_trace_height(len(message))
    tr = np.zeros((SHA256_WIDTH, h), dtype=np.uint64)
    dg = ctypes.create_string_buffer(32)
    nb = lib().stark_host_build_sha256_trace(message, len(message), tr.ctypes.data_as(u64p), dg)
    return tr, nb


def deep_ali_merge_sha256(trace_evals_on_lde, combination_coeffs, omega, n_trace, blowup, n_blocks,
                          phi_mode=PHI_RAYON, ctx=None):
    """deep_ali_merge_sha256 (lib.rs:453-545)."""
    n = n_trace * blowup
    lde = np.asarray(trace_evals_on_lde, dtype=np.uint64)
    assert lde.shape[0] == SHA256_WIDTH, "trace width mismatch"
    assert len(combination_coeffs) == SHA256_CONSTRAINTS, \
        "need one combination coefficient per constraint, got %d for %d constraints" % (
            len(combination_coeffs), SHA256_CONSTRAINTS)
    assert lde.shape[1] == n, "trace column length mismatch"
    del omega  # unused by the reference too (lib.rs:466)
    ctx = ctx or default_context()
    lde, lp = _u64(lde)
    co, cp = _u64(np.asarray(combination_coeffs, dtype=np.uint64))
    out = np.empty(n, dtype=np.uint64)
    ctx._check(lib().stark_merge_sha256(ctx.h, lp, n_trace, blowup, n_blocks, cp, SHA256_CONSTRAINTS, phi_mode,
                                        out.ctypes.data_as(u64p)))
    phi_b = 2 * n_trace
    qb = phi_b - n_trace
    return out, CompositionInfo(phi_b, qb, qb / n, SHA256_CONSTRAINTS, 2, SHA256_WIDTH)


SHA512_WIDTH, SHA512_CONSTRAINTS = 1510, 1526


def build_sha512_trace_multi(message):
    """build_sha512_trace_multi (sha512_air.rs:551-573), host side: returns
    (trace as a 1510 x n_trace uint64 array, n_blocks)."""
    message = bytes(message)
    h = lib().stark_host_sha512_trace_height(len(message))
    tr = np.zeros((SHA512_WIDTH, h), dtype=np.uint64)
    dg = ctypes.create_string_buffer(64)
    nb = lib().stark_host_build_sha512_trace(message, len(message), tr.ctypes.data_as(u64p), dg)
    return tr, nb


def deep_ali_merge_sha512(trace_evals_on_lde, combination_coeffs, omega, n_trace, blowup, n_blocks,
                          phi_mode=PHI_RAYON, ctx=None):
    """deep_ali_merge_sha512 (lib.rs:561-653)."""
    n = n_trace * blowup
    lde = np.asarray(trace_evals_on_lde, dtype=np.uint64)
    assert lde.shape[0] == SHA512_WIDTH, "trace width mismatch"
    assert len(combination_coeffs) == SHA512_CONSTRAINTS, \
        "need one combination coefficient per constraint, got %d for %d constraints" % (
            len(combination_coeffs), SHA512_CONSTRAINTS)
    assert lde.shape[1] == n, "trace column length mismatch"
    del omega
    ctx = ctx or default_context()
    lde, lp = _u64(lde)
    co, cp = _u64(np.asarray(combination_coeffs, dtype=np.uint64))
    out = np.empty(n, dtype=np.uint64)
    ctx._check(lib().stark_merge_sha512(ctx.h, lp, n_trace, blowup, n_blocks, cp, SHA512_CONSTRAINTS, phi_mode,
                                        out.ctypes.data_as(u64p)))
    return out, CompositionInfo(2 * n_trace, n_trace, n_trace / n, SHA512_CONSTRAINTS, 2, SHA512_WIDTH)


def merge_sha_dev(alg, d_lde, n_trace, blowup, n_blocks, combination_coeffs, d_out, phi_mode=PHI_RAYON, ctx=None):
    """deep_ali_merge_sha256 / _sha512 (alg 256 / 512) on device buffers
    (stark_merge_sha_dev): d_lde the width x n0 LDE, d_out receives n0 values."""
    ctx = ctx or default_context()
    co, cp = _u64(np.asarray(combination_coeffs, dtype=np.uint64))
    ctx._check(lib().stark_merge_sha_dev(ctx.h, int(alg), d_lde.ptr, n_trace, blowup, n_blocks, cp, len(co),
                                         phi_mode, d_out.ptr))


def merge_sha_from_trace_dev(alg, d_trace, n_trace, blowup, n_blocks, combination_coeffs, d_out, phi_mode=PHI_RAYON,
                             ctx=None):
    """The same c_eval as merge_sha_dev on the trace's LDE, computed from the
    width x n_trace trace (device) without that LDE -- the provers' route
    (stark_merge_sha_from_trace_dev). blowup >= 4."""
    ctx = ctx or default_context()
    co, cp = _u64(np.asarray(combination_coeffs, dtype=np.uint64))
    ctx._check(lib().stark_merge_sha_from_trace_dev(ctx.h, int(alg), d_trace.ptr, n_trace, blowup, n_blocks, cp,
                                                    len(co), phi_mode, d_out.ptr))


@dataclass
class ShaProof:
    digest: bytes
    root_f0: bytes
    n_blocks: int
    n_trace: int
    proof_blob: bytes


def prove_sha_batch(alg, messages, fs_bindings, heights=None, phi_mode=PHI_RAYON, ctx=None):
    """SHA-256 (alg=256) / SHA-512 (alg=512) AIR proofs with the
    prove_ds_ksk_binding pipeline (stark_prove_sha_batch); `alg` is one value
    or one per message; heights[i] is the trace height of message i (None:
    the minimum)."""
    ctx = ctx or default_context()
    msgs = [bytes(m) for m in messages]
    n = len(msgs)
    algs = [alg] * n if isinstance(alg, int) else list(alg)
    assert len(algs) == n
    assert len(fs_bindings) == n and (heights is None or len(heights) == n)
    offs, acc = [], 0
    for m in msgs:
        offs.append(acc)
        acc += len(m)
    outs = (_ShaOut * max(1, n))()
    hs = (sz * max(1, n))(*heights) if heights is not None else None
    ctx._check(lib().stark_prove_sha_batch(ctx.h, (ctypes.c_int * max(1, n))(*algs), b"".join(msgs),
                                           (sz * max(1, n))(*offs),
                                           (sz * max(1, n))(*[len(m) for m in msgs]), hs, n,
                                           b"".join(bytes(f) for f in fs_bindings), phi_mode, outs))
    res = []
    for k in range(n):
        o = outs[k]
        dl = 32 if algs[k] == 256 else 64
        res.append(ShaProof(bytes(o.digest)[:dl], bytes(o.root_f0), o.n_blocks, o.n_trace, _take(o.proof, o.proof_len)))
    return res


# ------------------------------------------------------------ RSA-2048 AIR
RSA_EXP_WIDTH, RSA_EXP_CONSTRAINTS, RSA_EXP_N_TRACE = 11041, 11280, 32


@dataclass
class RsaStackedRecord:
    """RsaStackedRecord (rsa2048_stacked_air.rs:58-62): modulus n, signature
    value s < n and the encoded message em = s^65537 mod n, as integers."""
    n: int
    s: int
    em: int


def _rsa_bytes(records):
    def be(x):
        return int(x).to_bytes(256, "big")
    return (b"".join(be(r.n) for r in records), b"".join(be(r.s) for r in records),
            b"".join(be(r.em) for r in records))


def fill_rsa_stacked(records, n_trace=RSA_EXP_N_TRACE):
    """fill_rsa_stacked (rsa2048_stacked_air.rs:64-78); one record is
    fill_rsa_exp_multirow (rsa2048_exp_air.rs:148-202). Returns the
    (11041 * len(records)) x n_trace trace."""
    records = list(records)
    tr = np.zeros((RSA_EXP_WIDTH * len(records), n_trace), dtype=np.uint64)
    nb, sb, eb = _rsa_bytes(records)
    rc = lib().stark_host_build_rsa_exp_trace(nb, sb, eb, len(records), n_trace, tr.ctypes.data_as(u64p))
    if rc != 0:
        raise StarkError(rc, "fill_rsa_stacked: need 0 < n, s < n and n_trace a power of two >= 32")
    return tr


def fill_rsa_exp_multirow(n, s, em, n_trace=RSA_EXP_N_TRACE):
    return fill_rsa_stacked([RsaStackedRecord(n, s, em)], n_trace)


def deep_ali_merge_rsa_stacked_streaming(trace_evals_on_lde, combination_coeffs, n_records, omega, n_trace, blowup,
                                         ctx=None):
    """deep_ali_merge_rsa_stacked_streaming (lib.rs:1949-2049); with one
    record it is deep_ali_merge_rsa_exp_multirow_streaming (lib.rs:1827-1944)."""
    n = n_trace * blowup
    k = RSA_EXP_CONSTRAINTS * n_records
    lde = np.asarray(trace_evals_on_lde, dtype=np.uint64)
    assert lde.shape == (RSA_EXP_WIDTH * n_records, n), "trace shape mismatch"
    assert len(combination_coeffs) == k, \
        "need one combination coefficient per constraint, got %d for %d constraints" % (len(combination_coeffs), k)
    del omega  # unused by the reference too (lib.rs:1838)
    ctx = ctx or default_context()
    lde, lp = _u64(lde)
    co, cp = _u64(np.asarray(combination_coeffs, dtype=np.uint64))
    out = np.empty(n, dtype=np.uint64)
    ctx._check(lib().stark_merge_rsa_exp(ctx.h, lp, n_trace, blowup, n_records, cp, k, out.ctypes.data_as(u64p)))
    return out, CompositionInfo(2 * n_trace, n_trace, n_trace / n, k, 2, RSA_EXP_WIDTH * n_records)


def deep_ali_merge_rsa_exp_multirow_streaming(trace_evals_on_lde, combination_coeffs, omega, n_trace, blowup,
                                              ctx=None):
    return deep_ali_merge_rsa_stacked_streaming(trace_evals_on_lde, combination_coeffs, 1, omega, n_trace, blowup,
                                                ctx)


def emsa_pkcs1_v1_5_encode_sha256(digest, k=256):
    """emsa_pkcs1_v1_5_encode_sha256 (rsa2048.rs) for k = 256."""
    assert k == 256 and len(digest) == 32
    em = ctypes.create_string_buffer(256)
    lib().stark_host_rsa_pkcs1_em_sha256(bytes(digest), em)
    return em.raw


def rsa_record_pk_hash(n_be, signature, message, em_bytes):
    """The per-record public-input hash of the mixed-algorithm zone example
    (mixed_algorithm_zone_rollup.rs:370-378): SHA3-256("RSA2048-MIXEDZONE"
    || n || signature || message || em)."""
    import hashlib
    h = hashlib.sha3_256()
    for part in (b"RSA2048-MIXEDZONE", n_be, signature, message, em_bytes):
        h.update(bytes(part))
    return h.digest()


@dataclass
class RsaProof:
    pi_hash: bytes
    root_f0: bytes
    n_trace: int
    proof_blob: bytes


def prove_rsa_exp_batch(records, public_inputs_hashes, records_per_proof=1, n_trace=RSA_EXP_N_TRACE, ctx=None):
    """RSA-2048 exponentiation proofs (stark_prove_rsa_exp_batch): the
    prove_record_rsa pipeline (mixed_algorithm_zone_rollup.rs:349-395) from
    the trace on, for len(records) / records_per_proof proofs of that many
    stacked records each; pi_hash = SHA3-256(public_inputs_hash || root_f0)."""
    ctx = ctx or default_context()
    records = list(records)
    assert records and len(records) % records_per_proof == 0
    n = len(records) // records_per_proof
    assert len(public_inputs_hashes) == n
    nb, sb, eb = _rsa_bytes(records)
    outs = (_RsaOut * n)()
    ctx._check(lib().stark_prove_rsa_exp_batch(ctx.h, nb, sb, eb, n, records_per_proof, n_trace,
                                               b"".join(bytes(x) for x in public_inputs_hashes), outs))
    return [RsaProof(bytes(o.pi_hash), bytes(o.root_f0), o.n_trace, _take(o.proof, o.proof_len)) for o in outs]


# ------------------------------------------------------ ECDSA-P256 AIR
P256_WIDTH, P256_CONSTRAINTS, P256_ECDSA_N_TRACE = 151334, 155774, 256


@dataclass
class P256Chains:
    """The inputs of fill_ecdsa_double_multirow
    (p256_ecdsa_double_multirow_air.rs:134-222), as integers below p: for
    chain A then chain B the initial accumulator and the base point
    (projective x, y, z), the scalar bits (0/1, one per row), and the six
    r_proj values; bind_outputs takes r_proj from the chain outputs at the
    chains' last bits (the callers' two-pass fill)."""
    a_initial: tuple
    a_base: tuple
    b_initial: tuple
    b_base: tuple
    a_bits: list
    b_bits: list
    rproj: tuple = (0, 0, 0, 0, 0, 0)
    bind_outputs: bool = True


def _p256_chains(c):
    x = _P256Chains()
    pts = list(c.a_initial) + list(c.a_base) + list(c.b_initial) + list(c.b_base)
    assert len(pts) == 12 and len(c.rproj) == 6
    for i, v in enumerate(pts):
        x.pts[i][:] = int(v).to_bytes(32, "big")
    for i, v in enumerate(c.rproj):
        x.rproj[i][:] = int(v).to_bytes(32, "big")
    a = (ctypes.c_uint8 * max(1, len(c.a_bits)))(*[int(b) for b in c.a_bits])
    b = (ctypes.c_uint8 * max(1, len(c.b_bits)))(*[int(v) for v in c.b_bits])
    x.a_bits, x.k_a = ctypes.cast(a, u8p), len(c.a_bits)
    x.b_bits, x.k_b = ctypes.cast(b, u8p), len(c.b_bits)
    x.bind_outputs = int(bool(c.bind_outputs))
    x._keep = (a, b)
    return x


def fill_ecdsa_double_multirow(chains: P256Chains, n_trace):
    """fill_ecdsa_double_multirow (p256_ecdsa_double_multirow_air.rs:134-222):
    the 151334 x n_trace trace and the r_proj values it binds."""
    x = _p256_chains(chains)
    tr = np.zeros((P256_WIDTH, n_trace), dtype=np.uint64)
    rc = lib().stark_host_build_ecdsa_p256_trace(ctypes.byref(x), n_trace, tr.ctypes.data_as(u64p))
    if rc != 0:
        raise StarkError(rc, "fill_ecdsa_double_multirow: need values below p, n_trace a power of two >= the bits")
    return tr, tuple(int.from_bytes(bytes(x.rproj[i]), "big") for i in range(6))


def deep_ali_merge_ecdsa_double_multirow_streaming(trace_evals_on_lde, combination_coeffs, omega, n_trace, blowup,
                                                   ctx=None):
    """deep_ali_merge_ecdsa_double_multirow_streaming (lib.rs:1421-1540)."""
    n = n_trace * blowup
    lde = np.asarray(trace_evals_on_lde, dtype=np.uint64)
    assert lde.shape == (P256_WIDTH, n), "trace shape mismatch"
    assert len(combination_coeffs) == P256_CONSTRAINTS, "need one combination coefficient per constraint"
    del omega  # unused by the reference too (lib.rs:1430)
    ctx = ctx or default_context()
    lde, lp = _u64(lde)
    co, cp = _u64(np.asarray(combination_coeffs, dtype=np.uint64))
    out = np.empty(n, dtype=np.uint64)
    ctx._check(lib().stark_merge_ecdsa_p256(ctx.h, lp, n_trace, blowup, cp, P256_CONSTRAINTS,
                                            out.ctypes.data_as(u64p)))
    return out, CompositionInfo(2 * n_trace, n_trace, n_trace / n, P256_CONSTRAINTS, 2, P256_WIDTH)


@dataclass
class P256Proof:
    pi_hash: bytes
    root_f0: bytes
    rproj: tuple
    n_trace: int
    proof_blob: bytes


def _p256_outs(outs):
    return [P256Proof(bytes(o.pi_hash), bytes(o.root_f0),
                      tuple(int.from_bytes(bytes(o.rproj[i]), "big") for i in range(6)), o.n_trace,
                      _take(o.proof, o.proof_len)) for o in outs]


def prove_p256_chains_batch(chains, n_trace, public_inputs_hashes, ctx=None):
    """Double-chain proofs (stark_prove_p256_chains_batch): the
    run_double_chain_streaming pipeline (p256_full_ecdsa_stark_bench.rs:
    133-240) per chain input; pi_hash = SHA3-256(public_inputs_hash || root_f0)."""
    ctx = ctx or default_context()
    chains = list(chains)
    assert chains and len(public_inputs_hashes) == len(chains)
    xs = [_p256_chains(c) for c in chains]
    arr = (_P256Chains * len(xs))(*xs)
    outs = (_P256Out * len(xs))()
    ctx._check(lib().stark_prove_p256_chains_batch(ctx.h, arr, len(xs), n_trace,
                                                   b"".join(bytes(x) for x in public_inputs_hashes), outs))
    return _p256_outs(outs)


@dataclass
class EcdsaRecord:
    """An ECDSA-P256 signature record: public key (qx, qy), signature (r, s)
    as integers, and the 32-byte SHA-256 digest of the signed message."""
    qx: int
    qy: int
    r: int
    s: int
    digest: bytes


def ecdsa_record_pk_hash(rec: EcdsaRecord):
    """pk_hash_pre of prove_record_ecdsa (mixed_algorithm_zone_rollup.rs:
    293-302): SHA3-256("ECDSAP256-MIXEDZONE" || qx || qy || r || s || digest)."""
    import hashlib
    h = hashlib.sha3_256(b"ECDSAP256-MIXEDZONE")
    for v in (rec.qx, rec.qy, rec.r, rec.s):
        h.update(int(v).to_bytes(32, "big"))
    h.update(bytes(rec.digest))
    return h.digest()


def prove_record_ecdsa_batch(records, ctx=None):
    """prove_record_ecdsa (mixed_algorithm_zone_rollup.rs:193-322) from the
    signature on, one proof per record (stark_prove_ecdsa_p256_batch):
    raises StarkError for a key off the curve or a signature that does not
    verify, where the example asserts."""
    ctx = ctx or default_context()
    records = list(records)
    assert records

    def cat(f):
        return b"".join(f(r) for r in records)
    outs = (_P256Out * len(records))()
    ctx._check(lib().stark_prove_ecdsa_p256_batch(
        ctx.h, cat(lambda r: int(r.qx).to_bytes(32, "big")), cat(lambda r: int(r.qy).to_bytes(32, "big")),
        cat(lambda r: int(r.r).to_bytes(32, "big")), cat(lambda r: int(r.s).to_bytes(32, "big")),
        cat(lambda r: bytes(r.digest)), len(records), outs))
    return _p256_outs(outs)


# --------------------------------------------------------------------- FRI
@dataclass
class FriDomain:
    """FriDomain::new_radix2 (fri.rs:160-175)."""
    omega: int
    size: int

    @classmethod
    def new_radix2(cls, size):
        return cls(pow(7, (P - 1) // size, P), size)


@dataclass
class DeepFriParams:
    """DeepFriParams (fri.rs:1303-1349)."""
    schedule: List[int]
    r: int
    seed_z: int
    coeff_commit_final: bool = False
    d_final: int = 1
    stir: bool = False
    s0: Optional[int] = None
    public_inputs_hash: Optional[bytes] = None

    def _c(self):
        arr = (ctypes.c_size_t * max(1, len(self.schedule)))(*self.schedule)
        p = _FriParams(arr, len(self.schedule), self.r, self.s0 if self.s0 is not None else self.r, self.seed_z,
                       int(self.coeff_commit_final), self.d_final, int(self.stir), self.public_inputs_hash)
        p._keep = arr
        return p


def deep_fri_prove(f0, domain0, params, ctx=None):
    """deep_fri_prove::<SexticExt> (fri.rs:1900-2055). Returns the proof's
    ark-serialize compressed bytes (what prover.rs serializes and ships)."""
    ctx = ctx or default_context()
    f0, fp = _u64(f0)
    assert f0.size == domain0.size
    p = params._c()
    pp, pl = u8p(), sz()
    ctx._check(lib().stark_deep_fri_prove(ctx.h, fp, f0.size, ctypes.byref(p), ctypes.byref(pp), ctypes.byref(pl)))
    return _take(pp, pl.value)


def merkle_commit(leaf_values, arity, tree_label, ctx=None):
    """MerkleTreeChannel push_leaves_parallel + finalize (merkle lib.rs:391-476) with
    layer_arities = [arity; depth]; leaf_values is (n_leaves, vals_per_leaf)."""
    ctx = ctx or default_context()
    v = np.ascontiguousarray(np.atleast_2d(leaf_values), dtype=np.uint64)
    root = ctypes.create_string_buffer(32)
    ctx._check(lib().stark_merkle_commit(ctx.h, v.ctypes.data_as(u64p), v.shape[0], v.shape[1], arity, tree_label,
                                         root))
    return root.raw


def merkle_depth(leaves, arity):
    """merkle_depth (fri.rs:950-959)."""
    d, cur = 1, leaves
    while cur > arity:
        cur = (cur + arity - 1) // arity
        d += 1
    return d


@dataclass
class MerkleChannelCfg:
    """MerkleChannelCfg (merkle lib.rs:117-128) with constant layer arity."""
    layer_arities: List[int]
    tree_label: int

    @classmethod
    def constant(cls, n_leaves, arity, tree_label):
        return cls([arity] * merkle_depth(n_leaves, arity), tree_label)

    def _arity(self):
        assert self.layer_arities and all(a == self.layer_arities[0] for a in self.layer_arities), \
            "this library builds constant-arity trees"
        return self.layer_arities[0]


def merkle_open(leaf_values, arity, tree_label, indices, ctx=None):
    """MerkleTreeChannel push_leaves_parallel + finalize + open (merkle lib.rs:391-512): the root and, per
    index, the serialized MerkleOpening (leaf digest, sibling path, index)."""
    ctx = ctx or default_context()
    v = np.ascontiguousarray(np.atleast_2d(leaf_values), dtype=np.uint64)
    idx = (sz * max(1, len(indices)))(*indices)
    root = ctypes.create_string_buffer(32)
    pp, pl = u8p(), sz()
    offs = (sz * (len(indices) + 1))()
    ctx._check(lib().stark_merkle_open(ctx.h, v.ctypes.data_as(u64p), v.shape[0], v.shape[1], arity, tree_label,
                                       idx, len(indices), root, ctypes.byref(pp), ctypes.byref(pl), offs))
    blob = _take(pp, pl.value)
    return root.raw, [blob[offs[q]:offs[q + 1]] for q in range(len(indices))]


def verify_opening(cfg, root, opening):
    """MerkleTreeChannel::verify_opening (merkle lib.rs:754-799) on a serialized MerkleOpening."""
    rc = lib().stark_merkle_verify_opening(cfg._arity(), len(cfg.layer_arities), cfg.tree_label, bytes(root),
                                           bytes(opening), len(opening))
    if rc < 0:
        raise StarkError(rc, "merkle opening does not deserialize")
    return rc == 1


def compute_leaf_hash(cfg, index, values):
    """compute_leaf_hash (merkle lib.rs:806-812)."""
    v = np.ascontiguousarray(values, dtype=np.uint64)
    out = ctypes.create_string_buffer(32)
    lib().stark_host_merkle_leaf_hash(cfg.layer_arities[0], cfg.tree_label, index, v.ctypes.data_as(u64p), v.size,
                                      out)
    return out.raw


def deep_fri_verify(params, proof):
    """deep_fri_verify::<SexticExt> (fri.rs:2136-2710) on the ark-serialize compressed proof: True / False.
    Bytes that do not deserialize raise StarkError (the reference's deserialize_with_mode error)."""
    p = params._c()
    rc = lib().stark_deep_fri_verify(ctypes.byref(p), bytes(proof), len(proof))
    if rc < 0:
        raise StarkError(rc, "proof bytes do not deserialize as DeepFriProof<SexticExt>")
    return rc == 1


def deep_fri_verify_batch(params, proofs, public_inputs_hashes=None):
    """deep_fri_verify over many proofs on the library's host threads; a list of True / False / None (None:
    the bytes do not deserialize). public_inputs_hashes binds proof i to its own 32-byte hash."""
    p = params._c()
    n = len(proofs)
    bufs = [bytes(x) for x in proofs]
    arr = (ctypes.c_char_p * max(1, n))(*bufs)
    lens = (sz * max(1, n))(*[len(x) for x in bufs])
    res = (ctypes.c_int * max(1, n))()
    pis = b"".join(public_inputs_hashes) if public_inputs_hashes is not None else None
    if pis is not None:
        assert len(pis) == 32 * n
    rc = lib().stark_deep_fri_verify_batch(ctypes.byref(p), arr, lens, n, pis, res)
    if rc < 0:
        raise StarkError(rc, "deep_fri_verify_batch failed")
    return [True if res[i] == 1 else (False if res[i] == 0 else None) for i in range(n)]


# --------------------------------------------------------------------- swarm-dns
class LdtMode(enum.Enum):
    Stir = _LDT_STIR
    Fri = _LDT_FRI

    def is_stir(self):
        return self is LdtMode.Stir


def make_schedule(n0, ldt=LdtMode.Stir):
    """make_schedule (prover.rs:53-67)."""
    assert n0 > 0 and n0 & (n0 - 1) == 0, "n0 must be a power of 2"
    lg = n0.bit_length() - 1
    if not ldt.is_stir():
        return [2] * lg
    s = [8] * (lg // 3)
    if lg % 3:
        s.append(1 << (lg % 3))
    return s


def build_params(n0, pk_hash_32, ldt=LdtMode.Stir):
    """build_params (prover.rs:87-98)."""
    return DeepFriParams(make_schedule(n0, ldt), NUM_QUERIES, SEED_Z, True, 1, ldt.is_stir(), NUM_QUERIES,
                         bytes(pk_hash_32))


@dataclass
class DnsRecord:
    """DnsRecord (dns.rs:24-31) with the constructors of dns.rs:33-41."""
    domain: str
    record_type: int
    ttl: int
    rdata: bytes

    @classmethod
    def a(cls, d, t, ip):
        return cls(d, 1, t, bytes(ip))

    @classmethod
    def aaaa(cls, d, t, ip):
        return cls(d, 28, t, bytes(ip))

    @classmethod
    def txt(cls, d, t, s):
        return cls(d, 16, t, s.encode())

    @classmethod
    def mx(cls, d, t, prio, ex):
        return cls(d, 15, t, prio.to_bytes(2, "big") + ex.encode())

    def _c(self):
        d = self.domain.encode()
        return _DnsRecord(d, len(d), self.record_type, self.ttl, self.rdata, len(self.rdata))

    def leaf_hash(self, salt):
        out = ctypes.create_string_buffer(32)
        r = self._c()
        lib().stark_host_dns_leaf_hash(ctypes.byref(r), bytes(salt), out)
        return out.raw


@dataclass
class InnerShardOutput:
    pi_hash: bytes
    merkle_root: bytes
    record_count: int
    n_trace: int
    proof_bytes: int
    root_f0: bytes
    proof_blob: bytes


@dataclass
class OuterRollupOutput:
    n_trace: int
    proof_bytes: int
    root_f0: bytes
    proof_blob: bytes
    n0: int


def prove_inner_shard(salt, records: Sequence[DnsRecord], fs_binding_32, ldt=LdtMode.Stir, phi_mode=PHI_RAYON,
                      ctx=None):
    """prove_inner_shard (prover.rs:148-227)."""
    ctx = ctx or default_context()
    arr = (_DnsRecord * max(1, len(records)))(*[r._c() for r in records])
    out = _InnerOut()
    ctx._check(lib().stark_prove_inner_shard(ctx.h, bytes(salt), arr, len(records), bytes(fs_binding_32),
                                             ldt.value, phi_mode, ctypes.byref(out)))
    blob = _take(out.proof, out.proof_len)
    return InnerShardOutput(bytes(out.pi_hash), bytes(out.merkle_root), out.record_count, out.n_trace, len(blob),
                            bytes(out.root_f0), blob)


@dataclass
class Nsec3Record:
    """Nsec3Record (prover.rs:321-325)"""
    owner_hash: bytes
    next_hash: bytes


@dataclass
class Nsec3Output:
    """Nsec3Output (prover.rs:327-338); prove_ms / local_verify_ms are not kept."""
    pi_hash: bytes
    chain_root: bytes
    record_count: int
    n_trace: int
    proof_bytes: int
    root_f0: bytes
    proof_blob: bytes


def _nsec3_out(o):
    blob = _take(o.proof, o.proof_len)
    return Nsec3Output(bytes(o.pi_hash), bytes(o.chain_root), o.record_count, o.n_trace, len(blob), bytes(o.root_f0),
                       blob)


def _nsec3_bytes(records):
    out = []
    for r in records:
        assert len(r.owner_hash) == 32 and len(r.next_hash) == 32
        out.append(bytes(r.owner_hash) + bytes(r.next_hash))
    return b"".join(out)


def prove_nsec3_completeness(records: Sequence[Nsec3Record], salt, fs_binding_32, ldt=LdtMode.Stir,
                             phi_mode=PHI_RAYON, ctx=None):
    """prove_nsec3_completeness (prover.rs:348-446); an empty or open chain
    raises StarkError where the reference panics."""
    ctx = ctx or default_context()
    out = _Nsec3Out()
    ctx._check(lib().stark_prove_nsec3_completeness(ctx.h, _nsec3_bytes(records), len(records), bytes(salt),
                                                    bytes(fs_binding_32), ldt.value, phi_mode, ctypes.byref(out)))
    return _nsec3_out(out)


def prove_nsec3_completeness_batch(chains, salt, fs_bindings, ldt=LdtMode.Stir, phi_mode=PHI_RAYON, ctx=None):
    """prove_nsec3_completeness for several chains in lockstep device batches
    (stark_prove_nsec3_completeness_batch); result i equals the single call."""
    ctx = ctx or default_context()
    chains = [list(c) for c in chains]
    n = len(chains)
    assert len(fs_bindings) == n
    outs = (_Nsec3Out * max(1, n))()
    ctx._check(lib().stark_prove_nsec3_completeness_batch(
        ctx.h, b"".join(_nsec3_bytes(c) for c in chains), (sz * max(1, n))(*[len(c) for c in chains]), n, bytes(salt),
        b"".join(bytes(f) for f in fs_bindings), ldt.value, phi_mode, outs))
    return [_nsec3_out(outs[i]) for i in range(n)]


@dataclass
class DsKskOutput:
    """DsKskOutput (prover.rs:464-490); prove_ms / local_verify_ms are not kept."""
    pi_hash: bytes
    asserted_digest: bytes
    n_blocks: int
    n_trace: int
    proof_bytes: int
    root_f0: bytes
    proof_blob: bytes


def _ds_ksk_out(o):
    blob = _take(o.proof, o.proof_len)
    return DsKskOutput(bytes(o.pi_hash), bytes(o.asserted_digest), o.n_blocks, o.n_trace, len(blob),
                       bytes(o.root_f0), blob)


def prove_ds_ksk_binding(dnskey_bytes, parent_ds_hash, fs_binding_32, ldt=LdtMode.Stir, phi_mode=PHI_RAYON,
                         ctx=None):
    """prove_ds_ksk_binding (prover.rs:512-585)."""
    ctx = ctx or default_context()
    out = _DsKskOut()
    key = bytes(dnskey_bytes)
    ctx._check(lib().stark_prove_ds_ksk_binding(ctx.h, key, len(key), bytes(parent_ds_hash), bytes(fs_binding_32),
                                                ldt.value, phi_mode, ctypes.byref(out)))
    return _ds_ksk_out(out)


def prove_ds_ksk_binding_batch(dnskeys, parent_ds_hashes, fs_bindings, ldt=LdtMode.Stir, phi_mode=PHI_RAYON,
                               ctx=None):
    """Many prove_ds_ksk_binding calls as lockstep device batches
    (stark_prove_ds_ksk_binding_batch); result k equals the single call on key k."""
    ctx = ctx or default_context()
    keys = [bytes(k) for k in dnskeys]
    n = len(keys)
    assert len(parent_ds_hashes) == n and len(fs_bindings) == n
    offs, acc = [], 0
    for k in keys:
        offs.append(acc)
        acc += len(k)
    outs = (_DsKskOut * max(1, n))()
    ctx._check(lib().stark_prove_ds_ksk_binding_batch(
        ctx.h, b"".join(keys), (sz * max(1, n))(*offs), (sz * max(1, n))(*[len(k) for k in keys]), n,
        b"".join(bytes(p) for p in parent_ds_hashes), b"".join(bytes(f) for f in fs_bindings), ldt.value, phi_mode,
        outs))
    return [_ds_ksk_out(outs[k]) for k in range(n)]


class PackedShards:
    """Records of many shards packed once into the C record array that
    stark_prove_inner_shard_batch reads; pass it in place of the list of
    shards to prove the same records again without re-marshaling them."""

    def __init__(self, shards):
        flat = [r._c() for s in shards for r in s]
        self.arr = (_DnsRecord * max(1, len(flat)))(*flat)
        self.counts = [len(s) for s in shards]
        self.sizes = (ctypes.c_size_t * max(1, len(shards)))(*self.counts)

    def __len__(self):
        return len(self.counts)


def prove_inner_shard_batch(salt, shards, fs_bindings, ldt=LdtMode.Stir, phi_mode=PHI_RAYON, ctx=None):
    """Many prove_inner_shard calls as one lockstep device batch
    (stark_prove_inner_shard_batch); result k equals prove_inner_shard on shard k.
    `shards` is a list of record lists or a PackedShards."""
    ctx = ctx or default_context()
    packed = shards if isinstance(shards, PackedShards) else PackedShards(shards)
    assert len(packed) == len(fs_bindings)
    arr, sizes = packed.arr, packed.sizes
    outs = (_InnerOut * max(1, len(shards)))()
    fsb = b"".join(bytes(f) for f in fs_bindings)
    ctx._check(lib().stark_prove_inner_shard_batch(ctx.h, bytes(salt), arr, sizes, len(shards), fsb, ldt.value,
                                                   phi_mode, outs))
    res = []
    for k in range(len(shards)):
        o = outs[k]
        blob = _take(o.proof, o.proof_len)
        res.append(InnerShardOutput(bytes(o.pi_hash), bytes(o.merkle_root), o.record_count, o.n_trace, len(blob),
                                    bytes(o.root_f0), blob))
    return res


def deep_fri_prove_batch(f0s, domain0, params, public_inputs_hashes=None, ctx=None):
    """deep_fri_prove over many f0 vectors of one domain in lockstep; result k
    equals deep_fri_prove(f0s[k]) with public_inputs_hash = public_inputs_hashes[k]."""
    ctx = ctx or default_context()
    f0s = np.ascontiguousarray(f0s, dtype=np.uint64)
    K, n0 = f0s.shape
    assert n0 == domain0.size
    p = params._c()
    pp = (u8p * K)()
    pl = (sz * K)()
    pis = b"".join(bytes(h) for h in public_inputs_hashes) if public_inputs_hashes is not None else None
    ctx._check(lib().stark_deep_fri_prove_batch(ctx.h, f0s.ctypes.data_as(u64p), K, n0, ctypes.byref(p), pis, pp, pl))
    return [_take(pp[k], pl[k]) for k in range(K)]


def prove_outer_rollup(pi_hashes, pk_hash_32, ldt=LdtMode.Stir, phi_mode=PHI_RAYON, ctx=None):
    """prove_outer_rollup (prover.rs:253-304)."""
    ctx = ctx or default_context()
    flat = b"".join(bytes(h) for h in pi_hashes)
    pp, pl = u8p(), sz()
    rf = ctypes.create_string_buffer(32)
    ctx._check(lib().stark_prove_outer_rollup(ctx.h, flat, len(pi_hashes), bytes(pk_hash_32), ldt.value, phi_mode,
                                              ctypes.byref(pp), ctypes.byref(pl), rf))
    blob = _take(pp, pl.value)
    n_trace = 8
    while n_trace < 4 * len(pi_hashes):
        n_trace *= 2
    return OuterRollupOutput(n_trace, len(blob), rf.raw, blob, n_trace * BLOWUP)


def shard_fs_binding(authority_pk_hash, job_id, shard_id, shard_nonce):
    """shard_fs_binding (crates/swarm-proto/src/messages.rs:162-176); host SHA3."""
    return hashlib.sha3_256(b"DNS-SWARM-SHARD-FS-V1" + bytes(authority_pk_hash) + bytes(job_id)
                            + int(shard_id).to_bytes(4, "little") + bytes(shard_nonce)).digest()


# --------------------------------------------------------------------- prover core
def prover_core(d_trace, w, n_trace, blowup, shift, params, d_lde, ctx=None):
    """Device-resident prover core: coset LDE -> row tree -> STIR prove -> row
    openings (stark_prover_core_dev). d_trace / d_lde are DeviceBuffers.
    Returns (row_root, proof bytes)."""
    ctx = ctx or default_context()
    p = params._c()
    root = ctypes.create_string_buffer(32)
    pp, pl = u8p(), sz()
    ctx._check(lib().stark_prover_core_dev(ctx.h, d_trace.ptr, w, n_trace, blowup, shift, ctypes.byref(p), d_lde.ptr,
                                           root, ctypes.byref(pp), ctypes.byref(pl)))
    return root.raw, _take(pp, pl.value)


# ---------------------------------------------------------------- Ed25519 (ed25519_verify_air.rs v16)

ED25519_K_SCALAR = 256  # prove_zsk_ksk_binding_v2's production k_scalar (prover.rs:835-840)


@dataclass
class VerifyAirShapeV16:
    """Sizes of verify_air_layout_v16 (ed25519_verify_air.rs:4741-4818)."""
    width: int
    height: int
    n_constraints: int


def verify_air_layout_v16(msg_len_for_sha512, k_scalar=ED25519_K_SCALAR):
    """Trace width / height and verify_v16_per_row_constraints for a SHA-512
    input (R || A || M) of msg_len_for_sha512 bytes."""
    assert msg_len_for_sha512 >= 64, "the SHA-512 input is R || A || M"
    w, h, k = sz(), sz(), sz()
    rc = lib().stark_ed25519_shape(msg_len_for_sha512 - 64, k_scalar, ctypes.byref(w), ctypes.byref(h), ctypes.byref(k))
    if rc != 0:
        raise StarkError(rc, "verify_air_layout_v16: k_scalar must be in [2, 256]")
    return VerifyAirShapeV16(w.value, h.value, k.value)


def ed25519_verify(pubkey, signature, message):
    """ed25519_verify::verify (ed25519_verify.rs:54-118): cofactored, s < L."""
    return bool(lib().stark_host_ed25519_verify(bytes(pubkey), bytes(signature), bytes(message), len(message)))


def ed25519_sign(seed, message):
    """RFC 8032 signing from a 32-byte seed: (public key, signature)."""
    pk, sig = ctypes.create_string_buffer(32), ctypes.create_string_buffer(64)
    lib().stark_host_ed25519_sign(bytes(seed), bytes(message), len(message), pk, sig)
    return pk.raw, sig.raw


def fill_verify_air_v16(pubkey, signature, message, k_scalar=ED25519_K_SCALAR, device=False, ctx=None):
    """fill_verify_air_v16 (ed25519_verify_air.rs:4820-4865) for the inputs
    prove_zsk_ksk_binding_v2 derives from (pubkey, signature, message): the
    width x height trace, built on the host or (device=True) on the GPU."""
    shape = verify_air_layout_v16(64 + len(message), k_scalar)
    tr = np.zeros((shape.width, shape.height), dtype=np.uint64)
    args = (bytes(pubkey), bytes(signature), bytes(message), len(message), k_scalar, tr.ctypes.data_as(u64p))
    if device:
        ctx = ctx or default_context()
        ctx._check(lib().stark_build_ed25519_trace(ctx.h, *args))
    else:
        rc = lib().stark_host_build_ed25519_trace(*args)
        if rc != 0:
            raise StarkError(rc, "fill_verify_air_v16: R or A does not decompress")
    return tr


def deep_ali_merge_ed25519_verify_streaming(trace_evals_on_lde, combination_coeffs, pubkey, signature, message,
                                            omega, n_trace, blowup, k_scalar=ED25519_K_SCALAR, ctx=None):
    """deep_ali_merge_ed25519_verify_streaming (lib.rs:854-974); the layout
    argument of the reference is given by the signature it was built for."""
    shape = verify_air_layout_v16(64 + len(message), k_scalar)
    n = n_trace * blowup
    assert n_trace == shape.height
    lde = np.asarray(trace_evals_on_lde, dtype=np.uint64)
    assert lde.shape == (shape.width, n), "trace shape mismatch"
    assert len(combination_coeffs) == shape.n_constraints, "need one combination coefficient per constraint"
    del omega  # unused by the reference too (lib.rs:866)
    ctx = ctx or default_context()
    lde, lp = _u64(lde)
    co, cp = _u64(np.asarray(combination_coeffs, dtype=np.uint64))
    out = np.empty(n, dtype=np.uint64)
    ctx._check(lib().stark_merge_ed25519(ctx.h, bytes(pubkey), bytes(signature), bytes(message), len(message),
                                         k_scalar, lp, blowup, cp, shape.n_constraints, out.ctypes.data_as(u64p)))
    return out, CompositionInfo(2 * n_trace, n_trace, n_trace / n, shape.n_constraints, 2, shape.width)


def ed25519_record_pk_hash(pubkey, signature, message):
    """pk_hash_pre of prove_record_ed25519 (mixed_algorithm_zone_rollup.rs:
    155-162): SHA3-256("ED25519-MIXEDZONE" || A || signature || message). That
    record proof is prove_zsk_ksk_binding_v2's proof with this hash as the
    FRI public-input hash (k_scalar 256, blowup 32, 54 queries)."""
    import hashlib
    h = hashlib.sha3_256(b"ED25519-MIXEDZONE")
    for part in (pubkey, signature, message):
        h.update(bytes(part))
    return h.digest()


def record_pi_hash(pk_hash_pre, root_f0):
    """Per-record pi_hash of the mixed-algorithm zone
    (mixed_algorithm_zone_rollup.rs:176-182): SHA3-256(pk_hash_pre || root_f0)."""
    import hashlib
    return hashlib.sha3_256(bytes(pk_hash_pre) + bytes(root_f0)).digest()


@dataclass
class ZskKskOutput:
    """ZskKskOutput of prove_zsk_ksk_binding_v2 (prover.rs:856-951)."""
    pi_hash: bytes
    verified: bool
    proof_blob: bytes
    proof_bytes: int
    root_f0: bytes
    n_trace: int


def prove_zsk_ksk_binding_v2_batch(requests, k_scalar=ED25519_K_SCALAR, blowup=BLOWUP, queries=NUM_QUERIES,
                                   ldt=LdtMode.Stir, ctx=None):
    """prove_zsk_ksk_binding_v2 for each (ksk_pubkey, signature, signed_data,
    fs_binding_32, merkle_root_32) request in one batched call; raises
    StarkError when a signature does not verify (the reference asserts)."""
    ctx = ctx or default_context()
    reqs = list(requests)
    assert reqs
    msgs = [bytes(r[2]) for r in reqs]
    off, o = [], 0
    for m in msgs:
        off.append(o)
        o += len(m)
    n = len(reqs)
    outs = (_ZskOut * n)()
    ctx._check(lib().stark_prove_zsk_ksk_binding_v2_batch(
        ctx.h, b"".join(bytes(r[0]) for r in reqs), b"".join(bytes(r[1]) for r in reqs), b"".join(msgs) or b"\0",
        (sz * n)(*off), (sz * n)(*[len(m) for m in msgs]), b"".join(bytes(r[3]) for r in reqs),
        b"".join(bytes(r[4]) for r in reqs), n, k_scalar, blowup, queries, ldt.value, outs))
    res = []
    for x in outs:
        blob = _take(x.proof, x.proof_len)
        res.append(ZskKskOutput(bytes(x.pi_hash), True, blob, len(blob), bytes(x.root_f0), x.n_trace))
    return res


def prove_zsk_ksk_binding_v2(ksk_pubkey, signature, signed_data, fs_binding_32, merkle_root_32,
                             k_scalar=ED25519_K_SCALAR, ldt=LdtMode.Stir, ctx=None):
    """prove_zsk_ksk_binding_v2 (prover.rs:856-951) for one signature."""
    return prove_zsk_ksk_binding_v2_batch([(ksk_pubkey, signature, signed_data, fs_binding_32, merkle_root_32)],
                                          k_scalar, ldt=ldt, ctx=ctx)[0]
