"""ProvingServer -- the device side of a swarm worker's request loop
(crates/swarm-worker/src/main.rs:220-335) on one GPU: requests are admitted
by an HBM budget and proven with continuous batching (stark_server_* in
include/stark_gpu.h). Each result equals the single-call result of
prove_ds_ksk_binding / prove_inner_shard / the RSA-2048 or ECDSA-P256 record
prover for the same request."""
import ctypes
from dataclasses import dataclass

from . import api as A

SERVE_DS_KSK, SERVE_INNER_SHARD, SERVE_RSA, SERVE_P256 = 0, 1, 2, 3
DRY_RUN = 1
E_TIMEOUT = -6


class _Req(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("dnskey", ctypes.c_char_p), ("dnskey_len", A.sz),
                ("parent_ds_hash", ctypes.c_char_p), ("salt", ctypes.c_char_p),
                ("records", ctypes.POINTER(A._DnsRecord)), ("n_records", A.sz), ("fs_binding", ctypes.c_char_p),
                ("rsa_n", ctypes.c_char_p), ("rsa_s", ctypes.c_char_p), ("rsa_em", ctypes.c_char_p),
                ("p256_qx", ctypes.c_char_p), ("p256_qy", ctypes.c_char_p), ("p256_r", ctypes.c_char_p),
                ("p256_s", ctypes.c_char_p), ("p256_digest", ctypes.c_char_p)]


class _Res(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int), ("kind", ctypes.c_int),
                ("pi_hash", ctypes.c_uint8 * 32), ("root_f0", ctypes.c_uint8 * 32), ("aux", ctypes.c_uint8 * 32),
                ("n_trace", A.sz), ("batch_size", A.sz), ("batch_seq", ctypes.c_uint64), ("proof", A.u8p),
                ("proof_len", A.sz),
                ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_done", ctypes.c_double)]


@dataclass
class ServeResult:
    id: int
    status: int
    kind: int
    pi_hash: bytes
    root_f0: bytes
    aux: bytes  # merkle_root (inner shard), asserted_digest (DS-KSK), zero (RSA, P-256)
    n_trace: int
    batch_size: int
    batch_seq: int
    proof_blob: bytes
    t_submit: float
    t_start: float
    t_done: float

    @property
    def latency(self):
        return self.t_done - self.t_submit


_bound = False


def _bind():
    global _bound
    if _bound:
        return A.lib()
    L = A.lib()
    vp = ctypes.c_void_p
    for name, res, args in [
        ("stark_server_create", ctypes.c_int, [ctypes.c_int, ctypes.c_int, A.sz, A.sz, ctypes.c_int,
                                               ctypes.POINTER(vp)]),
        ("stark_server_destroy", None, [vp]),
        ("stark_server_submit", ctypes.c_int, [vp, ctypes.POINTER(_Req), ctypes.POINTER(ctypes.c_uint64)]),
        ("stark_server_wait", ctypes.c_int, [vp, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(_Res)]),
        ("stark_server_stats", ctypes.c_int, [vp, ctypes.POINTER(A.sz), ctypes.POINTER(A.sz), ctypes.POINTER(A.sz)]),
        ("stark_server_last_error", ctypes.c_char_p, [vp]),
    ]:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _bound = True
    return L


class ProvingServer:
    def __init__(self, device=0, workers=4, hbm_budget=64 << 30, max_batch=64, dry_run=False):
        L = _bind()
        h = ctypes.c_void_p()
        rc = L.stark_server_create(device, workers, hbm_budget, max_batch, DRY_RUN if dry_run else 0,
                                   ctypes.byref(h))
        if rc:
            raise A.StarkError(rc, "stark_server_create failed")
        self.h = h

    def _submit(self, req):
        i = ctypes.c_uint64()
        rc = A.lib().stark_server_submit(self.h, ctypes.byref(req), ctypes.byref(i))
        if rc:
            raise A.StarkError(rc, A.lib().stark_server_last_error(self.h).decode())
        return i.value

    def submit_ds_ksk(self, dnskey, parent_ds_hash, fs_binding):
        key = bytes(dnskey)
        return self._submit(_Req(SERVE_DS_KSK, key, len(key), bytes(parent_ds_hash), None, None, 0,
                                 bytes(fs_binding)))

    def submit_inner_shard(self, salt, records, fs_binding):
        recs = records if isinstance(records, A.PackedShards) else A.PackedShards([records])
        assert len(recs) == 1
        return self._submit(_Req(SERVE_INNER_SHARD, None, 0, None, bytes(salt), recs.arr, recs.counts[0],
                                 bytes(fs_binding)))

    def submit_rsa(self, record, public_inputs_hash):
        """one RSA-2048 record (an api.RsaStackedRecord: n, s < n, em)"""
        def be(x):
            return int(x).to_bytes(256, "big")
        return self._submit(_Req(SERVE_RSA, None, 0, None, None, None, 0, bytes(public_inputs_hash),
                                 be(record.n), be(record.s), be(record.em)))

    def submit_p256(self, record):
        """one ECDSA-P256 record (an api.EcdsaRecord); its public-input hash is
        the record's pk_hash_pre"""
        def be(x):
            return int(x).to_bytes(32, "big")
        return self._submit(_Req(SERVE_P256, None, 0, None, None, None, 0, None, None, None, None,
                                 be(record.qx), be(record.qy), be(record.r), be(record.s), bytes(record.digest)))

    def wait(self, rid, timeout=600.0):
        r = _Res()
        rc = A.lib().stark_server_wait(self.h, rid, timeout, ctypes.byref(r))
        if rc == E_TIMEOUT:
            raise TimeoutError("request %d not done within %.1f s" % (rid, timeout))
        if rc:
            raise A.StarkError(rc, "stark_server_wait failed")
        blob = A._take(r.proof, r.proof_len)
        if r.status:
            raise A.StarkError(r.status, A.lib().stark_server_last_error(self.h).decode())
        return ServeResult(r.id, r.status, r.kind, bytes(r.pi_hash), bytes(r.root_f0), bytes(r.aux), r.n_trace,
                           r.batch_size, r.batch_seq, blob, r.t_submit, r.t_start, r.t_done)

    def stats(self):
        c, p, b = A.sz(), A.sz(), A.sz()
        A.lib().stark_server_stats(self.h, ctypes.byref(c), ctypes.byref(p), ctypes.byref(b))
        return {"completed": c.value, "peak_bytes_admitted": p.value, "budget": b.value}

    def close(self):
        if self.h:
            A.lib().stark_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
