"""ProvingServer -- the device side of a swarm worker's request loop
(crates/swarm-worker/src/main.rs:220-335) on one GPU: requests are admitted
by an HBM budget and proven with continuous batching (stark_server_* in
include/stark_gpu.h). Each result equals the single-call result of
prove_ds_ksk_binding / prove_inner_shard / prove_zsk_ksk_binding_v2 / the
RSA-2048 or ECDSA-P256 record prover for the same request.

Isolation: submit rejects a signature request whose signature does not
verify; a batch that still fails is proved again one request at a time, so
only the failing request raises (with its own error). A request submitted
with a deadline that no worker admits in time raises StarkError(E_TIMEOUT)."""
import ctypes
from dataclasses import dataclass

from . import api as A

SERVE_DS_KSK, SERVE_INNER_SHARD, SERVE_RSA, SERVE_P256, SERVE_ED25519 = 0, 1, 2, 3, 4
DRY_RUN, NO_VALIDATE = 1, 2
E_TIMEOUT = -6


class _Req(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("dnskey", ctypes.c_char_p), ("dnskey_len", A.sz),
                ("parent_ds_hash", ctypes.c_char_p), ("salt", ctypes.c_char_p),
                ("records", ctypes.POINTER(A._DnsRecord)), ("n_records", A.sz), ("fs_binding", ctypes.c_char_p),
                ("rsa_n", ctypes.c_char_p), ("rsa_s", ctypes.c_char_p), ("rsa_em", ctypes.c_char_p),
                ("p256_qx", ctypes.c_char_p), ("p256_qy", ctypes.c_char_p), ("p256_r", ctypes.c_char_p),
                ("p256_s", ctypes.c_char_p), ("p256_digest", ctypes.c_char_p),
                ("ed_pk", ctypes.c_char_p), ("ed_sig", ctypes.c_char_p), ("ed_msg", ctypes.c_char_p),
                ("ed_msg_len", A.sz), ("ed_merkle_root", ctypes.c_char_p), ("deadline_s", ctypes.c_double)]


class _Res(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int), ("kind", ctypes.c_int),
                ("pi_hash", ctypes.c_uint8 * 32), ("root_f0", ctypes.c_uint8 * 32), ("aux", ctypes.c_uint8 * 32),
                ("n_trace", A.sz), ("batch_size", A.sz), ("batch_seq", ctypes.c_uint64), ("proof", A.u8p),
                ("proof_len", A.sz),
                ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_done", ctypes.c_double),
                ("error", ctypes.c_char * 192)]


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
    def __init__(self, device=0, workers=4, hbm_budget=64 << 30, max_batch=64, dry_run=False, validate=True):
        """validate=False skips the signature checks at submit (only to test
        that a bad request in a batch fails alone)"""
        L = _bind()
        h = ctypes.c_void_p()
        flags = (DRY_RUN if dry_run else 0) | (0 if validate else NO_VALIDATE)
        rc = L.stark_server_create(device, workers, hbm_budget, max_batch, flags, ctypes.byref(h))
        if rc:
            raise A.StarkError(rc, "stark_server_create failed")
        self.h = h

    def _submit(self, req, deadline=None):
        if deadline is not None:
            req.deadline_s = float(deadline)
        i = ctypes.c_uint64()
        rc = A.lib().stark_server_submit(self.h, ctypes.byref(req), ctypes.byref(i))
        if rc:
            raise A.StarkError(rc, A.lib().stark_server_last_error(self.h).decode())
        return i.value

    def submit_ds_ksk(self, dnskey, parent_ds_hash, fs_binding, deadline=None):
        key = bytes(dnskey)
        return self._submit(_Req(SERVE_DS_KSK, key, len(key), bytes(parent_ds_hash), None, None, 0,
                                 bytes(fs_binding)), deadline)

    def submit_inner_shard(self, salt, records, fs_binding, deadline=None):
        recs = records if isinstance(records, A.PackedShards) else A.PackedShards([records])
        assert len(recs) == 1
        return self._submit(_Req(SERVE_INNER_SHARD, None, 0, None, bytes(salt), recs.arr, recs.counts[0],
                                 bytes(fs_binding)), deadline)

    def submit_rsa(self, record, public_inputs_hash, deadline=None):
        """one RSA-2048 record (an api.RsaStackedRecord: n, s < n, em)"""
        def be(x):
            return int(x).to_bytes(256, "big")
        return self._submit(_Req(SERVE_RSA, None, 0, None, None, None, 0, bytes(public_inputs_hash),
                                 be(record.n), be(record.s), be(record.em)), deadline)

    def submit_p256(self, record, deadline=None):
        """one ECDSA-P256 record (an api.EcdsaRecord); its public-input hash is
        the record's pk_hash_pre"""
        def be(x):
            return int(x).to_bytes(32, "big")
        return self._submit(_Req(SERVE_P256, None, 0, None, None, None, 0, None, None, None, None,
                                 be(record.qx), be(record.qy), be(record.r), be(record.s), bytes(record.digest)),
                            deadline)

    def submit_ed25519(self, pk, sig, msg, fs_binding, merkle_root, deadline=None):
        """one Ed25519 ZSK->KSK binding (prove_zsk_ksk_binding_v2 at k_scalar
        256): 32-byte public key, 64-byte signature over msg"""
        pk, sig, msg = bytes(pk), bytes(sig), bytes(msg)
        assert len(pk) == 32 and len(sig) == 64
        q = _Req(SERVE_ED25519, fs_binding=bytes(fs_binding), ed_pk=pk, ed_sig=sig, ed_msg=msg,
                 ed_msg_len=len(msg), ed_merkle_root=bytes(merkle_root))
        return self._submit(q, deadline)

    def wait(self, rid, timeout=600.0):
        r = _Res()
        rc = A.lib().stark_server_wait(self.h, rid, timeout, ctypes.byref(r))
        if rc == E_TIMEOUT:
            raise TimeoutError("request %d not done within %.1f s" % (rid, timeout))
        if rc:
            raise A.StarkError(rc, "stark_server_wait failed")
        blob = A._take(r.proof, r.proof_len)
        if r.status:
            raise A.StarkError(r.status, r.error.decode(errors="replace"))
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
