"""The oracle's threaded mode (the loops the reference runs with rayon, on
n threads: Merkle leaves / levels at >= 256, FRI folds, coset extraction and
merges at >= 2^12) computes the same bytes as one thread: the prover-core
proof, the SHA-256 and Ed25519 merges and a DS->KSK proof. The threaded mode
is what the CPU baseline times on the host's cores."""
import ctypes
import hashlib

import numpy as np

from tests import _oracle as O
from tests import ed25519_py as E


def _prover_core(tr, blowup, schedule, s0):
    w, n = tr.shape
    rr = ctypes.create_string_buffer(32)
    pp, pl = O.u8p(), O.sz()
    rc = O.lib().ref_prover_core(O.ptr(tr), w, n, blowup, 7, O.fri_params(schedule, r=s0, s0=s0), rr, None, pp, pl)
    assert rc == 0
    return rr.raw, O.take_bytes(pp, pl)


def test_prover_core_threads_equal_serial():
    rng = np.random.default_rng(3)
    tr = rng.integers(0, O.P, size=(8, 1 << 11), dtype=np.uint64)
    one = _prover_core(tr, 8, [8, 8, 8, 8, 4], 20)  # 2^14 points: every threaded site engages
    with O.threads(6):
        many = _prover_core(tr, 8, [8, 8, 8, 8, 4], 20)
    assert one == many


def test_merges_threads_equal_serial():
    rng = np.random.default_rng(4)
    h, blowup = 256, 32
    lde = rng.integers(0, O.P, size=(O.SHA256_W, h * blowup), dtype=np.uint64)
    one = O.merge_sha256(lde, h, blowup, 2)
    with O.threads(5):
        many = O.merge_sha256(lde, h, blowup, 2)
    assert np.array_equal(one, many)
    _, pk, msg, sig = E.rfc_vector(1)
    e = O.Ed25519(pk, sig, msg, 8)
    lde = rng.integers(0, O.P, size=(e.width, e.height * 4), dtype=np.uint64)
    one = e.merge(lde, 4)
    with O.threads(7):
        many = e.merge(lde, 4)
    assert np.array_equal(one, many)


def test_ds_ksk_proof_threads_equal_serial():
    key = bytes(range(70))
    one = O.prove_ds_ksk_binding(key, hashlib.sha256(key).digest(), b"\x11" * 32)
    with O.threads(4):
        many = O.prove_ds_ksk_binding(key, hashlib.sha256(key).digest(), b"\x11" * 32)
    assert one == many
