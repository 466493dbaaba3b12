"""GPU parity of the Ed25519 verify AIR v16 (stark_build_ed25519_trace,
stark_merge_ed25519, stark_prove_zsk_ksk_binding_v2_batch) against the oracle:
the device trace cell for cell, c_eval on random LDEs with random
coefficients (every gadget kind and every general-lambda path), proofs byte
for byte at k_scalar 8 / blowup 4 and, at the production shape (k_scalar 256,
blowup 32), against the oracle's stored outputs
(tests/golden/ed25519_zsk_ksk.json, written by make_ed25519_fixture.py)."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from tests import _oracle as O
from tests import ed25519_py as E

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FS = bytes([0x5A] * 32)
ROOT32 = bytes(range(32))


@pytest.fixture(scope="module")
def S():
    import stark_swarm_amd as S
    S.default_context()
    return S


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def two_block():
    seed = hashlib.sha256(b"ed25519-gpu-test").digest()
    msg = b"www.example.org. 300 IN A 192.0.2.7 " + bytes(range(60))
    pk, sig = E.sign(seed, msg)
    return pk, sig, msg


def test_device_trace_equals_host_and_oracle(S):
    _, pk, msg, sig = E.rfc_vector(0)
    say("ed25519 device trace, RFC 8032 TEST 1, k 256")
    dev = S.fill_verify_air_v16(pk, sig, msg, 256, device=True)
    assert np.array_equal(dev, S.fill_verify_air_v16(pk, sig, msg, 256))
    pk, sig, msg = two_block()
    say("ed25519 device trace, two-block message, k 8")
    dev = S.fill_verify_air_v16(pk, sig, msg, 8, device=True)
    assert np.array_equal(dev, O.Ed25519(pk, sig, msg, 8).trace())


@pytest.mark.parametrize("vec,blowup", [(1, 2), (None, 4)])
def test_merge_random_lde_random_coeffs(S, vec, blowup):
    """on a random LDE every constraint is non-zero; random coefficients
    exercise the general-lambda path of every block kind and of the SHA-512 rows"""
    if vec is None:
        pk, sig, msg = two_block()
    else:
        _, pk, msg, sig = E.rfc_vector(vec)
    e = O.Ed25519(pk, sig, msg, 8)
    n = e.height * blowup
    rng = np.random.default_rng(7 + blowup)
    say("merge_ed25519 random lde %d x %d, blowup %d" % (e.width, n, blowup))
    lde = rng.integers(0, O.P, size=(e.width, n), dtype=np.uint64)
    co = rng.integers(0, O.P, size=e.n_constraints, dtype=np.uint64)
    want = e.merge(lde, blowup, co)
    got, info = S.deep_ali_merge_ed25519_verify_streaming(lde, co, pk, sig, msg, 0, e.height, blowup, k_scalar=8)
    assert np.array_equal(got, want)
    assert info.num_constraints == e.n_constraints and info.trace_width == e.width


def test_merge_valid_trace_reference_coeffs(S):
    _, pk, msg, sig = E.rfc_vector(2)
    e = O.Ed25519(pk, sig, msg, 8)
    say("merge_ed25519 valid trace lde, blowup 4")
    lde = O.lde_columns(e.trace(), 4)
    co = np.arange(1, e.n_constraints + 1, dtype=np.uint64)
    got, _ = S.deep_ali_merge_ed25519_verify_streaming(lde, co, pk, sig, msg, 0, e.height, 4, k_scalar=8)
    assert np.array_equal(got, e.merge(lde, 4))


def test_prove_batch_matches_oracle_small(S):
    """k_scalar 8, blowup 4: the three RFC vectors and a two-block message (two
    trace shapes in one call), each byte-identical to the oracle's proof"""
    reqs = []
    for i in range(3):
        _, pk, msg, sig = E.rfc_vector(i)
        reqs.append((pk, sig, msg, FS, ROOT32))
    pk, sig, msg = two_block()
    reqs.append((pk, sig, msg, bytes([3] * 32), ROOT32))
    say("prove_zsk_ksk_binding_v2 batch of 4, k 8, blowup 4")
    outs = S.prove_zsk_ksk_binding_v2_batch(reqs, k_scalar=8, blowup=4)
    for r, o in zip(reqs, outs):
        proof, root_f0, pi_hash = O.prove_zsk_ksk_binding_v2(r[0], r[1], r[2], r[3], r[4], 8, 4)
        assert o.proof_blob == proof
        assert o.root_f0 == root_f0 and o.pi_hash == pi_hash


def test_prove_full_shape_matches_oracle_fixture(S):
    gold = json.load(open(os.path.join(GOLD, "ed25519_zsk_ksk.json")))["proofs"]
    reqs = [(bytes.fromhex(g["pk"]), bytes.fromhex(g["sig"]), bytes.fromhex(g["msg"]), bytes.fromhex(g["fs_binding"]),
             bytes.fromhex(g["merkle_root"])) for g in gold]
    say("prove_zsk_ksk_binding_v2 k 256 blowup 32, %d signatures batched" % len(reqs))
    outs = S.prove_zsk_ksk_binding_v2_batch(reqs)
    for g, o in zip(gold, outs):
        assert len(o.proof_blob) == g["proof_len"]
        assert hashlib.sha3_256(o.proof_blob).hexdigest() == g["proof_sha3_256"], g["name"]
        assert o.root_f0.hex() == g["root_f0"] and o.pi_hash.hex() == g["pi_hash"]
    say("prove_zsk_ksk_binding_v2 k 256, one at a time")
    for r, o in zip(reqs, outs):
        assert S.prove_zsk_ksk_binding_v2(*r).proof_blob == o.proof_blob


def test_invalid_signature_is_rejected(S):
    _, pk, msg, sig = E.rfc_vector(0)
    bad = bytearray(sig)
    bad[40] ^= 1
    with pytest.raises(S.StarkError) as ei:
        S.prove_zsk_ksk_binding_v2_batch([(pk, bytes(bad), msg, FS, ROOT32)])
    assert ei.value.code == -1
