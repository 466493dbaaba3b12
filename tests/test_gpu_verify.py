"""GPU: classic DEEP-FRI (LdtMode::Fri) proofs against the oracle byte for
byte, and every proof kind the library makes accepted by the library's host
verifier and the oracle's (deep_fri_verify, fri.rs:2136-2710), with tampered
copies rejected.

Marked gpu: needs an MI355X and the built libstark_gpu.so.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from tests import _oracle as O
from tests import ed25519_py as E
from tests.proofwalk import walk

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
P = O.P


@pytest.fixture(scope="module")
def S():
    import stark_swarm_amd as S
    S.default_context()
    return S


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def low_degree_f0(n0, degree, seed):
    rng = np.random.default_rng(seed)
    c = np.zeros(n0, dtype=np.uint64)
    c[:degree] = rng.integers(0, P, size=degree, dtype=np.uint64)
    return O.ntt(c)


def check(S, params, proof, oparams=None):
    """accepted by both verifiers; a flipped root_f0 bit and a flipped bit of
    the final polynomial are rejected by both"""
    op = oparams if oparams is not None else O.fri_params(params.schedule, r=params.r, s0=params.s0,
                                                           seed_z=params.seed_z,
                                                           coeff_commit_final=params.coeff_commit_final,
                                                           d_final=params.d_final, stir=params.stir,
                                                           pi_hash=params.public_inputs_hash)
    assert S.deep_fri_verify(params, proof) is True
    assert O.deep_fri_verify(op, proof) == 1
    spans, _ = walk(proof)
    fin0 = [a for a, b, name in spans if name == "fin.0"][0]
    for pos in (3, fin0):  # root_f0 and the final polynomial, read in both modes
        bad = bytearray(proof)
        bad[pos] ^= 0x10
        bad = bytes(bad)
        try:
            lv = S.deep_fri_verify(params, bad)
        except S.StarkError:
            lv = None
        ov = O.deep_fri_verify(op, bad)
        assert lv is not True and ov != 1
        assert (lv is None) == (ov == -1)


@pytest.mark.parametrize("n0,schedule,cc,d_final,deg", [
    (1 << 10, [2] * 10, True, 1, 1 << 10),
    (1 << 12, [8, 4, 8, 16], True, 1, 1 << 12),
    (1 << 11, [4, 8, 4], True, 4, 64),
    (1 << 9, [8, 8, 8], False, 1, 1 << 9),
    (1 << 14, [16, 16, 8, 8], True, 1, 1 << 14),
])
def test_classic_fri_matches_oracle(S, n0, schedule, cc, d_final, deg):
    say("classic fri n0=%d schedule=%s" % (n0, schedule))
    f0 = low_degree_f0(n0, deg, n0 ^ len(schedule))
    pi = hashlib.sha3_256(b"classic %d" % n0).digest()
    params = S.DeepFriParams(schedule, 20, 0xDEEFBAAD, cc, d_final, False, 20, pi)
    got = S.deep_fri_prove(f0, S.FriDomain.new_radix2(n0), params)
    want = O.deep_fri_prove(f0, O.fri_params(schedule, r=20, s0=20, coeff_commit_final=cc, d_final=d_final,
                                              stir=False, pi_hash=pi))
    assert got == want
    check(S, params, got)


def test_classic_fri_batch(S):
    n0, sch = 1 << 11, [2] * 11
    f0s = np.stack([low_degree_f0(n0, n0, 50 + k) for k in range(3)])
    pis = [bytes([k + 1]) * 32 for k in range(3)]
    params = S.DeepFriParams(sch, 54, 0xDEEFBAAD, True, 1, False, 54, None)
    got = S.deep_fri_prove_batch(f0s, S.FriDomain.new_radix2(n0), params, pis)
    for k in range(3):
        p1 = S.DeepFriParams(sch, 54, 0xDEEFBAAD, True, 1, False, 54, pis[k])
        assert got[k] == S.deep_fri_prove(f0s[k], S.FriDomain.new_radix2(n0), p1)
        assert S.deep_fri_verify(p1, got[k])
    assert S.deep_fri_verify_batch(params, got, pis) == [True, True, True]
    assert S.deep_fri_verify_batch(params, got, pis[::-1]) == [False, True, False]


def _golden_shard(S):
    d = json.load(open(os.path.join(GOLD, "primary_state.json")))
    from stark_swarm_amd.zone import build_synthetic_zone, split_evenly
    salt = bytes.fromhex(d["zone_salt_hex"])
    w = d["shards"][0]
    recs = split_evenly(build_synthetic_zone(d["record_count"]), 3)[0]
    fsb = S.shard_fs_binding(bytes.fromhex(d["authority_pk_hash_hex"]), bytes.fromhex(d["job_id_hex"]),
                             w["shard_id"], bytes.fromhex(w["shard_nonce_hex"]))
    return d, salt, recs, fsb


def test_inner_shard_and_outer_rollup_fri_mode(S):
    d, salt, recs, fsb = _golden_shard(S)
    say("prove_inner_shard LdtMode::Fri (1024 records, n0 2^17, 17 arity-2 layers)")
    out = S.prove_inner_shard(salt, recs, fsb, S.LdtMode.Fri)
    leaves = O.zone_leaf_hashes(0, 1024, salt)
    with O.ldt_fri():
        proof, pi, root, rf = O.prove_inner_shard(salt, leaves, fsb, phi_mode=1)
    assert out.proof_blob == proof and out.pi_hash == pi and out.root_f0 == rf
    stir = S.prove_inner_shard(salt, recs, fsb, S.LdtMode.Stir)
    assert stir.proof_blob != out.proof_blob and stir.merkle_root == out.merkle_root
    n0 = out.n_trace * 32
    check(S, S.build_params(n0, fsb, S.LdtMode.Fri), out.proof_blob)
    # the STIR parameters do not accept the classic proof, nor the reverse
    assert S.deep_fri_verify(S.build_params(n0, fsb, S.LdtMode.Stir), out.proof_blob) is False
    assert S.deep_fri_verify(S.build_params(n0, fsb, S.LdtMode.Fri), stir.proof_blob) is False
    say("prove_outer_rollup LdtMode::Fri")
    pis = [bytes.fromhex(h) for h in d["inner_pi_hashes_hex"]]
    pk = bytes.fromhex(d["authority_pk_hash_hex"])
    o = S.prove_outer_rollup(pis, pk, S.LdtMode.Fri)
    with O.ldt_fri():
        proof, rf = O.prove_outer_rollup(b"".join(pis), pk, phi_mode=1)
    assert o.proof_blob == proof
    check(S, S.build_params(o.n0, pk, S.LdtMode.Fri), o.proof_blob)


def test_ds_ksk_nsec3_zsk_fri_mode(S):
    from tests.test_gpu_parity import _nsec3_chain
    say("prove_ds_ksk_binding LdtMode::Fri")
    key = bytes(range(40))
    parent = hashlib.sha256(key).digest()
    fsb = hashlib.sha3_256(b"dsksk-fri").digest()
    out = S.prove_ds_ksk_binding(key, parent, fsb, S.LdtMode.Fri)
    with O.ldt_fri():
        proof, pi, dg, rf, nb, nt = O.prove_ds_ksk_binding(key, parent, fsb)
    assert out.proof_blob == proof and out.pi_hash == pi
    check(S, S.build_params(nt * 32, fsb, S.LdtMode.Fri), out.proof_blob)
    say("prove_nsec3_completeness LdtMode::Fri")
    recs = _nsec3_chain(9)
    salt, fsb = b"swarm-test-zone1", hashlib.sha3_256(b"nsec3-fri").digest()
    o = S.prove_nsec3_completeness([S.Nsec3Record(a, b) for a, b in recs], salt, fsb, S.LdtMode.Fri)
    with O.ldt_fri():
        proof, pi, cr, rf, nt = O.prove_nsec3_completeness(recs, salt, fsb)
    assert o.proof_blob == proof and o.pi_hash == pi
    check(S, S.build_params(nt * 32, fsb, S.LdtMode.Fri), o.proof_blob)
    say("prove_zsk_ksk_binding_v2 LdtMode::Fri (k_scalar 8, blowup 4)")
    seed = hashlib.sha256(b"ed25519-fri").digest()
    msg = b"fri.example. 300 IN A 192.0.2.9"
    pk, sig = E.sign(seed, msg)
    fs, mr = bytes([0x33] * 32), bytes(range(32))
    z = S.prove_zsk_ksk_binding_v2_batch([(pk, sig, msg, fs, mr)], k_scalar=8, blowup=4, ldt=S.LdtMode.Fri)[0]
    with O.ldt_fri():
        proof, rf, pi = O.prove_zsk_ksk_binding_v2(pk, sig, msg, fs, mr, 8, 4)
    assert z.proof_blob == proof and z.pi_hash == pi
    check(S, S.build_params(z.n_trace * 4, fs, S.LdtMode.Fri), z.proof_blob)


def test_every_stir_proof_kind_verifies(S):
    """the library's STIR proofs of every whole-proof caller, under the
    parameters each caller binds"""
    from tests.test_gpu_parity import _nsec3_chain, _rsa_rec
    from tests.test_p256_air import ecdsa_sign
    d, salt, recs, fsb = _golden_shard(S)
    say("verify: inner shard batch")
    outs = S.prove_inner_shard_batch(salt, [recs[:300], recs[300:]], [fsb, bytes(32)])
    for o, f in zip(outs, [fsb, bytes(32)]):
        check(S, S.build_params(o.n_trace * 32, f), o.proof_blob)
    say("verify: DS-KSK batch")
    keys = [bytes(range(i, i + 30 + 70 * i)) for i in range(3)]
    fsbs = [bytes([i + 7]) * 32 for i in range(3)]
    ds = S.prove_ds_ksk_binding_batch(keys, [hashlib.sha256(k).digest() for k in keys], fsbs)
    for o, f in zip(ds, fsbs):
        check(S, S.build_params(o.n_trace * 32, f), o.proof_blob)
    say("verify: uneven SHA-256/512 batch")
    msgs = [b"a" * 10, b"b" * 200, b"c" * 500]
    sha = S.prove_sha_batch([256, 512, 256], msgs, fsbs, heights=[1 << 10, 1 << 11, 1 << 12])
    for o, f in zip(sha, fsbs):
        check(S, S.build_params(o.n_trace * 32, f), o.proof_blob)
    say("verify: NSEC3")
    n3 = S.prove_nsec3_completeness([S.Nsec3Record(a, b) for a, b in _nsec3_chain(20)], salt, fsb)
    check(S, S.build_params(n3.n_trace * 32, fsb), n3.proof_blob)
    say("verify: RSA-2048")
    pis = [hashlib.sha3_256(b"rsa-verify").digest()]
    rsa = S.prove_rsa_exp_batch([S.RsaStackedRecord(*_rsa_rec(0x77))], pis)[0]
    check(S, S.build_params(rsa.n_trace * 32, pis[0]), rsa.proof_blob)
    say("verify: ECDSA-P256 record")
    rec = S.EcdsaRecord(*ecdsa_sign(33, b"verify me"))
    pe = S.prove_record_ecdsa_batch([rec])[0]
    check(S, S.build_params(pe.n_trace * 32, S.ecdsa_record_pk_hash(rec)), pe.proof_blob)
    say("verify: Ed25519 v16 (k_scalar 8, blowup 4)")
    seed = hashlib.sha256(b"ed25519-verify").digest()
    msg = b"v.example. 60 IN TXT hello"
    pk, sig = E.sign(seed, msg)
    fs = bytes([0x44] * 32)
    z = S.prove_zsk_ksk_binding_v2_batch([(pk, sig, msg, fs, bytes(32))], k_scalar=8, blowup=4)[0]
    check(S, S.build_params(z.n_trace * 4, fs), z.proof_blob)
    say("verify: outer rollup")
    o = S.prove_outer_rollup([x.pi_hash for x in outs], bytes(range(32)))
    check(S, S.build_params(o.n0, bytes(range(32))), o.proof_blob)
    say("verify: batch verifier over all of them")
    proofs = [outs[0].proof_blob, ds[1].proof_blob, rsa.proof_blob]
    # one parameter set per call: the same-n0 proofs of one kind
    assert S.deep_fri_verify_batch(S.build_params(ds[0].n_trace * 32, fsbs[0]), [x.proof_blob for x in ds[:1]]) \
        == [True]
    assert proofs


def test_merkle_open_matches_oracle(S):
    rng = np.random.default_rng(9)
    for n, vpl, arity in [(1, 3, 2), (37, 18, 4), (1000, 6, 8), (4096, 1, 16)]:
        vals = rng.integers(0, P, size=(n, vpl), dtype=np.uint64)
        idx = sorted({0, n // 3, n - 1})
        root, ops = S.merkle_open(vals, arity, 0x1234, idx)
        assert root == O.merkle_root(vals, arity, 0x1234)
        cfg = S.MerkleChannelCfg.constant(n, arity, 0x1234)
        for i, op in zip(idx, ops):
            assert op == O.merkle_open(vals, arity, 0x1234, i)
            assert S.verify_opening(cfg, root, op)
