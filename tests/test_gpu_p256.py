"""GPU parity of the double-chain ECDSA-P256 AIR (stark_merge_ecdsa_p256,
stark_prove_p256_chains_batch, stark_prove_ecdsa_p256_batch) against the
oracle: c_eval on random and valid LDEs, proofs byte for byte at small K,
and the full 256-row record proof against the oracle's stored output
(tests/golden/p256_ecdsa_record.json, written by make_p256_fixture.py)."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from tests import _oracle as O
from tests.test_p256_air import GX, GY, N, P as FP, ec_mul, ecdsa_sign, two_pass

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def S():
    import stark_swarm_amd as S
    S.default_context()
    return S


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def coeffs():
    return np.arange(1, O.p256_n_constraints() + 1, dtype=np.uint64)


@pytest.mark.parametrize("n_trace,blowup", [(4, 32), (8, 8)])
def test_merge_p256_random_lde(S, n_trace, blowup):
    """every one of the 155 774 constraints is non-zero on a random LDE, so
    this checks the regrouped sums of every gadget kind and the tail"""
    say("merge_p256 random lde %d x %d" % (n_trace, blowup))
    n = n_trace * blowup
    lde = np.random.default_rng(n).integers(0, O.P, size=(O.p256_width(), n), dtype=np.uint64)
    want = O.merge_p256(lde, n_trace, blowup)
    got, info = S.deep_ali_merge_ecdsa_double_multirow_streaming(lde, coeffs(), 0, n_trace, blowup)
    assert np.array_equal(got, want)
    assert info.num_constraints == 155774 and info.trace_width == 151334


def test_merge_p256_valid_and_planted_trace(S):
    say("merge_p256 valid trace")
    pts = [GX, GY, 1, GX, GY, 1] + [ec_mul(2, (GX, GY))[0], ec_mul(2, (GX, GY))[1], 1] * 2
    tr, _ = two_pass(pts, [1, 0, 1, 1], [0, 1, 1, 0], 4)
    lde = O.lde_columns(tr, 32)
    got, _ = S.deep_ali_merge_ecdsa_double_multirow_streaming(lde, coeffs(), 0, 4, 32)
    assert np.array_equal(got, O.merge_p256(lde, 4, 32))
    tr[O.p256_gadgets(1)[40][1][4] + 3, 2] = 7  # a carry cell of chain B
    lde = O.lde_columns(tr, 32)
    got, _ = S.deep_ali_merge_ecdsa_double_multirow_streaming(lde, coeffs(), 0, 4, 32)
    assert np.array_equal(got, O.merge_p256(lde, 4, 32))


def _oracle_chain_proof(pts, rproj, a_bits, b_bits, n_trace, pis):
    tr = O.p256_trace(pts, rproj, a_bits, b_bits, n_trace)
    lde = O.lde_columns(tr, 32)
    c = O.merge_p256(lde, n_trace, 32)
    lg = (32 * n_trace).bit_length() - 1
    sched = [8] * (lg // 3) + ([1 << (lg % 3)] if lg % 3 else [])
    return O.deep_fri_prove(c, O.fri_params(sched, r=54, s0=54, pi_hash=pis))


def test_prove_p256_chains_matches_oracle(S):
    """run_double_chain_streaming at K = 5 and K = 8 (n_trace 8), one with
    r_proj bound to the outputs, one with given r_proj values (a trace whose
    boundary does not hold), as one batch and alone"""
    say("prove_p256_chains")
    g, q = (GX, GY), ec_mul(2, (GX, GY))
    pts = [g[0], g[1], 1, g[0], g[1], 1, q[0], q[1], 1, q[0], q[1], 1]
    a1, b1 = [i % 2 == 0 for i in range(5)], [i % 3 == 0 for i in range(5)]
    a2, b2 = [1, 1, 0, 1, 0, 0, 1, 1], [0, 0, 1, 1, 1, 0, 1, 0]
    _, rp1 = two_pass(pts, [int(b) for b in a1], [int(b) for b in b1], 8)
    rp2 = [12345 + i for i in range(6)]
    pis = [bytes(32), hashlib.sha3_256(b"p256").digest()]
    chains = [S.P256Chains(tuple(pts[0:3]), tuple(pts[3:6]), tuple(pts[6:9]), tuple(pts[9:12]), a1, b1),
              S.P256Chains(tuple(pts[0:3]), tuple(pts[3:6]), tuple(pts[6:9]), tuple(pts[9:12]), a2, b2,
                           rproj=tuple(rp2), bind_outputs=False)]
    outs = S.prove_p256_chains_batch(chains, 8, pis)
    assert list(outs[0].rproj) == rp1 and list(outs[1].rproj) == rp2
    want = [_oracle_chain_proof(pts, rp1, [int(b) for b in a1], [int(b) for b in b1], 8, pis[0]),
            _oracle_chain_proof(pts, rp2, a2, b2, 8, pis[1])]
    for o, w, p in zip(outs, want, pis):
        assert o.proof_blob == w and o.n_trace == 8
        assert o.pi_hash == hashlib.sha3_256(p + o.root_f0).digest()
    one = S.prove_p256_chains_batch(chains[1:], 8, pis[1:])[0]
    assert one.proof_blob == outs[1].proof_blob


def _fixture():
    with open(os.path.join(GOLD, "p256_ecdsa_record.json")) as f:
        return json.load(f)["records"]


def _record(S, d):
    return S.EcdsaRecord(int(d["qx"], 16), int(d["qy"], 16), int(d["r"], 16), int(d["s"], 16),
                         bytes.fromhex(d["digest"]))


def test_prove_ecdsa_record_equals_oracle_fixture(S):
    """the 256-row per-record prover, byte for byte against the oracle's proof
    of the same seeded record; alone and in a batch with a second record"""
    recs = _fixture()
    say("prove_record_ecdsa fixture records %d" % len(recs))
    mine = [_record(S, d) for d in recs]
    outs = S.prove_record_ecdsa_batch(mine)
    for o, d, r in zip(outs, recs, mine):
        assert o.n_trace == 256 and len(o.proof_blob) == d["proof_len"]
        assert hashlib.sha3_256(o.proof_blob).hexdigest() == d["proof_sha3_256"]
        assert o.root_f0.hex() == d["root_f0"] and o.pi_hash.hex() == d["pi_hash"]
        assert o.pi_hash == hashlib.sha3_256(S.ecdsa_record_pk_hash(r) + o.root_f0).digest()
    say("prove_record_ecdsa batch of 3")
    extra = [S.EcdsaRecord(*ecdsa_sign(100 + i, b"batch %d" % i)) for i in range(2)]
    batch = S.prove_record_ecdsa_batch(mine[:1] + extra)
    assert batch[0].proof_blob == outs[0].proof_blob
    alone = S.prove_record_ecdsa_batch(extra[1:])[0]
    assert batch[2].proof_blob == alone.proof_blob and batch[2].pi_hash == alone.pi_hash
    # the bound r_proj is (2^256 + u) times the base point, per chain
    r = extra[1]
    w = pow(r.s, -1, N)
    u1, u2 = int.from_bytes(r.digest, "big") % N * w % N, r.r * w % N
    for ch, (u, base) in enumerate(((u1, (GX, GY)), (u2, (r.qx, r.qy)))):
        X, Y, Z = alone.rproj[3 * ch:3 * ch + 3]
        x, y = ec_mul(2**256 + u, base)
        assert (X - x * Z) % FP == 0 and (Y - y * Z) % FP == 0


def test_prove_ecdsa_rejects_bad_records(S):
    qx, qy, r, s, dg = ecdsa_sign(5, b"bad")
    with pytest.raises(S.StarkError):
        S.prove_record_ecdsa_batch([S.EcdsaRecord(qx, qy, r, (s + 1) % N, dg)])
    with pytest.raises(S.StarkError):
        S.prove_record_ecdsa_batch([S.EcdsaRecord(qx, (qy + 1) % FP, r, s, dg)])
    # the signature checks run beside the proof pipeline: a batch with one
    # record that does not verify still fails, naming that record
    with pytest.raises(S.StarkError, match="record 1"):
        S.prove_record_ecdsa_batch([S.EcdsaRecord(qx, qy, r, s, dg), S.EcdsaRecord(qx, qy, r, (s + 1) % N, dg)])


def test_prove_p256_chains_low_degree_path(S):
    """n_trace 32: the prover composes the gadget constraints on the
    blowup-2 LDE and interpolates, the tail on the full LDE; byte for byte
    against the oracle's direct evaluation at all 1024 points (uneven chain
    lengths; a second input with r_proj values whose boundary fails)"""
    say("prove_p256_chains n_trace 32")
    g, q = (GX, GY), ec_mul(3, (GX, GY))
    pts = [g[0], g[1], 1, g[0], g[1], 1, q[0], q[1], 1, q[0], q[1], 1]
    a = [int((i * 7) % 3 == 0) for i in range(20)]
    b = [int((i * 5) % 4 == 1) for i in range(27)]
    tr = O.p256_trace(pts, [0] * 6, a, b, 32)
    c = O.p256_columns()
    rp = [sum(int(tr[c["sel"][ch][i] + l, (19, 26)[ch]]) << (26 * l) for l in range(10))
          for ch in range(2) for i in range(3)]
    bad = [v ^ 5 for v in rp]
    pis = [hashlib.sha3_256(b"low-degree %d" % i).digest() for i in range(2)]
    chains = [S.P256Chains(tuple(pts[0:3]), tuple(pts[3:6]), tuple(pts[6:9]), tuple(pts[9:12]), a, b),
              S.P256Chains(tuple(pts[0:3]), tuple(pts[3:6]), tuple(pts[6:9]), tuple(pts[9:12]), a, b,
                           rproj=tuple(bad), bind_outputs=False)]
    outs = S.prove_p256_chains_batch(chains, 32, pis)
    assert list(outs[0].rproj) == rp
    assert outs[0].proof_blob == _oracle_chain_proof(pts, rp, a, b, 32, pis[0])
    assert outs[1].proof_blob == _oracle_chain_proof(pts, bad, a, b, 32, pis[1])
