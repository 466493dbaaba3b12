"""Pins for the oracle's deep_fri_verify (fri.rs:2136-2710), its classic-FRI
prover branch (fri.rs:1461-1600, 1724-1807) and the Merkle opening check
(crates/merkle/src/lib.rs:478-512, 754-799).

The reference's own proofs (tests/golden/*.bin, made and self-verified by the
reference prover, prover.rs:189-192) must be accepted under the reference's
parameters, and the verifier must react to exactly the fields the reference
reads: a bit flipped in any bound field rejects, while the fields
deep_fri_verify never reads (omega0; fz_per_layer in STIR mode; a classic
proof's separate f0_openings list and the child_pos / parent_index /
parent_pos of its query refs) leave the verdict at accept.
"""
import json
import os
import re

import numpy as np
import pytest

from tests import _oracle as O
from tests.proofwalk import walk

P = O.P
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def ps():
    return json.load(open(os.path.join(GOLD, "primary_state.json")))


def stir_schedule(n0):
    lg = n0.bit_length() - 1
    s = [8] * (lg // 3)
    if lg % 3:
        s.append(1 << (lg % 3))
    return s


def golden_inner():
    d = ps()
    w = d["shards"][0]
    fsb = O.shard_fs_binding(bytes.fromhex(d["authority_pk_hash_hex"]), bytes.fromhex(d["job_id_hex"]),
                             w["shard_id"], bytes.fromhex(w["shard_nonce_hex"]))
    n0 = w["n_trace"] * 32
    blob = open(os.path.join(GOLD, w["proof"]), "rb").read()
    return blob, O.fri_params(stir_schedule(n0), pi_hash=fsb), n0


def golden_outer():
    d = ps()
    blob = open(os.path.join(GOLD, "outer-rollup.bin"), "rb").read()
    return blob, O.fri_params(stir_schedule(d["outer_n0"]), pi_hash=bytes.fromhex(d["authority_pk_hash_hex"]))


def test_golden_proofs_verify():
    blob, p, _ = golden_inner()
    assert O.deep_fri_verify(p, blob) == 1
    blob, p = golden_outer()
    assert O.deep_fri_verify(p, blob) == 1


def test_golden_walk_covers_every_byte():
    blob, _, _ = golden_inner()
    spans, end = walk(blob)
    assert end == len(blob)
    assert sum(e - s for s, e, _ in spans) == len(blob)


def test_wrong_parameters_reject():
    blob, p, n0 = golden_inner()
    sch = stir_schedule(n0)
    fsb = p.public_inputs_hash
    bad = bytes([fsb[0] ^ 1]) + fsb[1:]
    assert O.deep_fri_verify(O.fri_params(sch, pi_hash=bad), blob) == 0
    assert O.deep_fri_verify(O.fri_params(sch, pi_hash=None), blob) == 0
    assert O.deep_fri_verify(O.fri_params(sch, pi_hash=fsb, seed_z=0xDEEFBAAE), blob) == 0
    assert O.deep_fri_verify(O.fri_params(sch, pi_hash=fsb, s0=53), blob) == 0
    assert O.deep_fri_verify(O.fri_params(sch, pi_hash=fsb, d_final=2), blob) == 0
    assert O.deep_fri_verify(O.fri_params(sch, pi_hash=fsb, coeff_commit_final=False), blob) == 0
    assert O.deep_fri_verify(O.fri_params(sch, pi_hash=fsb, stir=False), blob) == 0
    assert O.deep_fri_verify(O.fri_params([8, 8, 8, 8, 4, 8], pi_hash=fsb), blob) == 0
    # truncated bytes do not deserialize
    assert O.deep_fri_verify(p, blob[:-1]) == -1
    assert O.deep_fri_verify(p, blob[:100]) == -1


def flip(blob, start, end):
    """flip one low bit inside [start, end) keeping a field value canonical"""
    b = bytearray(blob)
    if end - start == 8 or (end - start) % 8 == 0 and end - start >= 48:
        v = int.from_bytes(b[start:start + 8], "little")
        for bit in range(8):
            nv = v ^ (1 << bit)
            if nv < P:
                break
        b[start:start + 8] = nv.to_bytes(8, "little")
    else:
        b[start] ^= 1
    return bytes(b)


def group_of(name):
    return re.sub(r"\.\d+", ".#", name)


def tamper_sweep(blob, params, unbound, per_group=3):
    spans, end = walk(blob)
    assert end == len(blob)
    by_group = {}
    for s, e, name in spans:
        by_group.setdefault(group_of(name), []).append((s, e, name))
    seen = set()
    for g, lst in sorted(by_group.items()):
        picks = [lst[0], lst[len(lst) // 2], lst[-1]][:per_group]
        for s, e, name in picks:
            if (s, e) in seen:
                continue
            seen.add((s, e))
            v = O.deep_fri_verify(params, flip(blob, s, e))
            if g in unbound:
                assert v == 1, "unread field %s changed the verdict" % name
            else:
                assert v != 1, "tampered %s was accepted" % name
    return by_group


def test_tamper_every_field_stir_golden():
    blob, p, _ = golden_inner()
    groups = tamper_sweep(blob, p, unbound={"omega0", "fz.#"})
    # the STIR proof carries the proximity queries, coset evaluations and coefficient tuples
    for g in ("sp.#.f0p.path.#.#", "sp.#.l1.path.#.#", "sp.#.fval.#", "sp.#.fnext", "ce.#.#", "ct.#.#", "croot"):
        assert g in groups


CLASSIC_UNBOUND = {"omega0", "q.#.refs.#.child_pos", "q.#.refs.#.parent_index", "q.#.refs.#.parent_pos",
                   "f0o.#.leaf", "f0o.#.path.#.#", "f0o.#.index"}


def low_degree_f0(n0, degree, seed):
    rng = np.random.default_rng(seed)
    c = np.zeros(n0, dtype=np.uint64)
    c[:degree] = rng.integers(0, P, size=degree, dtype=np.uint64)
    return O.ntt(c)


@pytest.mark.parametrize("n0,schedule,cc,d_final,deg", [
    (1 << 10, [2] * 10, True, 1, 1 << 10),   # LdtMode::Fri: arity-2 folds to one point
    (1 << 9, [8, 8, 8], True, 1, 1 << 9),
    (1 << 9, [8, 8, 8], False, 1, 1 << 9),
    (1 << 11, [4, 8, 4], True, 4, 64),        # n_final 16, a final polynomial of degree 3
    (1 << 8, [16, 2], True, 2, 32),
])
def test_classic_proofs_verify_and_bind(n0, schedule, cc, d_final, deg):
    f0 = low_degree_f0(n0, deg, n0 + len(schedule))
    pi = bytes(range(32))
    p = O.fri_params(schedule, r=12, s0=12, coeff_commit_final=cc, d_final=d_final, stir=False, pi_hash=pi)
    proof = O.deep_fri_prove(f0, p)
    assert O.deep_fri_verify(p, proof) == 1
    spans, end = walk(proof)
    assert end == len(proof)
    names = {group_of(n) for _, _, n in spans}
    assert "q.#.pay.#.q" in names and "lp.#.#.leaf" in names and "fz.#" in names
    assert "sp.#.base" not in names and "ce.#.#" not in names
    if n0 <= 1 << 9:
        tamper_sweep(proof, p, CLASSIC_UNBOUND, per_group=2)
    # the same f0 under STIR parameters is a different, also accepted, proof
    ps_ = O.fri_params(schedule, r=12, s0=12, coeff_commit_final=cc, d_final=d_final, stir=True, pi_hash=pi)
    sproof = O.deep_fri_prove(f0, ps_)
    assert O.deep_fri_verify(ps_, sproof) == 1
    assert O.deep_fri_verify(p, sproof) == 0


def test_high_degree_final_layer_rejects():
    # f0 of full degree folded to 16 points cannot meet d_final = 1
    f0 = low_degree_f0(1 << 9, 1 << 9, 3)
    for stir in (True, False):
        p = O.fri_params([8, 4], r=8, s0=8, stir=stir, d_final=1)
        assert O.deep_fri_verify(p, O.deep_fri_prove(f0, p)) == 0


def test_classic_deep_quotient_values():
    # q_i (x_i - z) = f_i - f(z) is what the verifier checks; here the payload
    # values are also checked against a direct big-integer Horner evaluation
    n0, schedule = 1 << 6, [4, 4, 4]
    f0 = low_degree_f0(n0, n0, 11)
    p = O.fri_params(schedule, r=4, s0=4, stir=False)
    proof = O.deep_fri_prove(f0, p)
    spans, _ = walk(proof)
    pos = {n: (s, e) for s, e, n in spans}

    def fp6_at(name):
        s, e = pos[name]
        return [int.from_bytes(proof[s + 8 * i:s + 8 * i + 8], "little") for i in range(6)]

    i0 = int.from_bytes(proof[pos["q.0.refs.0.i"][0]:pos["q.0.refs.0.i"][1]], "little")
    f = fp6_at("q.0.pay.0.f")
    assert f == [int(f0[i0]), 0, 0, 0, 0, 0]
    # s of layer 0 at i0 equals f of layer 1 at i0 mod n/m (the fold, repeated)
    assert fp6_at("q.0.pay.0.s") == fp6_at("q.0.pay.1.f")


def test_merkle_open_and_verify():
    rng = np.random.default_rng(21)
    for n, vpl, arity in [(1, 1, 2), (5, 3, 2), (16, 6, 4), (100, 2, 8), (37, 18, 16), (64, 1, 16)]:
        vals = rng.integers(0, P, size=(n, vpl), dtype=np.uint64)
        depth = O.merkle_depth(n, arity)
        root = O.merkle_root(vals, arity, 7)
        for idx in sorted({0, n // 2, n - 1}):
            op = O.merkle_open(vals, arity, 7, idx)
            assert O.merkle_verify_opening(arity, depth, 7, root, op) == 1
            assert O.merkle_verify_opening(arity, depth, 8, root, op) == (1 if n == 1 else 0)
            bad = bytearray(op)
            bad[0] ^= 1  # leaf digest
            assert O.merkle_verify_opening(arity, depth, 7, root, bytes(bad)) == 0
            if len(op) > 48:
                bad = bytearray(op)
                bad[48] ^= 1  # first sibling
                assert O.merkle_verify_opening(arity, depth, 7, root, bytes(bad)) == 0
            assert O.merkle_verify_opening(arity, depth, 7, root, op[:-1]) == -1


def test_golden_openings_verify():
    # the proximity queries' openings of the reference's proof, checked one by one
    blob, p, n0 = golden_inner()
    spans, _ = walk(blob)
    first = {}
    last = {}
    for s, e, name in spans:
        m = re.match(r"sp\.(\d+)\.(f0p|l1)\.", name)
        if m and not name.endswith(".some"):
            key = (int(m.group(1)), m.group(2))
            first.setdefault(key, s)
            last[key] = e
    root_f0 = blob[:32]
    root1 = blob[40 + 32:40 + 64]
    nn = n0 // 8
    a0 = O.pick_arity(nn, 16)
    a1 = O.pick_arity(nn, 8)
    for q in range(54):
        op = blob[first[(q, "f0p")]:last[(q, "f0p")]]
        assert O.merkle_verify_opening(a0, O.merkle_depth(nn, a0), 0xFF, root_f0, op) == 1
        op1 = blob[first[(q, "l1")]:last[(q, "l1")]]
        assert O.merkle_verify_opening(a1, O.merkle_depth(nn, a1), 1, root1, op1) == 1
        assert O.merkle_verify_opening(a1, O.merkle_depth(nn, a1), 1, root_f0, op1) == 0


# ---------------------------------------------------------------- the library's host verifier
def lib_params(p):
    """the same parameters as an oracle FriParams, for the library mirror"""
    import stark_swarm_amd as S
    sch = [p.schedule[i] for i in range(p.n_layers)]
    return S.DeepFriParams(sch, p.r, p.seed_z, bool(p.coeff_commit_final), p.d_final, bool(p.stir), p.s0,
                           p.public_inputs_hash)


def lib_verdict(p, proof):
    import stark_swarm_amd as S
    try:
        return 1 if S.deep_fri_verify(lib_params(p), proof) else 0
    except S.StarkError:
        return -1


def corpus():
    """(params, proof bytes) pairs: the golden proofs, tampered copies of them
    (a few per field kind), wrong parameters and the oracle's classic and STIR
    proofs"""
    out = []
    blob, p, n0 = golden_inner()
    out.append((p, blob))
    spans, _ = walk(blob)
    seen = set()
    for s, e, name in spans:
        g = group_of(name)
        if g in seen:
            continue
        seen.add(g)
        out.append((p, flip(blob, s, e)))
    out.append((p, blob[:-3]))
    out.append((O.fri_params(stir_schedule(n0), pi_hash=p.public_inputs_hash, s0=53), blob))
    blob2, p2 = golden_outer()
    out.append((p2, blob2))
    for n0, sch, cc, df, deg, stir in [(1 << 9, [2] * 9, True, 1, 1 << 9, False), (1 << 10, [8, 8, 4], False, 4,
                                       512, False), (1 << 8, [4, 4, 4, 4], True, 1, 256, True)]:
        f0 = low_degree_f0(n0, deg, n0)
        pp = O.fri_params(sch, r=10, s0=10, coeff_commit_final=cc, d_final=df, stir=stir, pi_hash=bytes(32))
        proof = O.deep_fri_prove(f0, pp)
        out.append((pp, proof))
        sp, _ = walk(proof)
        seen = set()
        for s, e, name in sp:
            g = group_of(name)
            if g not in seen:
                seen.add(g)
                out.append((pp, flip(proof, s, e)))
    return out


def test_library_verifier_equals_oracle():
    c = corpus()
    verdicts = [(O.deep_fri_verify(p, b), lib_verdict(p, b)) for p, b in c]
    assert all(a == b for a, b in verdicts), [i for i, (a, b) in enumerate(verdicts) if a != b]
    assert sum(1 for a, _ in verdicts if a == 1) >= 5 and sum(1 for a, _ in verdicts if a == 0) >= 20


def test_library_verify_batch():
    import stark_swarm_amd as S
    blob, p, n0 = golden_inner()
    d = ps()
    fsbs = []
    for w in d["shards"]:
        fsbs.append(O.shard_fs_binding(bytes.fromhex(d["authority_pk_hash_hex"]), bytes.fromhex(d["job_id_hex"]),
                                       w["shard_id"], bytes.fromhex(w["shard_nonce_hex"])))
    bad = flip(blob, 40, 48)
    lp = lib_params(p)
    got = S.deep_fri_verify_batch(lp, [blob, blob, bad, blob[:50]], [fsbs[0], fsbs[1], fsbs[0], fsbs[0]])
    assert got == [True, False, False, None]


def test_library_merkle_verify_opening_equals_oracle():
    import stark_swarm_amd as S
    rng = np.random.default_rng(4)
    for n, vpl, arity in [(1, 1, 2), (9, 3, 2), (64, 6, 8), (33, 18, 16)]:
        vals = rng.integers(0, P, size=(n, vpl), dtype=np.uint64)
        depth = O.merkle_depth(n, arity)
        root = O.merkle_root(vals, arity, 3)
        cfg = S.MerkleChannelCfg([arity] * depth, 3)
        for idx in sorted({0, n - 1}):
            op = O.merkle_open(vals, arity, 3, idx)
            assert S.verify_opening(cfg, root, op) is True
            assert S.compute_leaf_hash(cfg, idx, vals[idx]) == op[:32]
            bad = bytearray(op)
            bad[-1] ^= 1  # the index's top byte: the parent positions change (no path: nothing to check)
            assert S.verify_opening(cfg, root, bytes(bad)) is (n == 1)
            assert O.merkle_verify_opening(arity, depth, 3, root, bytes(bad)) == (n == 1)
