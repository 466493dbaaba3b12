"""GPU parity: every HIP stage against the oracle on the same seeded inputs,
bit-exact, and whole proofs against the reference's golden bytes.

Marked gpu: needs an MI355X and the built libstark_gpu.so (no CPU fallback).
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from tests import _oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
P = O.P


@pytest.fixture(scope="module")
def S():
    import stark_swarm_amd as S
    S.default_context()
    return S


def rnd(seed, shape):
    return np.random.default_rng(seed).integers(0, P, size=shape, dtype=np.uint64)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# ---------------------------------------------------------------- NTT
@pytest.mark.parametrize("lg,ncols", [(0, 3), (1, 2), (3, 5), (4, 1), (7, 9), (10, 4), (12, 2), (13, 3), (16, 2),
                                      (20, 1)])
@pytest.mark.parametrize("inverse", [False, True])
def test_ntt_matches_oracle(S, lg, ncols, inverse):
    say("ntt lg=%d ncols=%d inverse=%s" % (lg, ncols, inverse))
    a = rnd(100 + lg, (ncols, 1 << lg))
    got = S.ntt(a, inverse=inverse)
    for c in range(ncols):
        assert np.array_equal(got[c], O.ntt(a[c], inverse=inverse)), "column %d" % c


def test_ntt_edge_values(S):
    n = 1 << 14
    for v in [0, 1, P - 1, P - 2, 2**32, 2**32 - 1]:
        a = np.full((1, n), v, dtype=np.uint64)
        assert np.array_equal(S.ntt(a)[0], O.ntt(a[0]))
        assert np.array_equal(S.ntt(a, inverse=True)[0], O.ntt(a[0], inverse=True))


# ---------------------------------------------------------------- LDE
@pytest.mark.parametrize("w,lg,blowup", [(1, 3, 2), (4, 4, 32), (3, 8, 8), (5, 12, 4), (2, 13, 8), (4, 16, 32),
                                         (2, 18, 8)])
def test_lde_matches_oracle(S, w, lg, blowup):
    say("lde w=%d n=2^%d blowup=%d" % (w, lg, blowup))
    cols = rnd(200 + lg, (w, 1 << lg))
    got = S.lde_trace_columns(cols, 1 << lg, blowup)
    assert np.array_equal(got, O.lde_columns(cols, blowup))


@pytest.mark.parametrize("w,lg,blowup", [(33, 4, 2), (7, 5, 32), (33, 6, 4), (9, 7, 16), (3, 8, 32), (40, 9, 2),
                                         (5, 9, 16), (2, 9, 32)])
def test_small_lde_sizes(S, w, lg, blowup):
    """the one-pass small LDE (blowup x 2^lg <= 8192: all cosets of a column
    in one workgroup), with column counts that leave the last workgroup
    partly empty, and the first size past it (2^9 x 32)"""
    say("small lde w=%d n=2^%d blowup=%d" % (w, lg, blowup))
    cols = rnd(400 + lg + w, (w, 1 << lg))
    got = S.lde_trace_columns(cols, 1 << lg, blowup)
    assert np.array_equal(got, O.lde_columns(cols, blowup))


@pytest.mark.parametrize("lg,blowup", [(5, 8), (12, 8), (14, 8), (16, 4)])
def test_coset_lde_matches_oracle(S, lg, blowup):
    say("coset lde n=2^%d blowup=%d shift=7" % (lg, blowup))
    ctx = S.default_context()
    w, n = 3, 1 << lg
    cols = rnd(300 + lg, (w, n))
    d = ctx.alloc(cols.nbytes).upload(cols)
    o = ctx.alloc(cols.nbytes * blowup)
    ctx._check(S.lib().stark_lde_columns_dev(ctx.h, d.ptr, w, n, blowup, 7, o.ptr))
    assert np.array_equal(o.download((w, n * blowup)), O.lde_columns(cols, blowup, shift=7))


def test_lde_errors_like_reference(S):
    with pytest.raises(ValueError):
        S.lde_trace_columns([], 8, 4)
    with pytest.raises(ValueError):
        S.lde_trace_columns(rnd(1, (2, 12)), 12, 4)
    with pytest.raises(ValueError):
        S.lde_trace_columns(rnd(1, (2, 16)), 16, 3)
    with pytest.raises(ValueError):
        S.lde_trace_columns(rnd(1, (2, 16)), 8, 4)


# ---------------------------------------------------------------- merge
@pytest.mark.parametrize("lg,blowup,phi_mode", [(3, 32, 1), (7, 32, 1), (7, 32, 0), (8, 8, 0), (12, 32, 1),
                                                (10, 8, 2)])
def test_merge_matches_oracle(S, lg, blowup, phi_mode):
    say("merge n_trace=2^%d blowup=%d phi_mode=%d" % (lg, blowup, phi_mode))
    n = 1 << lg
    leaves = np.random.default_rng(lg).integers(0, 2**64, size=n, dtype=np.uint64) % np.uint64(P)
    tr = O.hash_rollup_trace(n, leaves)
    lde = O.lde_columns(tr, blowup)
    got, info = S.deep_ali_merge_general(lde, [1, 2, 3], S.AirType.HashRollup, pow(7, (P - 1) // (n * blowup), P),
                                         n, blowup, phi_mode=phi_mode)
    assert np.array_equal(got, O.merge_hash_rollup(lde, n, blowup, phi_mode=phi_mode))
    assert info.num_constraints == 3 and info.trace_width == 4


# ---------------------------------------------------------------- Merkle
# 256 / 258 leaves of arity 2: the first node level has exactly 128 nodes (hashed
# with the tree's top in one launch) or 129 (hashed alone, the top after it)
@pytest.mark.parametrize("n,vpl,arity,label", [(1, 3, 2, 0), (7, 1, 2, 3), (64, 6, 8, 1), (100, 8, 16, 0xFF),
                                               (1000, 13, 4, 7), (4096, 6, 8, 2), (33, 48, 16, 0xFE),
                                               (256, 2, 2, 5), (258, 3, 2, 9)])
def test_merkle_matches_oracle(S, n, vpl, arity, label):
    vals = rnd(n + vpl, (n, vpl))
    import ctypes
    root = ctypes.create_string_buffer(32)
    depth = 1
    cur = n
    while cur > arity:
        cur = (cur + arity - 1) // arity
        depth += 1
    O.lib().ref_merkle_root(O.ptr(vals), n, vpl, arity, depth, label, root)
    assert S.merkle_commit(vals, arity, label) == root.raw


# ---------------------------------------------------------------- FRI / STIR
@pytest.mark.parametrize("lg,schedule,s0,cc", [(6, [8, 8], 5, True), (9, [8, 8, 8], 12, True),
                                               (10, [8, 8, 8, 2], 20, True), (11, [4, 8, 4, 8, 2], 9, False),
                                               (13, [16, 8, 8, 8], 30, True), (17, [8, 8, 8, 8, 8, 4], 54, True)])
def test_deep_fri_prove_matches_oracle(S, lg, schedule, s0, cc):
    say("deep_fri_prove n0=2^%d schedule=%s" % (lg, schedule))
    f0 = rnd(400 + lg, 1 << lg)
    pi = hashlib.sha3_256(b"pi%d" % lg).digest()
    ref = O.deep_fri_prove(f0, O.fri_params(schedule, r=s0, s0=s0, coeff_commit_final=cc, pi_hash=pi))
    got = S.deep_fri_prove(f0, S.FriDomain.new_radix2(1 << lg),
                           S.DeepFriParams(schedule, s0, S.SEED_Z, cc, 1, True, s0, pi))
    assert got == ref


# ---------------------------------------------------------------- golden proofs
def _gold():
    return json.load(open(os.path.join(GOLD, "primary_state.json")))


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_inner_shard_golden_on_gpu(S, idx):
    from stark_swarm_amd.zone import build_synthetic_zone, split_evenly
    d = _gold()
    w = d["shards"][idx]
    say("golden inner shard %d" % w["shard_id"])
    salt = bytes.fromhex(d["zone_salt_hex"])
    recs = split_evenly(build_synthetic_zone(d["record_count"]), 3)[idx]
    fsb = S.shard_fs_binding(bytes.fromhex(d["authority_pk_hash_hex"]), bytes.fromhex(d["job_id_hex"]),
                             w["shard_id"], bytes.fromhex(w["shard_nonce_hex"]))
    out = S.prove_inner_shard(salt, recs, fsb, S.LdtMode.Stir)
    assert out.merkle_root.hex() == w["merkle_root_hex"]
    assert out.pi_hash.hex() == w["pi_hash_hex"]
    assert out.n_trace == w["n_trace"]
    assert len(out.proof_blob) == d["blobs"][w["proof"]]["len"]
    assert hashlib.sha3_256(out.proof_blob).hexdigest() == d["blobs"][w["proof"]]["sha3_256"]
    if idx == 0:
        assert out.proof_blob == open(os.path.join(GOLD, w["proof"]), "rb").read()


def test_outer_rollup_golden_on_gpu(S):
    d = _gold()
    say("golden outer rollup")
    pis = [bytes.fromhex(h) for h in d["inner_pi_hashes_hex"]]
    out = S.prove_outer_rollup(pis, bytes.fromhex(d["authority_pk_hash_hex"]), S.LdtMode.Stir)
    assert out.proof_blob == open(os.path.join(GOLD, "outer-rollup.bin"), "rb").read()
    assert out.root_f0.hex() == d["outer_root_f0_hex"]
    assert out.n0 == d["outer_n0"] and out.n_trace == d["outer_n_trace"]


def _oracle_threads():
    import os
    return max(1, min(os.cpu_count() or 1, 16))


def test_prover_core_c2_rows_matches_oracle(S):
    """The C2 shape on 8 of its 128 columns: 2^20 rows, blowup 8, STIR fold
    [8, 8, 8, 8, 8, 4], 80 queries -- the whole proof byte for byte against
    the oracle (run on the host's cores in the reference's threading)"""
    import ctypes
    ctx = S.default_context()
    w, n, blowup, schedule, s0 = 8, 1 << 20, 8, [8, 8, 8, 8, 8, 4], 80
    tr = rnd(520, (w, n))
    params = S.DeepFriParams(schedule, s0, S.SEED_Z, True, 1, True, s0, None)
    d_tr = ctx.alloc(tr.nbytes).upload(tr)
    d_lde = ctx.alloc(tr.nbytes * blowup)
    say("prover core 8 x 2^20, blowup 8 (GPU)")
    root, blob = S.prover_core(d_tr, w, n, blowup, 7, params, d_lde)
    rr = ctypes.create_string_buffer(32)
    pp, pl = O.u8p(), O.sz()
    say("prover core 8 x 2^20, blowup 8 (oracle, %d threads)" % _oracle_threads())
    with O.threads(_oracle_threads()):
        rc = O.lib().ref_prover_core(O.ptr(tr), w, n, blowup, 7, O.fri_params(schedule, r=s0, s0=s0), rr, None, pp, pl)
    assert rc == 0
    assert root == rr.raw
    assert blob == O.take_bytes(pp, pl)


def test_prover_core_c2_full_lde_columns(S):
    """The full C2 shape (2^20 x 128, blowup 8, shift 7): three whole LDE
    columns of the proof's device LDE against the oracle's coset LDE of
    those columns, and the proof repeats byte for byte"""
    ctx = S.default_context()
    w, n, blowup, schedule, s0 = 128, 1 << 20, 8, [8, 8, 8, 8, 8, 4], 80
    tr = rnd(521, (w, n))
    params = S.DeepFriParams(schedule, s0, S.SEED_Z, True, 1, True, s0, None)
    d_tr = ctx.alloc(tr.nbytes).upload(tr)
    d_lde = ctx.alloc(tr.nbytes * blowup)
    say("prover core 128 x 2^20, blowup 8 (GPU, twice)")
    root, blob = S.prover_core(d_tr, w, n, blowup, 7, params, d_lde)
    assert S.prover_core(d_tr, w, n, blowup, 7, params, d_lde) == (root, blob)
    n0 = n * blowup
    cols = [0, 77, 127]
    say("oracle coset LDE of columns %s" % cols)
    want = O.lde_columns(np.ascontiguousarray(tr[cols]), blowup, shift=7)
    for i, c in enumerate(cols):
        assert np.array_equal(d_lde.download((n0,), offset=8 * c * n0), want[i]), c
    d_tr.free()
    d_lde.free()


def test_unknown_ldt_mode_is_an_error(S):
    # LdtMode has two variants (prover.rs:37-41); any other code is refused
    import ctypes
    ctx = S.default_context()
    pp, pl = ctypes.POINTER(ctypes.c_uint8)(), ctypes.c_size_t()
    rf = ctypes.create_string_buffer(32)
    rc = S.lib().stark_prove_outer_rollup(ctx.h, b"\x00" * 32, 1, b"\x01" * 32, 7, 1, ctypes.byref(pp),
                                          ctypes.byref(pl), rf)
    assert rc == -1


# ---------------------------------------------------------------- prover core
@pytest.mark.parametrize("w,lg,blowup,schedule,s0", [(8, 10, 8, [8, 8, 8, 4], 10), (16, 11, 8, [8, 8, 8, 8], 12),
                                                     (3, 8, 4, [8, 8, 2], 6)])
def test_prover_core_matches_oracle(S, w, lg, blowup, schedule, s0):
    say("prover core w=%d n=2^%d blowup=%d" % (w, lg, blowup))
    import ctypes
    ctx = S.default_context()
    n = 1 << lg
    tr = rnd(500 + lg, (w, n))
    params = S.DeepFriParams(schedule, s0, S.SEED_Z, True, 1, True, s0, None)
    d_tr = ctx.alloc(tr.nbytes).upload(tr)
    d_lde = ctx.alloc(tr.nbytes * blowup)
    root, blob = S.prover_core(d_tr, w, n, blowup, 7, params, d_lde)
    rr = ctypes.create_string_buffer(32)
    pp, pl = O.u8p(), O.sz()
    f0 = np.empty(n * blowup, dtype=np.uint64)
    rc = O.lib().ref_prover_core(O.ptr(tr), w, n, blowup, 7,
                                 O.fri_params(schedule, r=s0, s0=s0), rr, O.ptr(f0), pp, pl)
    assert rc == 0
    ref = O.take_bytes(pp, pl)
    assert root == rr.raw
    assert blob == ref
    assert np.array_equal(d_lde.download((w, n * blowup)), O.lde_columns(tr, blowup, shift=7))


# ---------------------------------------------------------------- lockstep batches
def test_golden_shards_as_one_batch(S):
    from stark_swarm_amd.zone import build_synthetic_zone, split_evenly
    d = _gold()
    say("golden shards 1-3 as one lockstep batch")
    salt = bytes.fromhex(d["zone_salt_hex"])
    shards = split_evenly(build_synthetic_zone(d["record_count"]), 3)
    fsbs = [S.shard_fs_binding(bytes.fromhex(d["authority_pk_hash_hex"]), bytes.fromhex(d["job_id_hex"]),
                               w["shard_id"], bytes.fromhex(w["shard_nonce_hex"])) for w in d["shards"]]
    outs = S.prove_inner_shard_batch(salt, shards, fsbs)
    for w, o in zip(d["shards"], outs):
        assert o.pi_hash.hex() == w["pi_hash_hex"]
        assert o.merkle_root.hex() == w["merkle_root_hex"]
        assert hashlib.sha3_256(o.proof_blob).hexdigest() == d["blobs"][w["proof"]]["sha3_256"]


def test_mixed_size_shard_batch_equals_singles(S):
    from stark_swarm_amd.zone import build_synthetic_zone
    say("mixed-size shard batch vs single proofs")
    zone = build_synthetic_zone(3000)
    salt = b"0123456789abcdef"
    sizes = [1024, 700, 1, 2048, 1024, 33, 300]
    shards, fsbs, off = [], [], 0
    for i, s in enumerate(sizes):
        shards.append(zone[off:off + s])
        off = (off + 131 * (i + 1)) % (3000 - 2048)
        fsbs.append(hashlib.sha3_256(b"fsb%d" % i).digest())
    outs = S.prove_inner_shard_batch(salt, shards, fsbs)
    for sh, f, o in zip(shards, fsbs, outs):
        single = S.prove_inner_shard(salt, sh, f)
        assert o.proof_blob == single.proof_blob
        assert o.pi_hash == single.pi_hash and o.n_trace == single.n_trace


@pytest.mark.parametrize("K,lg,schedule", [(5, 9, [8, 8, 8]), (3, 13, [8, 8, 8, 8, 2]), (16, 8, [4, 4, 4, 4])])
def test_fri_batch_matches_oracle(S, K, lg, schedule):
    say("deep_fri_prove_batch K=%d n0=2^%d" % (K, lg))
    f0s = rnd(600 + K + lg, (K, 1 << lg))
    pis = [hashlib.sha3_256(b"pi%d" % k).digest() for k in range(K)]
    got = S.deep_fri_prove_batch(f0s, S.FriDomain.new_radix2(1 << lg),
                                 S.DeepFriParams(schedule, 7, S.SEED_Z, True, 1, True, 7, None), pis)
    for k in range(K):
        ref = O.deep_fri_prove(f0s[k], O.fri_params(schedule, r=7, s0=7, pi_hash=pis[k]))
        assert got[k] == ref, "proof %d" % k


def test_dns_commitments_with_long_records(S):
    """Device leaf hashes / DNS trees against the host restatement and the
    oracle's tree, including records whose message spans several SHA3 blocks."""
    say("dns commitments, long records")
    salt = b"zone-salt-16byte"
    recs = []
    for i in range(37):
        ln = [0, 1, 60, 80, 81, 82, 100, 108, 109, 110, 200, 244, 245, 300, 500][i % 15]
        recs.append(S.DnsRecord.txt("t%d.example.org" % i, 60 + i, "x" * ln))
    recs.append(S.DnsRecord.aaaa("v6.example.org", 300, bytes(range(16))))
    leaves = b"".join(r.leaf_hash(salt) for r in recs)
    ref_root = O.dns_merkle_root(leaves)
    fsb = hashlib.sha3_256(b"long").digest()
    outs = S.prove_inner_shard_batch(salt, [recs, recs[:5], recs[7:8]], [fsb, fsb, fsb])
    assert outs[0].merkle_root == ref_root
    assert outs[1].merkle_root == O.dns_merkle_root(leaves[:5 * 32])
    assert outs[2].merkle_root == leaves[7 * 32:8 * 32]  # a one-leaf tree is its leaf
    # the proof itself against the oracle on the same leaf hashes
    proof, pi, root, rf = O.prove_inner_shard(salt, leaves, fsb, phi_mode=1)
    assert outs[0].proof_blob == proof and outs[0].pi_hash == pi


def test_packed_shards_and_empty_shard(S):
    """A PackedShards batch proves the same bytes as the record lists, twice in
    a row; a shard without records is an error (dns.rs merkle_root unwraps)."""
    from stark_swarm_amd.zone import build_synthetic_zone
    say("packed shards, empty shard")
    salt = b"swarm-test-zone1"
    zone = build_synthetic_zone(300)
    shards = [zone[:40], zone[40:41], zone[41:300]]
    fsbs = [hashlib.sha3_256(bytes([i])).digest() for i in range(3)]
    ref = [o.proof_blob for o in S.prove_inner_shard_batch(salt, shards, fsbs)]
    packed = S.PackedShards(shards)
    for _ in range(2):
        assert [o.proof_blob for o in S.prove_inner_shard_batch(salt, packed, fsbs)] == ref
    with pytest.raises(S.StarkError):
        S.prove_inner_shard_batch(salt, [zone[:3], []], fsbs[:2])
    with pytest.raises(S.StarkError):
        S.prove_inner_shard(salt, [], fsbs[0])


@pytest.mark.parametrize("msg,phi_mode", [(b"abc", 0), (b"abc", 1), (bytes(range(68)), 1), (b"r" * 130, 0)],
                         ids=["1blk-phi0", "1blk-rayon", "2blk-rayon", "3blk-phi0"])
def test_merge_sha256_matches_oracle(S, msg, phi_mode):
    """Device SHA-256 composition + quotient against the oracle's merge."""
    say("merge_sha256 len %d phi %d" % (len(msg), phi_mode))
    tr, nb, _ = O.sha256_trace(msg)
    h = tr.shape[1]
    lde = O.lde_columns(tr, 32)
    want = O.merge_sha256(lde, h, 32, nb, phi_mode)
    got, info = S.deep_ali_merge_sha256(lde, list(range(1, 767)), pow(7, (O.P - 1) // (32 * h), O.P), h, 32, nb,
                                        phi_mode=phi_mode)
    assert np.array_equal(got, want)
    assert info.num_constraints == 766 and info.trace_width == 756


@pytest.mark.parametrize("key,fsb", [(b"\x42" * 44, b"\xAA" * 32), (bytes(range(68)), b"\xBB" * 32)],
                         ids=["ed25519-44B", "p256-68B"])
def test_ds_ksk_binding_matches_oracle(S, key, fsb):
    """The reference's own DS->KSK cases (prover.rs:2946-2975): proof bytes,
    pi_hash and the asserted digest (= SHA-256 of the key)."""
    say("ds_ksk %d bytes" % len(key))
    parent = hashlib.sha256(key).digest()
    out = S.prove_ds_ksk_binding(key, parent, fsb)
    proof, pi, dg, rf, nb, nt = O.prove_ds_ksk_binding(key, parent, fsb)
    assert out.asserted_digest == parent == dg
    assert (out.n_blocks, out.n_trace) == (nb, nt)
    assert out.root_f0 == rf and out.pi_hash == pi
    assert out.proof_blob == proof


def test_ds_ksk_batch_matches_singles(S):
    """Keys of 1, 2 and 3 blocks (three trace heights) in one batch call."""
    say("ds_ksk batch")
    keys = [b"\x42" * 44, bytes(range(68)), b"k" * 10, b"m" * 130, b"n" * 60]
    parents = [hashlib.sha256(k).digest() for k in keys]
    fsbs = [bytes([i]) * 32 for i in range(len(keys))]
    outs = S.prove_ds_ksk_binding_batch(keys, parents, fsbs)
    for k, o in zip(keys, outs):
        single = S.prove_ds_ksk_binding(k, parents[keys.index(k)], fsbs[keys.index(k)])
        assert o.proof_blob == single.proof_blob and o.pi_hash == single.pi_hash
    proof, pi, dg, rf, nb, nt = O.prove_ds_ksk_binding(keys[3], parents[3], fsbs[3])
    assert outs[3].proof_blob == proof and outs[3].pi_hash == pi and nb == 3


@pytest.mark.parametrize("msg,phi_mode", [(b"abc", 0), (b"x" * 112, 1)], ids=["1blk-phi0", "2blk-rayon"])
def test_merge_sha512_matches_oracle(S, msg, phi_mode):
    say("merge_sha512 len %d phi %d" % (len(msg), phi_mode))
    tr, nb, _ = O.sha512_trace(msg)
    h = tr.shape[1]
    lde = O.lde_columns(tr, 32)
    want = O.merge_sha512(lde, h, 32, nb, phi_mode)
    got, info = S.deep_ali_merge_sha512(lde, list(range(1, 1527)), 0, h, 32, nb, phi_mode=phi_mode)
    assert np.array_equal(got, want)


def _oracle_sha_proof(alg, msg, fsb, height=None):
    """The pipeline of prove_ds_ksk_binding restated with oracle pieces: trace
    (padded to `height` by repeating the last useful row) -> LDE x32 -> merge
    -> STIR proof with 54 queries bound to fsb."""
    tr, nb, dg = (O.sha256_trace if alg == 256 else O.sha512_trace)(msg)
    useful = 128 * nb
    h = height or tr.shape[1]
    if h > tr.shape[1]:
        pad = np.repeat(tr[:, useful - 1:useful], h - tr.shape[1], axis=1)
        tr = np.concatenate([tr, pad], axis=1)
    lde = O.lde_columns(tr, 32)
    c = (O.merge_sha256 if alg == 256 else O.merge_sha512)(lde, h, 32, nb, 1)
    n0 = 32 * h
    lg = n0.bit_length() - 1
    sched = [8] * (lg // 3) + ([1 << (lg % 3)] if lg % 3 else [])
    return O.deep_fri_prove(c, O.fri_params(sched, r=54, s0=54, pi_hash=fsb)), dg, nb


def test_sha_batch_matches_oracle(S):
    """prove_sha_batch: SHA-512 and SHA-256 proofs, one with a trace taller
    than its minimum (padding rows), against the oracle pipeline."""
    say("prove_sha_batch")
    msgs512 = [b"abc", b"x" * 112]
    fsb = [hashlib.sha3_256(m).digest() for m in msgs512]
    outs = S.prove_sha_batch(512, msgs512, fsb)
    for m, f, o in zip(msgs512, fsb, outs):
        proof, dg, nb = _oracle_sha_proof(512, m, f)
        assert o.digest == hashlib.sha512(m).digest() == dg and o.n_blocks == nb
        assert o.proof_blob == proof
    msgs256 = [b"k" * 10, b"m" * 70]
    fsb2 = [b"\x01" * 32, b"\x02" * 32]
    outs = S.prove_sha_batch(256, msgs256, fsb2, heights=[256, 256])
    for m, f, o in zip(msgs256, fsb2, outs):
        proof, dg, nb = _oracle_sha_proof(256, m, f, height=256)
        assert o.n_trace == 256 and o.digest == dg and o.proof_blob == proof
    # both hash AIRs and several heights in one call, against the single calls
    algs = [256, 512, 256, 512]
    msgs = [b"a" * 5, b"b" * 200, b"c" * 100, b"d"]
    hs = [512, 256, 256, 128]
    fsb3 = [bytes([i]) * 32 for i in range(4)]
    mixed = S.prove_sha_batch(algs, msgs, fsb3, heights=hs)
    for i in range(4):
        one = S.prove_sha_batch(algs[i], [msgs[i]], [fsb3[i]], heights=[hs[i]])[0]
        assert mixed[i].proof_blob == one.proof_blob and mixed[i].n_trace == hs[i]


def test_sha_tall_traces_match_oracle(S):
    """the provers' merge from the traces at heights whose blowup-2 LDE and
    interpolation take the two-pass transforms (SHA-256 at 2^12 rows, 11
    blocks then padding rows; SHA-512 at 2^11 rows), against the oracle's
    direct evaluation on the blowup-32 LDE"""
    say("prove_sha_batch tall traces")
    m256, m512 = b"q" * 700, b"z" * 300
    f256, f512 = b"\x05" * 32, b"\x06" * 32
    o256 = S.prove_sha_batch(256, [m256], [f256], heights=[4096])[0]
    assert o256.proof_blob == _oracle_sha_proof(256, m256, f256, height=4096)[0]
    o512 = S.prove_sha_batch(512, [m512], [f512], heights=[2048])[0]
    assert o512.proof_blob == _oracle_sha_proof(512, m512, f512, height=2048)[0]


def test_sha256_two_pass_odd_coset_matches_oracle(S):
    """SHA-256 at 2^13 rows: the provers' odd-coset transform and the
    interpolation both take the two-pass NTT, against the oracle"""
    say("prove_sha_batch 2^13 rows")
    m, f = b"q" * 700, b"\x07" * 32
    o = S.prove_sha_batch(256, [m], [f], heights=[8192])[0]
    assert o.proof_blob == _oracle_sha_proof(256, m, f, height=8192)[0]


def test_server_results_equal_standalone(S):
    """Requests admitted by the proving server in a mix (several workers,
    continuous batching, submitters on two threads) give the same bytes as
    each request proven alone."""
    import threading

    from stark_swarm_amd.server import ProvingServer
    from stark_swarm_amd.zone import build_synthetic_zone
    say("proving server mix")
    zone = build_synthetic_zone(1600)
    salt = b"swarm-test-zone1"
    reqs = []
    from stark_swarm_amd.zone import synthetic_ecdsa_records
    from stark_swarm_amd.zone import synthetic_ed25519_signatures
    ecdsa = synthetic_ecdsa_records(4, 0x5EED0009)
    eds = synthetic_ed25519_signatures(3, 0x5EED000A)
    for i in range(21):
        fsb = hashlib.sha3_256(b"req%d" % i).digest()
        if i % 7 == 6:
            reqs.append(("ed25519", eds[i // 7], fsb))
        elif i % 3 == 0:
            reqs.append(("shard", zone[80 * i: 80 * i + 20 + 7 * i], fsb))
        elif i % 4 == 1:
            reqs.append(("rsa", S.RsaStackedRecord(*_rsa_rec(0x900 + i)), fsb))
        elif i % 8 == 2:
            reqs.append(("p256", ecdsa[i // 8], None))
        else:
            key = bytes([(i * 13 + j) & 255 for j in range(20 + 9 * i)])
            reqs.append(("ksk", key, fsb))
    ids = [None] * len(reqs)
    with ProvingServer(workers=3, hbm_budget=16 << 30, max_batch=6) as srv:
        def go(part):
            for j in part:
                kind, x, fsb = reqs[j]
                if kind == "shard":
                    ids[j] = srv.submit_inner_shard(salt, x, fsb)
                elif kind == "rsa":
                    ids[j] = srv.submit_rsa(x, fsb)
                elif kind == "p256":
                    ids[j] = srv.submit_p256(x)
                elif kind == "ed25519":
                    ids[j] = srv.submit_ed25519(*x, fsb, ROOT_ED)
                else:
                    ids[j] = srv.submit_ds_ksk(x, hashlib.sha256(x).digest(), fsb)
        th = [threading.Thread(target=go, args=(range(t, len(reqs), 2),)) for t in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        res = [srv.wait(i, timeout=300) for i in ids]
    assert max(r.batch_size for r in res) > 1, "the mix must have produced shared batches"
    for (kind, x, fsb), r in zip(reqs, res):
        if kind == "shard":
            alone = S.prove_inner_shard(salt, x, fsb)
            assert r.aux == alone.merkle_root
        elif kind == "rsa":
            alone = S.prove_rsa_exp_batch([x], [fsb])[0]
            assert r.aux == bytes(32) and r.kind == 2
        elif kind == "p256":
            alone = S.prove_record_ecdsa_batch([x])[0]
            assert r.kind == 3 and r.n_trace == 256
        elif kind == "ed25519":
            alone = S.prove_zsk_ksk_binding_v2(*x, fsb, ROOT_ED)
            assert r.kind == 4 and r.root_f0 == alone.root_f0
        else:
            alone = S.prove_ds_ksk_binding(x, hashlib.sha256(x).digest(), fsb)
            assert r.aux == alone.asserted_digest
        assert r.proof_blob == alone.proof_blob and r.pi_hash == alone.pi_hash


ROOT_ED = bytes(range(100, 132))


def test_server_bad_request_in_a_batch_fails_alone(S):
    """With the submit-time signature checks off, a batch holding one Ed25519
    request with a wrong signature fails; the server proves its requests
    again one at a time, so only that request raises (with its own error)
    and the others' proofs equal the single-call proofs."""
    from stark_swarm_amd.server import ProvingServer
    from stark_swarm_amd.zone import synthetic_ed25519_signatures
    eds = synthetic_ed25519_signatures(5, 0x5EED000B)
    bad = 3
    pk, sig, msg = eds[bad]
    eds[bad] = (pk, sig[:40] + bytes([sig[40] ^ 1]) + sig[41:], msg)
    say("proving server: one bad Ed25519 request among 5, checks at submit off")
    with ProvingServer(workers=1, hbm_budget=16 << 30, max_batch=8, validate=False) as srv:
        ids = [srv.submit_ed25519(*e, bytes([i]) * 32, ROOT_ED) for i, e in enumerate(eds)]
        out = []
        for i in ids:
            try:
                out.append(srv.wait(i, timeout=300))
            except S.StarkError as e:
                out.append(e)
    assert isinstance(out[bad], S.StarkError) and out[bad].code == -1
    assert "verif" in str(out[bad]).lower()
    for i, (e, r) in enumerate(zip(eds, out)):
        if i == bad:
            continue
        assert not isinstance(r, Exception), r
        assert r.proof_blob == S.prove_zsk_ksk_binding_v2(*e, bytes([i]) * 32, ROOT_ED).proof_blob
    assert max(r.batch_size for i, r in enumerate(out) if i != bad) > 1, "the bad request must have shared a batch"


# ---- RSA-2048 exponentiation AIR ----

def _rsa_rec(seed, bits=2047):
    import random
    rng = random.Random(seed)
    n = (rng.getrandbits(bits - 1) << 1) | 1 | (1 << (bits - 1))
    s = rng.randrange(n)
    return n, s, pow(s, 65537, n)


def _oracle_rsa_proof(recs, pis, n_trace=32):
    """trace of the stacked records -> LDE x32 -> merge -> STIR (54 queries)
    bound to pis, with oracle pieces"""
    tr = O.rsa_exp_trace(recs, n_trace)
    lde = O.lde_columns(tr, 32)
    c = O.merge_rsa(lde, n_trace, 32, len(recs))
    lg = (32 * n_trace).bit_length() - 1
    sched = [8] * (lg // 3) + ([1 << (lg % 3)] if lg % 3 else [])
    return O.deep_fri_prove(c, O.fri_params(sched, r=54, s0=54, pi_hash=pis))


@pytest.mark.parametrize("nrec", [1, 2])
def test_merge_rsa_matches_oracle(S, nrec):
    say("merge_rsa records %d" % nrec)
    recs = [_rsa_rec(0x700 + i, bits=2047 + i) for i in range(nrec)]
    tr = O.rsa_exp_trace(recs)
    lde = O.lde_columns(tr, 32)
    want = O.merge_rsa(lde, 32, 32, nrec)
    got, info = S.deep_ali_merge_rsa_stacked_streaming(lde, list(range(1, 11280 * nrec + 1)), nrec, 0, 32, 32)
    assert np.array_equal(got, want)
    # a trace that is not valid (one carry bit planted) composes the same way
    tr[4721 + 40 * 7 + 2, 5] = 2
    lde = O.lde_columns(tr, 32)
    got, _ = S.deep_ali_merge_rsa_stacked_streaming(lde, list(range(1, 11280 * nrec + 1)), nrec, 0, 32, 32)
    assert np.array_equal(got, O.merge_rsa(lde, 32, 32, nrec))


def test_prove_rsa_matches_oracle(S):
    """single-record proofs (the per-record prover of the mixed-algorithm zone
    example) and a two-record stacked proof, byte for byte against the oracle;
    a three-proof batch against the single calls"""
    say("prove_rsa")
    recs = [_rsa_rec(0x800 + i) for i in range(3)]
    pis = [hashlib.sha3_256(bytes([i])).digest() for i in range(3)]
    outs = S.prove_rsa_exp_batch([S.RsaStackedRecord(*r) for r in recs], pis)
    for r, p, o in zip(recs, pis, outs):
        proof, rf, pi = O.prove_rsa_exp(*r, p)
        assert o.proof_blob == proof and o.root_f0 == rf and o.pi_hash == pi and o.n_trace == 32
    assert outs[0].proof_blob == _oracle_rsa_proof(recs[:1], pis[0])
    one = S.prove_rsa_exp_batch([S.RsaStackedRecord(*recs[1])], [pis[1]])[0]
    assert one.proof_blob == outs[1].proof_blob
    say("prove_rsa stacked")
    st = S.prove_rsa_exp_batch([S.RsaStackedRecord(*r) for r in recs[:2]], [pis[2]], records_per_proof=2)[0]
    assert st.proof_blob == _oracle_rsa_proof(recs[:2], pis[2])
    assert st.pi_hash == hashlib.sha3_256(pis[2] + st.root_f0).digest()


def test_prove_rsa_real_signature(S):
    """an RSASSA-PKCS1-v1_5 SHA-256 signature from a fresh 2048-bit key, bound
    with the example's public-input recipe, at a taller trace (64 rows)"""
    from tests.test_rsa_air import rsa_key
    say("prove_rsa real signature")
    n, d = rsa_key(0xA1)
    msg = b"DNSSEC-RRSIG-V0|legacy.example.com.|A|10.0.0.3|epoch-0|alg-8"
    em_b = S.emsa_pkcs1_v1_5_encode_sha256(hashlib.sha256(msg).digest())
    em = int.from_bytes(em_b, "big")
    s = pow(em, d, n)
    pk = S.rsa_record_pk_hash(n.to_bytes(256, "big"), s.to_bytes(256, "big"), msg, em_b)
    o = S.prove_rsa_exp_batch([S.RsaStackedRecord(n, s, em)], [pk], n_trace=64)[0]
    assert o.n_trace == 64 and o.proof_blob == _oracle_rsa_proof([(n, s, em)], pk, n_trace=64)


def test_prove_rsa_rejects_bad_record(S):
    n, s, em = _rsa_rec(5)
    with pytest.raises(S.StarkError):
        S.prove_rsa_exp_batch([S.RsaStackedRecord(n, n + 1, em)], [bytes(32)])


def test_prove_rsa_stacked_three_and_batch_of_stacked(S):
    """stacked_rsa_n3 records in one proof (rsa2048_stacked_air.rs:125-157
    shape), and two stacked proofs of two records in one batch call"""
    say("prove_rsa stacked 3")
    recs = [_rsa_rec(0xA00 + i, bits=2046 + i) for i in range(3)]
    pis = hashlib.sha3_256(b"stacked3").digest()
    o = S.prove_rsa_exp_batch([S.RsaStackedRecord(*r) for r in recs], [pis], records_per_proof=3)[0]
    assert o.proof_blob == _oracle_rsa_proof(recs, pis)
    say("prove_rsa batch of stacked")
    recs4 = [_rsa_rec(0xB00 + i) for i in range(4)]
    pis2 = [bytes([7]) * 32, bytes([9]) * 32]
    two = S.prove_rsa_exp_batch([S.RsaStackedRecord(*r) for r in recs4], pis2, records_per_proof=2)
    for k in range(2):
        one = S.prove_rsa_exp_batch([S.RsaStackedRecord(*r) for r in recs4[2 * k:2 * k + 2]], [pis2[k]],
                                    records_per_proof=2)[0]
        assert two[k].proof_blob == one.proof_blob


# ---- NSEC3 chain completeness ----

def _nsec3_chain(n, tag=b"NSEC3-DEMO-OWNER"):
    owners = sorted(hashlib.sha3_256(tag + i.to_bytes(8, "little")).digest() for i in range(n))
    return [(owners[i], owners[(i + 1) % n]) for i in range(n)]


@pytest.mark.parametrize("n,phi_mode", [(1, 1), (5, 1), (200, 1), (200, 0), (64, 2)])
def test_prove_nsec3_matches_oracle(S, n, phi_mode):
    say("prove_nsec3 n=%d phi %d" % (n, phi_mode))
    recs = _nsec3_chain(n)
    salt, fsb = b"swarm-test-zone1", hashlib.sha3_256(b"nsec3-%d" % n).digest()
    proof, pi, cr, rf, nt = O.prove_nsec3_completeness(recs, salt, fsb, phi_mode)
    o = S.prove_nsec3_completeness([S.Nsec3Record(a, b) for a, b in recs], salt, fsb, phi_mode=phi_mode)
    assert (o.proof_blob, o.pi_hash, o.chain_root, o.root_f0, o.n_trace, o.record_count) == (proof, pi, cr, rf, nt, n)


def test_prove_nsec3_batch_and_errors(S):
    say("prove_nsec3 batch")
    chains = [_nsec3_chain(n, b"T%d" % n) for n in (3, 40, 9, 300)]
    fsbs = [bytes([i]) * 32 for i in range(4)]
    salt = b"swarm-test-zone2"
    outs = S.prove_nsec3_completeness_batch([[S.Nsec3Record(a, b) for a, b in c] for c in chains], salt, fsbs)
    for c, f, o in zip(chains, fsbs, outs):
        proof, pi, cr, rf, nt = O.prove_nsec3_completeness(c, salt, f)
        assert o.proof_blob == proof and o.pi_hash == pi and o.chain_root == cr
    open_chain = _nsec3_chain(6)
    open_chain[-1] = (open_chain[-1][0], bytes(32))
    with pytest.raises(S.StarkError):
        S.prove_nsec3_completeness([S.Nsec3Record(a, b) for a, b in open_chain], salt, fsbs[0])
    with pytest.raises(S.StarkError):
        S.prove_nsec3_completeness([], salt, fsbs[0])


def test_merge_nsec3_matches_oracle(S):
    say("merge nsec3")
    n = 256
    tr = rnd(77, (8, n))
    lde = O.lde_columns(tr, 32)
    got, info = S.deep_ali_merge_general(lde, [1, 2, 3, 4], S.AirType.Nsec3Chain, pow(7, (P - 1) // (32 * n), P), n, 32)
    assert np.array_equal(got, O.merge_nsec3(lde, n, 32)) and info.max_constraint_degree == 1
