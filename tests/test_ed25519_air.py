"""CPU pins of the oracle's Ed25519 cofactored-verify AIR, v16 (the checker of
the GPU path).

Reference: crates/deep_ali/src/ed25519_verify_air.rs (v0..v16 layout, fill and
per-row evaluator), ed25519_field_air.rs / ed25519_group_air.rs /
ed25519_scalar_air.rs / ed25519_scalar_mult_air.rs (gadgets),
crates/swarm-dns/src/prover.rs:826-951 (prove_zsk_ksk_binding_v2).

No Ed25519 proof or c_eval is stored in the reference, so the oracle is pinned
by (1) RFC 8032 section 7.1 vectors read through the trace against Python
big-integer Ed25519 (tests/ed25519_py.py): the SHA-512 digest bits, k mod L in
the reduce gadget, the decompressed R and A, every ladder accumulator, the sB /
kA / R threads and the identity verdict; (2) all 70 606 constraints vanishing
on every row of valid traces (the reference's v16 end-to-end tests); (3)
planted mistakes whose slot and value follow from the reference formulas
(booleanity b(1-b), packings, the mul carry chain with its 2^w radix and 19
wrap, the sub chain's signed carries), so a flipped sign or a moved slot fails.
"""
import hashlib

import numpy as np
import pytest

from tests import _oracle as O
from tests import ed25519_py as E

GP = 2**64 - 2**32 + 1
LW = [26, 25, 26, 25, 26, 25, 26, 25, 26, 25]
LOFF = [0, 26, 51, 77, 102, 128, 153, 179, 204, 230]

# constraint-slot offsets of the v16 vector (ed25519_verify_air.rs: verify_v*_per_row_constraints)
SHA_K, REDUCE_K, FMT_K = 1526, 1674, 560
MUL_K, SUB_K, ADD_K = 265 + 10 + 360, 265 + 40, 265 + 20
PADD_K = 10 + 4 * SUB_K + 5 * ADD_K + 9 * MUL_K
DBL_K = 8 * MUL_K + 4 * ADD_K + 2 * SUB_K
DECOMP_K = 10 + 265 + 4 * MUL_K + 3 * SUB_K + 10 + 2
SMULT_K = DBL_K + 4 * 11 + PADD_K
OFF_FMT = SHA_K + REDUCE_K
OFF_DEC_R = OFF_FMT + FMT_K
OFF_SB = OFF_DEC_R + 2 * DECOMP_K


def off_v10(k):
    return OFF_SB + 2 * (SMULT_K + 80 + 40) + 80 + 22 + 7 * k


def limbs(v):
    return [(v >> LOFF[i]) & ((1 << LW[i]) - 1) for i in range(10)]


def from_limbs(cells):
    return sum(int(c) << LOFF[i] for i, c in enumerate(cells))


def fm(x):
    return x % GP


@pytest.fixture(scope="module")
def rfc1():
    seed, pk, msg, sig = E.rfc_vector(0)
    e = O.Ed25519(pk, sig, msg, 256)
    return e, e.trace(), pk, sig, msg


def test_python_reference_signs_rfc_vectors():
    for i in range(3):
        seed, pk, msg, sig = E.rfc_vector(i)
        assert E.public_key(seed) == pk
        assert E.sign(seed, msg) == (pk, sig)
        assert E.verify(pk, sig, msg)


def test_native_verify_matches_rfc_and_rejects_tampering():
    for i in range(3):
        _, pk, msg, sig = E.rfc_vector(i)
        assert O.ed25519_verify(pk, sig, msg)
        bad = bytearray(sig)
        bad[40] ^= 1  # the reference's corrupted_signature test flips a bit of S
        assert not O.ed25519_verify(pk, bytes(bad), msg)
        assert not E.verify(pk, bytes(bad), msg)
        assert not O.ed25519_verify(pk, sig, msg + b"x")
    for n in (0, 1, 111, 112, 127, 128, 300):
        m = bytes(range(256)) * 2
        assert O.sha512(m[:n]) == hashlib.sha512(m[:n]).digest()


def test_sizes_match_reference_constants():
    # verify_v16_per_row_constraints(256) and the width of the v16 layout
    assert O.ed25519_v16_constraints(256) == 70606
    assert O.ed25519_v16_constraints(8) == 70606 - 7 * 248
    _, pk, msg, sig = E.rfc_vector(0)
    e = O.Ed25519(pk, sig, msg, 256)
    assert (e.width, e.height) == (41519, 1024)
    assert e.rows == {"k_hash": 128, "reduce": 128, "dec_r": 129, "dec_a": 130, "sb_first": 131, "ka_first": 387,
                      "residual": 643, "dbl1": 644, "dbl2": 645, "dbl3": 646, "result": 647}
    e8 = O.Ed25519(pk, sig, msg, 8)
    assert e8.height == 256 and e8.width == 41519 - 3 * 248


def test_trace_reads_rfc_vector(rfc1):
    e, tr, pk, sig, msg = rfc1
    R, c = e.rows, e.cols
    dg = hashlib.sha512(sig[:32] + pk + msg).digest()
    pub_dg, kc, s_bits, k_bits = e.public()
    assert pub_dg == dg
    k = int.from_bytes(dg, "little") % E.L
    assert int.from_bytes(kc, "little") == k
    # digest bits at row k_hash - 1: bit i of the big-endian word w
    row = R["k_hash"] - 1
    for w in range(8):
        word = int.from_bytes(dg[8 * w:8 * w + 8], "big")
        assert [int(tr[c["digest_bits"] + 64 * w + i, row]) for i in range(64)] == [(word >> i) & 1 for i in range(64)]
    # reduce gadget: input limbs, r = digest mod L, q = digest div L (16-bit limbs)
    D = int.from_bytes(dg, "little")
    rr = R["reduce"]
    assert [int(tr[i, rr]) for i in range(32)] == [(D >> (16 * i)) & 0xFFFF for i in range(32)]
    rbits = [int(tr[c["reduce_r_bits"] + i, rr]) for i in range(256)]
    assert sum(b << i for i, b in enumerate(rbits)) == k
    assert sum(int(tr[c["reduce_q"] + i, rr]) << (16 * i) for i in range(17)) == D // E.L
    # decompressed R and A, signs
    for key, comp in (("dec_r", sig[:32]), ("dec_a", pk)):
        x, y = E.decompress(comp)
        assert from_limbs(tr[c["dec_x"]:c["dec_x"] + 10, R[key]]) == x
        assert from_limbs(tr[0:10, R[key]]) == y
        assert int(tr[c["dec_sign"], R[key]]) == comp[31] >> 7 == x & 1
    # the ladders: accumulator at row first + j = (s >> (256 - j)) B, base = B / A
    s = int.from_bytes(sig[32:], "little")
    A = E.decompress(pk)
    for first, scal, base in ((R["sb_first"], s, E.B), (R["ka_first"], k, A)):
        for j in (0, 1, 2, 17, 128, 255):
            xyz = [from_limbs(tr[c["acc_x"] + 10 * q:c["acc_x"] + 10 * q + 10, first + j]) for q in range(3)]
            assert E.proj_eq(xyz, E.mul(scal >> (256 - j), base))
        assert [int(v) for v in tr[c["sblk_sb" if base == E.B else "sblk_ka"]:, first][:4]] == \
            [(scal >> (255 - i)) & 1 for i in range(4)]
    # threads on every row: sB, kA, R (X, Y, Z = 1, T = XY)
    for r in (0, 500, 1023):
        sb = [from_limbs(tr[c["th_sb_x"] + 10 * q:c["th_sb_x"] + 10 * q + 10, r]) for q in range(4)]
        ka = [from_limbs(tr[c["th_ka_x"] + 10 * q:c["th_ka_x"] + 10 * q + 10, r]) for q in range(4)]
        assert E.proj_eq(sb, E.mul(s, E.B)) and sb[3] * sb[2] % E.P == sb[0] * sb[1] % E.P
        assert E.proj_eq(ka, E.mul(k, A))
        rx, ry = E.decompress(sig[:32])
        assert from_limbs(tr[c["th_r_x"]:c["th_r_x"] + 10, r]) == rx
        assert from_limbs(tr[c["th_r_t"]:c["th_r_t"] + 10, r]) == rx * ry % E.P
        assert [int(v) for v in tr[c["r_thread"]:c["r_thread"] + 256, r]] == [(k >> (255 - i)) & 1 for i in range(256)]
    # identity verdict at the result row: X = 0, T = 0, Y = Z != 0
    rs = R["result"]
    X, Y, Z, T = [from_limbs(tr[c[n]:c[n] + 10, rs]) for n in ("res_x", "res_y", "res_z", "res_t")]
    assert X == 0 and T == 0 and Y == Z != 0
    pts = e.points()
    assert E.proj_eq(pts["res2"][:3], E.add(E.add(E.mul(s, E.B), E.neg(E.decompress(sig[:32]))), E.neg(E.mul(k, A))))


@pytest.mark.parametrize("vec", [0, 1, 2])
def test_all_constraints_vanish_on_valid_trace(vec):
    _, pk, msg, sig = E.rfc_vector(vec)
    e = O.Ed25519(pk, sig, msg, 256)
    tr = e.trace()
    h = e.height
    for r in range(h):
        v = e.constraints(tr[:, r], tr[:, (r + 1) % h], r)
        assert not v.any(), (r, np.nonzero(v)[0][:8])


def test_invalid_signature_fires_only_the_verdict(rfc1):
    """a scalar that does not verify: all gadgets still hold, the identity checks at the result row fire"""
    _, pk, msg, sig = E.rfc_vector(0)
    bad = bytearray(sig)
    bad[40] ^= 1
    e = O.Ed25519(pk, bytes(bad), msg, 256)
    tr = e.trace()
    base = off_v10(256) + 10 + 3 * SUB_K
    for r in range(e.height):
        v = e.constraints(tr[:, r], tr[:, (r + 1) % e.height], r)
        nz = set(np.nonzero(v)[0].tolist())
        if r == e.rows["result"]:
            assert nz and nz <= set(range(base, base + 30))
        else:
            assert not nz, r


def planted(e, tr, row, col, value):
    cur = tr[:, row].copy()
    old = int(cur[col])
    cur[col] = value
    v = e.constraints(cur, tr[:, (row + 1) % e.height], row)
    return old, {int(i): int(v[i]) for i in np.nonzero(v)[0]}


def test_planted_digest_bit(rfc1):
    """digest bit b -> 2: booleanity b(1-b) = -2; the H-state pack and the input-limb bind move by the bit weight"""
    e, tr, *_ = rfc1
    row = e.rows["k_hash"] - 1
    for w, i in ((0, 0), (3, 37), (7, 63)):
        old, nz = planted(e, tr, row, e.cols["digest_bits"] + 64 * w + i, 2)
        delta = 2 - old
        m = next(m for m in range(4) if 56 - 16 * m <= i < 64 - 16 * m or 48 - 16 * m <= i < 56 - 16 * m)
        wt = i - (56 - 16 * m) if 56 - 16 * m <= i < 64 - 16 * m else 8 + i - (48 - 16 * m)
        assert nz == {OFF_FMT + 64 * w + i: fm(-2),
                      OFF_FMT + 512 + 2 * w + i // 32: fm(-(delta << (i % 32))),
                      OFF_FMT + 528 + 4 * w + m: fm(-(delta << wt))}


def test_planted_mul_carry_bit(rfc1):
    """ladder doubling, first mul (X^2): flipping bit b of carry k moves identity k by -2^w_k d and
    identity k+1 by +d (identity 0 by +19 d for k = 9), d = +-2^b"""
    e, tr, *_ = rfc1
    row = e.rows["sb_first"] + 5
    ident = OFF_SB + 265
    for k, b in ((3, 7), (9, 0), (0, 35)):
        col = e.cols["dbl_sqa_carry_bits"] + 36 * k + b
        old = int(tr[col, row])
        _, nz = planted(e, tr, row, col, 1 - old)
        d = (1 - 2 * old) << b
        want = {ident + k: fm(-(d << LW[k])), ident + (k + 1) % 10: fm(19 * d if k == 9 else d)}
        assert nz == want


def test_planted_sub_signed_carry(rfc1):
    """result row, sub X - 0: a borrow flag neg_4 = 1 moves identity 4 by +2^26 * (-1) and identity 5
    by +1; with pos_4 = 1 as well the net carry is 0 again and only the exclusion pos * neg = 1 fires"""
    e, tr, *_ = rfc1
    row = e.rows["result"]
    off = off_v10(256) + 10  # sub_X slots
    pos_col = e.cols["sub_x_c"] + 265
    neg_col = pos_col + 10
    old, nz = planted(e, tr, row, neg_col + 4, 1)
    assert old == 0 and int(tr[pos_col + 4, row]) == 0
    assert nz == {off + 265 + 4: fm(-(1 << 26)), off + 265 + 5: 1}
    cur = tr[:, row].copy()
    cur[neg_col + 4] = 1
    cur[pos_col + 4] = 1
    v = e.constraints(cur, tr[:, row + 1], row)
    assert {int(i): int(v[i]) for i in np.nonzero(v)[0]} == {off + 265 + 30 + 4: 1}


def test_planted_cond_bit_and_shift_chain(rfc1):
    """cond-add bit vs the scalar block head, and the block's shift transition (v6)"""
    e, tr, pk, sig, msg = rfc1
    k = 256
    first = e.rows["sb_first"]
    row = first + 10
    old, nz = planted(e, tr, row, e.cols["sblk_sb"], 2)
    v6 = OFF_SB + 2 * (SMULT_K + 80 + 40) + 80 + 22
    # slot k: cond bit - block[0]; the transition slots of the previous row are not on this row
    assert nz[v6 + k] == fm(int(tr[e.cols["cadd_bit"], row]) - 2)
    assert set(nz) == {v6 + k}
    # the previous row's shift transition nxt[0] - cur[1] sees it
    cur, nxt = tr[:, row - 1], tr[:, row].copy()
    nxt[e.cols["sblk_sb"]] = 2
    v = e.constraints(cur, nxt, row - 1)
    assert {int(i): int(v[i]) for i in np.nonzero(v)[0]} == {v6 + k + 1: fm(2 - int(tr[e.cols["sblk_sb"] + 1, row - 1]))}


def test_planted_reduce_and_thread(rfc1):
    """r thread: r_thread[i] - r_bit[255 - i] at the reduce row (v7); the ungated constancy of the thread"""
    e, tr, *_ = rfc1
    k = 256
    v7 = OFF_SB + 2 * (SMULT_K + 80 + 40) + 80 + 22 + 4 * k
    row = e.rows["reduce"]
    old, nz = planted(e, tr, row, e.cols["r_thread"] + 3, 5)
    assert nz == {v7 + 3: fm(5 - old), v7 + k + 3: fm(old - 5)}
    # at a plain row only the constancy slot of the previous transition and of this one see a changed cell
    old, nz = planted(e, tr, 900, e.cols["r_thread"] + 3, 7)
    assert nz == {v7 + k + 3: fm(old - 7)}
