"""CPU pins of the oracle's double-chain ECDSA-P256 AIR (the checker of the
GPU path).

Reference: crates/deep_ali/src/p256_field_air.rs (gadgets and their sizes),
p256_group_air.rs (RCB-2016 double / add layouts), p256_scalar_mul_air.rs
(the step), p256_ecdsa_double_multirow_air.rs (rows, boundary, transitions),
crates/swarm-dns/examples/mixed_algorithm_zone_rollup.rs:193-322 (record
prover). The arithmetic is pinned by Python big integers (curve law, ECDSA),
the sizes by the reference's own constants and layout tests, the constraint
order by planted mistakes at slots computed from the reference's documented
layout, and vanishing by the reference's double_multirow_k4_all_zero /
double_multirow_wrong_rproj_violates_boundary tests.
"""
import hashlib
import random

import numpy as np
import pytest

from tests import _oracle as O

P = 2**256 - 2**224 + 2**192 + 2**96 - 1
N = 0xFFFFFFFF00000000FFFFFFFFFFFFFFFFBCE6FAADA7179E84F3B9CAC2FC632551
B = 0x5AC635D8AA3A93E7B3EBBD55769886BC651D06B0CC53B0F63BCE3C3E27D2604B
GX = 0x6B17D1F2E12C4247F8BCE6E563A440F277037D812DEB33A0F4A13945D898C296
GY = 0x4FE342E2FE1A7F9B8EE7EB4A7C0F9E162BCE33576B315ECECBB6406837BF51F5
GP = 2**64 - 2**32 + 1


def ec_add(p1, p2):
    """affine addition on y^2 = x^3 - 3x + b; None is the identity"""
    if p1 is None:
        return p2
    if p2 is None:
        return p1
    (x1, y1), (x2, y2) = p1, p2
    if x1 == x2 and (y1 + y2) % P == 0:
        return None
    if p1 == p2:
        lam = (3 * x1 * x1 - 3) * pow(2 * y1, -1, P) % P
    else:
        lam = (y2 - y1) * pow(x2 - x1, -1, P) % P
    x3 = (lam * lam - x1 - x2) % P
    return x3, (lam * (x1 - x3) - y1) % P


def ec_mul(k, pt):
    acc = None
    for bit in bin(k)[2:] if k else "":
        acc = ec_add(acc, acc)
        if bit == "1":
            acc = ec_add(acc, pt)
    return acc


def ecdsa_sign(seed, msg):
    """(qx, qy, r, s, digest) from a seeded key and nonce"""
    rng = random.Random(seed)
    d = rng.randrange(1, N)
    q = ec_mul(d, (GX, GY))
    digest = hashlib.sha256(msg).digest()
    e = int.from_bytes(digest, "big") % N
    while True:
        k = rng.randrange(1, N)
        r = ec_mul(k, (GX, GY))[0] % N
        s = pow(k, -1, N) * (e + r * d) % N
        if r and s:
            return q[0], q[1], r, s, digest


def limbs(v):
    return [(v >> (26 * i)) & (2**26 - 1) for i in range(10)]


def read_fe(tr, base, row):
    return sum(int(tr[base + i, row]) << (26 * i) for i in range(10))


def rows_nonzero(tr, n_trace):
    out = []
    for r in range(n_trace):
        cons = O.p256_constraints(tr[:, r], tr[:, (r + 1) % n_trace], r, n_trace)
        out.append(np.nonzero(cons)[0])
    return out


def test_curve_constants():
    assert (GY * GY - (GX**3 - 3 * GX + B)) % P == 0
    assert ec_mul(N, (GX, GY)) is None


def test_sizes_match_reference_constants():
    # p256_field_air.rs tests: 270 / 280 / 290 / 1188 / 271 / 560 cells;
    # 270 / 290 / 311 / 1207 / 281 / 591 constraints
    per_kind = {0: 1207, 1: 290, 2: 311, 3: 591, 4: 281}
    g = O.p256_gadgets(0)
    kinds = [k for k, _ in g]
    dbl, add = kinds[:55], kinds[55:127]
    # group_double_layout_consistency / group_add_layout_consistency
    assert [dbl.count(k) for k in range(4)] == [13, 15, 6, 21]
    assert [add.count(k) for k in range(4)] == [14, 20, 9, 29]
    assert kinds[127:] == [4, 4, 4]
    step = sum(per_kind[k] for k in kinds)
    assert sum(per_kind[k] for k in dbl) == 13 * 1207 + 15 * 290 + 6 * 311 + 21 * 591 == 34318
    assert sum(per_kind[k] for k in add) == 14 * 1207 + 20 * 290 + 9 * 311 + 29 * 591 == 42636
    assert step == 77797
    # ecdsa_double_multirow_constraints = 2 steps + 60 boundary + 120 transitions
    assert O.p256_n_constraints() == 2 * step + 60 + 120 == 155774
    cells_dbl = 270 + 13 * 1188 + 15 * 280 + 6 * 290 + 21 * 560
    cells_add = 270 + 14 * 1188 + 20 * 280 + 9 * 290 + 29 * 560
    chain = 6 * 10 + 1 + cells_dbl + cells_add + 3 * 270
    assert O.p256_width() == 2 * chain + 60 == 151334
    c = O.p256_columns()
    # chain A cells: acc x,y,z, base x,y,z, the bit, then the step gadgets
    assert c["acc"][0] == [0, 10, 20] and c["base"][0] == [30, 40, 50] and c["bit"][0] == 60
    assert c["acc"][1][0] == chain and c["rproj"][0] == 2 * chain
    # the first mul (t0 = X X) of the double sits after the b block
    assert g[0][1][0] == 61 + 270 and g[0][1][1] == 61 + 280


def two_pass(points, a_bits, b_bits, n_trace):
    tr = O.p256_trace(points, [0] * 6, a_bits, b_bits, n_trace)
    c = O.p256_columns()
    last = len(a_bits) - 1
    rp = [read_fe(tr, c["sel"][ch][i], last) for ch in range(2) for i in range(3)]
    return O.p256_trace(points, rp, a_bits, b_bits, n_trace), rp


def k4_points():
    g = (GX, GY)
    q = ec_mul(2, g)
    return [GX, GY, 1, GX, GY, 1, q[0], q[1], 1, q[0], q[1], 1], g, q


def test_double_multirow_k4_all_zero():
    # double_multirow_k4_all_zero (p256_ecdsa_double_multirow_air.rs:339-418)
    pts, _, _ = k4_points()
    tr, _ = two_pass(pts, [1, 0, 1, 0], [0, 1, 1, 0], 4)
    assert all(len(nz) == 0 for nz in rows_nonzero(tr, 4))


def proj_eq(X, Y, Z, pt):
    x, y = pt
    return Z % P != 0 and (X - x * Z) % P == 0 and (Y - y * Z) % P == 0


def test_chain_outputs_follow_group_law():
    # acc' = 2 acc + bit * base from acc_0 = base: after k steps the chain
    # holds (2^k + bits) * base (scalar_mul_chain_2_steps_g_to_7g generalised)
    pts, g, q = k4_points()
    a_bits, b_bits = [1, 0, 1, 1], [0, 1, 1, 0]
    tr, rp = two_pass(pts, a_bits, b_bits, 4)
    c = O.p256_columns()
    for ch, (bits, base) in enumerate(((a_bits, g), (b_bits, q))):
        u = 0
        for r, bit in enumerate(bits):
            u = 2 * u + bit
            X, Y, Z = (read_fe(tr, c["sel"][ch][i], r) for i in range(3))
            assert proj_eq(X, Y, Z, ec_mul(2 ** (r + 1) + u, base))
            assert all(read_fe(tr, c["sel"][ch][i], r) < P for i in range(3))
        assert rp[3 * ch:3 * ch + 3] == [read_fe(tr, c["sel"][ch][i], 3) for i in range(3)]


def test_wrong_rproj_fires_only_boundary():
    # double_multirow_wrong_rproj_violates_boundary: a bogus r_proj fires the
    # boundary slots at the last row and nothing else
    pts, _, _ = k4_points()
    bits = [1, 0, 1, 0]
    tr, rp = two_pass(pts, bits, bits, 4)
    bogus = list(rp)
    bogus[4] ^= 1 << 30  # r_b y, limb 1
    tr2 = O.p256_trace(pts, bogus, bits, bits, 4)
    nz = rows_nonzero(tr2, 4)
    assert [len(v) for v in nz[:3]] == [0, 0, 0]
    assert list(nz[3]) == [2 * 77797 + 40 + 1]


def row_fired(tr, n_trace, r):
    """{slot: value} of the constraints that do not vanish on row r"""
    cons = O.p256_constraints(tr[:, r], tr[:, (r + 1) % n_trace], r, n_trace)
    return {int(i): int(cons[i]) for i in np.nonzero(cons)[0]}


RADIX = 1 << 26
P_LIMBS = [(P >> (26 * k)) & (RADIX - 1) for k in range(10)]  # P_LIMBS_TIGHT


def plant(tr, col, row, value):
    t = tr.copy()
    t[col, row] = value
    return t


def test_planted_mistakes_fire_documented_slots():
    """Each planted cell fires the slots its gadget's formulas put on it, with
    the values those formulas give (p256_field_air.rs: booleanity
    b (1 - b), limb pack c - sum 2^b bit (205-224, 404-451), add chain
    a + b + carry_in - c - 2^26 carry and carry[9] (425-451), mul position
    identities with carries biased by 2^35 (1044-1120), select c - b -
    sel (a - b) (1243-1272), freeze chain a - p - diff + net_in - 2^26 net_out
    and mux c - diff - neg_9 (a - diff) (1449-1530); the row transition
    nxt.acc - cur.select_c, p256_ecdsa_double_multirow_air.rs:274-289)."""
    pts, _, _ = k4_points()
    bits = [1, 1, 0, 1]
    tr, _ = two_pass(pts, bits, bits, 4)
    g = O.p256_gadgets(0)
    row = 1
    G = O.P
    # (1) a c bit of the first mul of chain A's double (slot base 0): its
    # booleanity slot 26 i + b (2 (1 - 2)) and its limb packing slot 260 + i
    col = g[0][1][1] + 26 * 3 + 5
    old = int(tr[col, row])
    t = plant(tr, col, row, 2)
    assert row_fired(t, 4, row) == {26 * 3 + 5: G - 2, 260 + 3: (-(2 - old) * 32) % G}
    # (2) the last carry of the first add (step 5, after the 13 muls): chain
    # identity k = 9 (slot 270 + 9: -2^26) and the "carry[9] = 0" slot 289 (1)
    base = 13 * 1207
    assert int(tr[g[13][1][4] + 9, row]) == 0
    t = plant(tr, g[13][1][4] + 9, row, 1)
    assert row_fired(t, 4, row) == {base + 279: (-RADIX) % G, base + 289: 1}
    # (3) a carry bit of the last mul of the add gadget (index 55 + 13 in
    # eval order) := 5: booleanity 5 (1 - 5) at 540 + 19 + 36 k + b, the signed
    # carry at position k moves by (5 - bit) 2^b: -2^26 of it at position k,
    # + it at position k + 1
    gi = 55 + 13
    base = 34318 + 13 * 1207
    col = g[gi][1][4] + 36 * 7 + 3
    d = (5 - int(tr[col, row])) << 3
    t = plant(tr, col, row, 5)
    assert row_fired(t, 4, row) == {base + 540 + 7: (-RADIX * d) % G, base + 540 + 8: d % G,
                                    base + 559 + 36 * 7 + 3: (5 * (1 - 5)) % G}
    # (4) chain B's z select output, limb 0, at a row with a successor: its
    # packing slot (+d), its select identity (+d), and the transition into
    # the next row (nxt.acc - cur.c: -d)
    gb = O.p256_gadgets(1)
    base = 77797 + 77797 - 281
    col = gb[-1][1][0]
    old = int(tr[col, row])
    d = (old ^ 1) - old
    t = plant(tr, col, row, old ^ 1)
    assert row_fired(t, 4, row) == {base + 260: d % G, base + 270: d % G, 2 * 77797 + 60 + 50: (-d) % G}
    # (5) the neg carry of limb 9 in the first freeze after a sub of chain A's
    # add gadget -- freeze slots after 14 muls, 20 adds, 9 subs and 20
    # freezes-after-adds: the chain identity k = 9 moves by 2^26 (neg_9' -
    # neg_9) and each mux slot k by (neg_9 - neg_9')(a_k - diff_k), a_k read
    # back from the valid chain identity
    fz = [i for i, (k, _) in enumerate(g) if k == 3 and i >= 55][20]
    base = 34318 + 14 * 1207 + 20 * 290 + 9 * 311 + 20 * 591
    cd = g[fz][1]
    dl, pos, neg = cd[2], cd[4], cd[5]
    net = [int(tr[pos + k, row]) - int(tr[neg + k, row]) for k in range(10)]
    a = [P_LIMBS[k] + int(tr[dl + k, row]) - (net[k - 1] if k else 0) + RADIX * net[k] for k in range(10)]
    n_old = int(tr[neg + 9, row])
    dn = (1 - n_old) - n_old
    t = plant(tr, neg + 9, row, 1 - n_old)
    want = {base + 540 + 9: (RADIX * dn) % G}
    for k in range(10):
        v = -dn * (a[k] - int(tr[dl + k, row]))
        if v % G:
            want[base + 581 + k] = v % G
    assert row_fired(t, 4, row) == want


def test_ecdsa_bits_and_native_verify():
    qx, qy, r, s, digest = ecdsa_sign(7, b"DNSSEC-RRSIG-V0|a.example|A|1.2.3.4|epoch-0|alg-13")
    u1b, u2b = O.p256_ecdsa_bits(qx, qy, r, s, digest)
    w = pow(s, -1, N)
    e = int.from_bytes(digest, "big") % N
    u1, u2 = e * w % N, r * w % N
    assert u1b == [int(ch) for ch in format(u1, "0256b")]
    assert u2b == [int(ch) for ch in format(u2, "0256b")]
    assert O.p256_ecdsa_bits(qx, qy, r, (s + 1) % N, digest) == -30
    assert O.p256_ecdsa_bits(qx, (qy + 1) % P, r, s, digest) == -31
    assert O.p256_ecdsa_bits(qx, qy, 0, s, digest) == -30


@pytest.mark.parametrize("seed", [3])
def test_ecdsa_trace_vanishes_and_binds_the_scalar_multiples(seed):
    qx, qy, r, s, digest = ecdsa_sign(seed, b"record %d" % seed)
    tr, rp = O.p256_ecdsa_trace(qx, qy, r, s, digest)
    w = pow(s, -1, N)
    e = int.from_bytes(digest, "big") % N
    u1, u2 = e * w % N, r * w % N
    # chain A ends at (2^256 + u1) G, chain B at (2^256 + u2) Q
    assert proj_eq(*rp[0:3], ec_mul(2**256 + u1, (GX, GY)))
    assert proj_eq(*rp[3:6], ec_mul(2**256 + u2, (qx, qy)))
    # every constraint vanishes on a sample of rows incl. the last two
    for row in (0, 1, 128, 254, 255):
        cons = O.p256_constraints(tr[:, row], tr[:, (row + 1) % 256], row, 256)
        assert not cons.any(), row


def test_merge_quotient_identity_small():
    # the merge over a K = 4 trace: phi at every point, then the Z_H quotient
    # (pinned code); on a satisfying trace phi vanishes on the trace domain
    pts, _, _ = k4_points()
    tr, _ = two_pass(pts, [1, 0, 0, 1], [1, 1, 0, 1], 4)
    lde = np.zeros((tr.shape[0], 4 * 8), dtype=np.uint64)
    O.lib().ref_lde_columns(O.ptr(np.ascontiguousarray(tr)), tr.shape[0], 4, 8, 1, O.ptr(lde))
    c = O.merge_p256(lde, 4, 8)
    assert c.shape == (32,) and c.any()
    # lambda = 0 everywhere gives phi = 0 and c = 0
    z = O.merge_p256(lde, 4, 8, np.zeros(O.p256_n_constraints(), dtype=np.uint64))
    assert not z.any()
