"""RSA-2048 exponentiation AIR (crates/deep_ali/src/rsa2048_exp_air.rs,
rsa2048_field_air.rs, rsa2048_stacked_air.rs) on the CPU: the oracle's trace
builder and 11 280-constraint evaluator.

No stored RSA proof exists in the reference, so the oracle is pinned by
what the mathematics fixes: every row's product, quotient and remainder
against Python big integers, s^65537 mod n at the output row, vanishing of
every constraint on valid traces (single and stacked, as the reference's
own tests state), the output-row binding firing for a wrong EM (the
reference's exp_chain_wrong_em_violates_boundary), and planted mistakes
firing exactly the slots that the layout's constraint order assigns.
"""
import hashlib
import random

import numpy as np
import pytest

from tests import _oracle as O

P = O.P
NL, LB = 80, 26
# columns of one record
C_ACC, C_S, C_N, C_PHASE, C_OP2, C_EM = 0, 80, 160, 240, 241, 321
C_C, C_Q, C_QBITS, C_CBITS, C_CARRY, W = 401, 481, 561, 2641, 4721, 11041
# constraint slots of one record
K_PHASE, K_SEL, K_CBOOL, K_CPACK, K_QBOOL, K_QPACK = 0, 1, 81, 2161, 2241, 4321
K_POS, K_CARRY, K_EM, K_TRANS, K = 4401, 4560, 10880, 10960, 11280


def rand_record(seed, bits=2047):
    """(n, s, em) as in the reference's tests: n odd with `bits` bits, s < n,
    em = s^65537 mod n"""
    rng = random.Random(seed)
    n = (rng.getrandbits(bits - 1) << 1) | 1 | (1 << (bits - 1))
    s = rng.randrange(n)
    return n, s, pow(s, 65537, n)


def limbs_value(col_vals):
    return sum(int(v) << (LB * i) for i, v in enumerate(col_vals))


def nonzero_slots(tr, row, base=0, n_trace=None):
    return set(fired(tr, row, base, n_trace))


def fired(tr, row, base=0, n_trace=None):
    """{slot: value} of the record's constraints that do not vanish on `row`"""
    h = tr.shape[1]
    n_trace = n_trace or h
    c = O.rsa_exp_constraints(tr[base:base + W, row], tr[base:base + W, (row + 1) % h], row, n_trace)
    return {int(i): int(c[i]) for i in np.nonzero(c)[0]}


def test_sizes():
    assert O.lib().ref_rsa_exp_width() == W
    assert O.lib().ref_rsa_exp_constraints_per_record() == K
    # rsa2048_field_air.rs:80-86 and rsa2048_exp_air.rs:80-86
    assert 2 * NL * 27 + 159 + 158 * 40 == 10799
    assert 1 + NL + 10799 + NL + 4 * NL == K


@pytest.mark.parametrize("seed", [0xDEADBEEF, 7])
def test_rows_are_modular_products(seed):
    n, s, em = rand_record(seed)
    tr = O.rsa_exp_trace([(n, s, em)])
    acc = s
    for r in range(32):
        assert limbs_value(tr[C_N:C_N + NL, r]) == n
        assert limbs_value(tr[C_S:C_S + NL, r]) == s
        assert limbs_value(tr[C_EM:C_EM + NL, r]) == em
        assert limbs_value(tr[C_ACC:C_ACC + NL, r]) == acc
        op2 = s if r == 16 else acc
        assert int(tr[C_PHASE, r]) == (r == 16)
        assert limbs_value(tr[C_OP2:C_OP2 + NL, r]) == op2
        c = limbs_value(tr[C_C:C_C + NL, r])
        q = limbs_value(tr[C_Q:C_Q + NL, r])
        assert (c, q) == ((acc * op2) % n, (acc * op2) // n)
        for i in range(NL):
            assert int(tr[C_C + i, r]) == sum(int(tr[C_CBITS + LB * i + b, r]) << b for b in range(LB))
            assert int(tr[C_Q + i, r]) == sum(int(tr[C_QBITS + LB * i + b, r]) << b for b in range(LB))
        acc = c
        if r == 16:
            assert c == pow(s, 65537, n) == em
    assert tr.max() <= (1 << LB) - 1


def test_carries_close_the_product():
    """the biased carry bits make a*b - q*n - c vanish position by position"""
    n, s, em = rand_record(11)
    tr = O.rsa_exp_trace([(n, s, em)])
    r = 3
    a = [int(v) for v in tr[C_ACC:C_ACC + NL, r]]
    b = [int(v) for v in tr[C_OP2:C_OP2 + NL, r]]
    q = [int(v) for v in tr[C_Q:C_Q + NL, r]]
    nn = [int(v) for v in tr[C_N:C_N + NL, r]]
    c = [int(v) for v in tr[C_C:C_C + NL, r]]
    carry = [sum(int(tr[C_CARRY + 40 * k + bb, r]) << bb for bb in range(40)) - (1 << 39) for k in range(158)]
    for k in range(159):
        lhs = sum(a[i] * b[k - i] - q[i] * nn[k - i] for i in range(max(0, k - 79), min(79, k) + 1))
        lhs += (-c[k] if k < NL else 0) + (carry[k - 1] if k else 0) - ((carry[k] << LB) if k < 158 else 0)
        assert lhs == 0


@pytest.mark.parametrize("seed", [0xDEADBEEF, 0xCAFEF00D])
def test_constraints_vanish(seed):
    """exp_chain_computes_s_to_65537 (rsa2048_exp_air.rs:325-356): no
    constraint fires on any row, the wrap row included (transitions are gated
    there)"""
    tr = O.rsa_exp_trace([rand_record(seed)])
    for r in range(32):
        assert nonzero_slots(tr, r) == set(), r


def test_stacked_records_vanish():
    """stacked_rsa_n3_all_zero (rsa2048_stacked_air.rs:125-157)"""
    recs = [rand_record(0xDEAD + i) for i in range(3)]
    tr = O.rsa_exp_trace(recs)
    assert tr.shape == (3 * W, 32)
    for r in range(32):
        for k in range(3):
            assert nonzero_slots(tr, r, base=k * W) == set()


def test_wrong_em_fires_output_binding_only():
    """exp_chain_wrong_em_violates_boundary (rsa2048_exp_air.rs:358-387):
    only the EM slots of the limbs that differ fire, and only at row 16"""
    n, s, em = rand_record(0xCAFEF00D)
    bogus = (em + 1) % n
    tr = O.rsa_exp_trace([(n, s, bogus)])
    diff = {i for i in range(NL) if ((em >> (LB * i)) & (2**LB - 1)) != ((bogus >> (LB * i)) & (2**LB - 1))}
    assert diff
    for r in range(32):
        assert nonzero_slots(tr, r) == ({K_EM + i for i in diff} if r == 16 else set())


def test_small_values_and_short_modulus():
    """a one-limb modulus and s = 0 (mul_gadget_satisfies_on_zero_a)"""
    n = 0xFFFFFFFF * 0xFFFFFFFE + 3
    for s in (0, 1, 0xCAFEBABE % n):
        tr = O.rsa_exp_trace([(n, s, pow(s, 65537, n))])
        for r in range(32):
            assert nonzero_slots(tr, r) == set()


def test_rejects_s_not_below_n():
    n = (1 << 2047) | 1
    with pytest.raises(ValueError):
        O.rsa_exp_trace([(n, n, 0)])
    with pytest.raises(ValueError):
        O.rsa_exp_trace([(0, 0, 0)])


# ---- planted mistakes fix the slot order ----

@pytest.fixture(scope="module")
def good():
    return O.rsa_exp_trace([rand_record(0x5EED)])


# Values: the reference's formulas (rsa2048_field_air.rs:345-415 -- bit
# booleanity b (1 - b), limb pack limb - sum 2^b bit, position identity
# p_ab - p_qn - c_k + carry_in - 2^26 carry_out with biased carries, carry
# booleanity; rsa2048_exp_air.rs:210-272 -- phase booleanity, select
# op - (phase (s - acc) + acc), transitions nxt - cur) at the planted cell,
# every other cell valid.

def test_planted_c_bit(good):
    tr = good.copy()
    i, b, r = 3, 5, 2
    old = int(tr[C_CBITS + LB * i + b, r])
    tr[C_CBITS + LB * i + b, r] = 2
    assert fired(tr, r) == {K_CBOOL + LB * i + b: P - 2, K_CPACK + i: (-(2 - old) * (1 << b)) % P}


def test_planted_q_bit_flip(good):
    tr = good.copy()
    i, b, r = 7, 0, 5
    old = int(tr[C_QBITS + LB * i + b, r])
    tr[C_QBITS + LB * i + b, r] ^= 1
    assert fired(tr, r) == {K_QPACK + i: (-((old ^ 1) - old)) % P}


def test_planted_carry_bit(good):
    tr = good.copy()
    m, b, r = 10, 3, 1
    old = int(tr[C_CARRY + 40 * m + b, r])
    tr[C_CARRY + 40 * m + b, r] = 2
    d = (2 - old) << b  # change of the signed carry at position m
    assert fired(tr, r) == {K_CARRY + 40 * m + b: P - 2,      # booleanity
                            K_POS + m: (-(1 << LB) * d) % P,   # - 2^26 carry_out
                            K_POS + m + 1: d % P}              # + carry_in


def test_planted_c_limb(good):
    """a changed output limb breaks its pack (+1), its position identity
    (-c_k: -1) and the acc transition into the next row (nxt.acc - c: -1)"""
    tr = good.copy()
    i, r = 9, 4
    tr[C_C + i, r] = (int(tr[C_C + i, r]) + 1) % P
    assert fired(tr, r) == {K_CPACK + i: 1, K_POS + i: P - 1, K_TRANS + i: P - 1}
    assert fired(tr, r + 1) == {}


def test_planted_phase(good):
    """phase := 2: booleanity 2 (1 - 2) = -2; each select limb moves by
    (phase - 2)(s_i - acc_i)"""
    tr = good.copy()
    r = 3
    ph = int(tr[C_PHASE, r])
    tr[C_PHASE, r] = 2
    want = {K_PHASE: P - 2}
    for i in range(NL):
        v = (ph - 2) * (int(tr[C_S + i, r]) - int(tr[C_ACC + i, r]))
        if v % P:
            want[K_SEL + i] = v % P
    assert fired(tr, r) == want


def test_planted_modulus_and_em_limbs(good):
    """n and em are constant columns: a change of +5 at row r fires the
    transition into r (+5) and out of r (-5); for n, position k's q.n sum also
    moves by 5 q_(k-i), so its identity by -5 q_(k-i)"""
    tr = good.copy()
    i, r = 12, 9
    tr[C_EM + i, r] = (int(tr[C_EM + i, r]) + 5) % P
    assert fired(tr, r - 1) == {K_TRANS + 3 * NL + i: 5}
    assert fired(tr, r) == {K_TRANS + 3 * NL + i: P - 5}
    tr = good.copy()
    tr[C_N + i, r] = (int(tr[C_N + i, r]) + 5) % P
    assert fired(tr, r - 1) == {K_TRANS + 2 * NL + i: 5}
    want = {K_TRANS + 2 * NL + i: P - 5}
    for k in range(159):
        j = k - i
        if 0 <= j < NL and int(tr[C_Q + j, r]):
            want[K_POS + k] = (-5 * int(tr[C_Q + j, r])) % P
    assert fired(tr, r) == want


def test_planted_s_limb(good):
    tr = good.copy()
    i, r = 0, 20
    ph = int(tr[C_PHASE, r])
    tr[C_S + i, r] = (int(tr[C_S + i, r]) + 1) % P
    assert fired(tr, r - 1) == {K_TRANS + NL + i: 1}
    want = {K_TRANS + NL + i: P - 1}
    if ph:
        want[K_SEL + i] = P - 1  # op - (s - acc + acc) with s + 1
    assert fired(tr, r) == want


# ---- PKCS#1 v1.5 and a real signature ----

def test_pkcs1_em_shape():
    """pkcs1_v1_5_encoding_shape (rsa2048.rs tests)"""
    d = bytes(range(32))
    em = O.rsa_pkcs1_em_sha256(d)
    ps = 256 - 51 - 3
    assert em[:2] == b"\x00\x01" and em[2:2 + ps] == b"\xff" * ps and em[2 + ps] == 0
    assert em[3 + ps:3 + ps + 19] == bytes.fromhex("3031300d060960864801650304020105000420")
    assert em[-32:] == d


def _prime(rng, bits):
    while True:
        c = rng.getrandbits(bits) | (1 << (bits - 1)) | (1 << (bits - 2)) | 1
        if all(c % p for p in (3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)) and all(
                pow(a, c - 1, c) == 1 for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41)):
            return c


def rsa_key(seed):
    rng = random.Random(seed)
    while True:
        p, q = _prime(rng, 1024), _prime(rng, 1024)
        phi = (p - 1) * (q - 1)
        if p != q and phi % 65537:
            return p * q, pow(65537, -1, phi)


def test_real_signature_chain():
    """an RSASSA-PKCS1-v1_5 SHA-256 signature made with a fresh 2048-bit key:
    the chain's output row equals the encoded message"""
    n, d = rsa_key(0xA1)
    msg = b"DNSSEC-RRSIG-V0|legacy.example.com.|A|10.0.0.3|epoch-0|alg-8"
    em = int.from_bytes(O.rsa_pkcs1_em_sha256(hashlib.sha256(msg).digest()), "big")
    s = pow(em, d, n)
    tr = O.rsa_exp_trace([(n, s, em)])
    assert limbs_value(tr[C_C:C_C + NL, 16]) == em
    for r in (0, 15, 16, 17, 31):
        assert nonzero_slots(tr, r) == set()


# ---- the library's host side against the oracle ----

def test_library_host_trace_matches_oracle():
    import stark_swarm_amd as S
    recs = [rand_record(0x11), rand_record(0x22, bits=2048)]
    got = S.fill_rsa_stacked([S.RsaStackedRecord(*r) for r in recs])
    assert np.array_equal(got, O.rsa_exp_trace(recs))
    got64 = S.fill_rsa_exp_multirow(*recs[0], n_trace=64)
    assert np.array_equal(got64, O.rsa_exp_trace(recs[:1], n_trace=64))


def test_library_host_trace_rejects_bad_input():
    import stark_swarm_amd as S
    n, s, em = rand_record(3)
    with pytest.raises(S.StarkError):
        S.fill_rsa_exp_multirow(n, n, em)
    with pytest.raises(S.StarkError):
        S.fill_rsa_exp_multirow(n, s, em, n_trace=16)
    with pytest.raises(S.StarkError):
        S.fill_rsa_exp_multirow(0, 0, 0)


def test_library_pkcs1_matches_oracle():
    import stark_swarm_amd as S
    for d in (bytes(32), bytes(range(32)), hashlib.sha256(b"x").digest()):
        assert S.emsa_pkcs1_v1_5_encode_sha256(d) == O.rsa_pkcs1_em_sha256(d)
