"""SHA-512 AIR (crates/deep_ali/src/sha512_air.rs) on the CPU: the oracle's
trace builder and 1526-constraint evaluator, and the library's host trace
builder. Pinned the same way as the SHA-256 AIR (no stored proof exists):
FIPS 180-4 SHA-512 digests read from the trace (hashlib, around every
padding boundary), vanishing of every constraint except the 16 H-state slots
of the wrap row, and planted mistakes firing exactly the slots that the
reference's constraint budget (sha512_air.rs:463-476) assigns.
"""
import ctypes
import hashlib

import numpy as np
import pytest

from tests import _oracle as O

P = O.P
BOOL, PACK, SIG0, SIG1, SSIG0, SSIG1 = 0, 512, 528, 658, 788, 918
CH, MAJ, T1, T2, SHIFT, NEWAE, HST, WWIN, SCHED = 1048, 1114, 1436, 1446, 1452, 1464, 1472, 1488, 1518
BIT_A, BIT_B, BIT_C, W_LO = 16, 80, 144, 1048

MSGS = [b"abc", b"", b"x" * 111, b"x" * 112, b"y" * 128, b"z" * 239, b"q" * 240, b"r" * 300]


def nonzero_slots(tr, nb, row):
    return set(fired(tr, nb, row))


def fired(tr, nb, row):
    """{slot: value} of the constraints that do not vanish on `row`"""
    h = tr.shape[1]
    c = O.sha512_constraints(tr[:, row], tr[:, (row + 1) % h], row, nb)
    return {int(i): int(c[i]) for i in np.nonzero(c)[0]}


def xor(x, y):
    """xor_poly (sha512_air.rs:508): x + y - 2xy"""
    return x + y - 2 * x * y


@pytest.mark.parametrize("msg", MSGS, ids=lambda m: "len%d" % len(m))
def test_trace_digest_is_sha512(msg):
    tr, nb, dg = O.sha512_trace(msg)
    assert dg == hashlib.sha512(msg).digest()
    assert nb == (len(msg) + 17 + 127) // 128
    assert tr.max() < 2**32


@pytest.mark.parametrize("msg", [b"abc", b"x" * 112, b"r" * 300], ids=lambda m: "len%d" % len(m))
def test_constraints_vanish_except_wrap(msg):
    tr, nb, _ = O.sha512_trace(msg)
    h = tr.shape[1]
    for r in range(h - 1):
        assert nonzero_slots(tr, nb, r) == set(), "row %d" % r
    assert nonzero_slots(tr, nb, h - 1) == set(range(HST, HST + 16))


@pytest.mark.parametrize("r", [3, 40, 90])
def test_planted_bit_fires_its_slots(r):
    """a_5 := 2: booleanity 5, the lo pack of a, Sigma0 t_41 and t_35 (inputs
    a_(i+28), a_(i+34)), Sigma0 bit 30 (input a_(i+39)), Maj ab_5 / ac_5
    where b_5 / c_5 is 1."""
    tr, nb, _ = O.sha512_trace(b"abc")
    a = [int(tr[BIT_A + i, r]) for i in range(64)]
    b5, c5 = int(tr[BIT_B + 5, r]), int(tr[BIT_C + 5, r])
    tr[BIT_A + 5, r] = 2
    # values: the reference's formulas (sha512_air.rs:933-1066) at a_5 = 2
    t0_30 = xor(a[58], a[0])
    want = {BOOL + 5: 2 * (1 - 2),                           # b (1 - b)
            PACK + 0: -(2 - a[5]) * 32,                      # a_lo - sum_{i<32} 2^i a_i
            SIG0 + 41: xor(a[5], a[11]) - xor(2, a[11]),     # t_41 - xor(a_5, a_11)
            SIG0 + 35: xor(a[63], a[5]) - xor(a[63], 2),     # t_35 - xor(a_63, a_5)
            SIG0 + 64 + 30: xor(t0_30, a[5]) - xor(t0_30, 2)}  # S0_30 - xor(t_30, a_5)
    if b5:
        want[MAJ + 15] = a[5] * b5 - 2 * b5                  # ab_5 - a_5 b_5
    if c5:
        want[MAJ + 16] = a[5] * c5 - 2 * c5                  # ac_5 - a_5 c_5
    assert fired(tr, nb, r) == {k: v % P for k, v in want.items()}
    assert fired(tr, nb, r - 1) == {}


def test_planted_word_fires_its_slots():
    """W[20]_lo + 1: row 20's T1 lo sum (T1_lo_lo + 2^32 cw - sum = -1) and the
    W-window lo feed (nxt W_win[14]_lo - cur W_lo = -1); row 19's schedule lo
    sum, whose next-row W_lo is the planted word (+1) (sha512_air.rs:1081-1082,
    1231, 1246)."""
    tr, nb, _ = O.sha512_trace(b"abc")
    r = 20
    tr[W_LO, r] = (int(tr[W_LO, r]) + 1) % P
    assert fired(tr, nb, r) == {T1: P - 1, WWIN + 28: P - 1}
    assert fired(tr, nb, r - 1) == {SCHED: 1}


def test_phi_quirk_zero():
    tr, nb, _ = O.sha512_trace(b"abc")
    lde = O.lde_columns(tr, 32)
    assert not O.merge_sha512(lde, 128, 32, nb, phi_mode=2).any()


@pytest.mark.parametrize("msg", MSGS, ids=lambda m: "len%d" % len(m))
def test_library_host_trace_matches_oracle(msg):
    import stark_swarm_amd.api as A
    h = A.lib().stark_host_sha512_trace_height(len(msg))
    tr = np.full((1510, h), 0xDEADBEEFDEADBEEF, dtype=np.uint64)  # every cell must be written
    dg = ctypes.create_string_buffer(64)
    nb = A.lib().stark_host_build_sha512_trace(msg, len(msg), tr.ctypes.data_as(A.u64p), dg)
    tr2, nb2, dg2 = O.sha512_trace(msg)
    assert nb == nb2 and dg.raw == dg2 and np.array_equal(tr, tr2)
