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
    h = tr.shape[1]
    c = O.sha512_constraints(tr[:, row], tr[:, (row + 1) % h], row, nb)
    return set(int(i) for i in np.nonzero(c)[0])


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
    b5, c5 = int(tr[BIT_B + 5, r]), int(tr[BIT_C + 5, r])
    tr[BIT_A + 5, r] = 2
    want = {BOOL + 5, PACK + 0, SIG0 + 41, SIG0 + 35, SIG0 + 64 + 30}
    if b5:
        want.add(MAJ + 15)
    if c5:
        want.add(MAJ + 16)
    assert nonzero_slots(tr, nb, r) == want
    assert nonzero_slots(tr, nb, r - 1) == set()


def test_planted_word_fires_its_slots():
    """W[20]_lo + 1: row 20's T1 lo sum and the W-window lo feed; row 19's
    schedule lo sum."""
    tr, nb, _ = O.sha512_trace(b"abc")
    r = 20
    tr[W_LO, r] = (int(tr[W_LO, r]) + 1) % P
    assert nonzero_slots(tr, nb, r) == {T1, WWIN + 28}
    assert nonzero_slots(tr, nb, r - 1) == {SCHED}


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
