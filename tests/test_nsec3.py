"""NSEC3 chain completeness (AirType::Nsec3Chain, crates/deep_ali/src/
air_workloads.rs:789-870, and prove_nsec3_completeness, crates/swarm-dns/
src/prover.rs:348-446) on the CPU: the oracle's trace builder, padding rule,
merge and closure check.

Pins: the reference's own trace test (nsec3_chain_constraints_zero_on_valid_
trace: every chain link holds on every row, the wrap row included), the
padding rule stated at air_workloads.rs:825-850, and one consequence of the
reference's degree bounds that the merge must reproduce: all four
constraints have degree 1, so phi has degree below n_trace and the Z_H
quotient (remainder dropped, lib.rs:190-228) is zero whatever the trace.
"""
import hashlib

import numpy as np

from tests import _oracle as O

P = O.P


def links_hold(tr):
    n = tr.shape[1]
    return all(int(tr[k, (i + 1) % n]) == int(tr[4 + k, i]) for i in range(n) for k in range(4))


def test_reference_chain_trace_vanishes():
    """nsec3_chain_constraints_zero_on_valid_trace (air_workloads.rs:760-786)"""
    n = 16
    owners = [[i, (31 * i) % 2**64, (17 * i) % 2**64, (7 * i) % 2**64] for i in range(n)]
    chain = [(owners[i], owners[(i + 1) % n]) for i in range(n)]
    tr = O.nsec3_trace(n, chain)
    assert links_hold(tr)
    assert np.array_equal(tr[:4, 3], np.array(owners[3], dtype=np.uint64))


def test_padding_repeats_last_next_and_closes():
    owners = [[10 + i, 20 + i, 30 + i, 40 + i] for i in range(5)]
    chain = [(owners[i], owners[(i + 1) % 5]) for i in range(5)]
    tr = O.nsec3_trace(8, chain)
    last_next = owners[0]
    for i in range(5, 8):
        assert [int(v) for v in tr[:4, i]] == last_next
    for i in range(5, 7):
        assert [int(v) for v in tr[4:, i]] == last_next
    assert [int(v) for v in tr[4:, 7]] == [int(v) for v in tr[:4, 0]]
    assert links_hold(tr)


def test_planted_links_fire_with_their_values():
    """C_k = nxt[k] - cur[4 + k] (air_workloads.rs:862-870): owner limb 2 of
    row 5 plus 3 fires row 4's C_2 with +3; next limb 1 of row 5 plus 3 fires
    row 5's C_1 with -3; every other row and slot vanishes"""
    n = 16
    owners = [[i, (31 * i) % 2**64, (17 * i) % 2**64, (7 * i) % 2**64] for i in range(n)]
    tr = O.nsec3_trace(n, [(owners[i], owners[(i + 1) % n]) for i in range(n)])

    def fired(t):
        out = {}
        for r in range(n):
            c = O.nsec3_constraints(t[:, r], t[:, (r + 1) % n])
            out.update({(r, k): int(c[k]) for k in range(4) if c[k]})
        return out
    assert fired(tr) == {}
    t = tr.copy()
    t[2, 5] = (int(t[2, 5]) + 3) % P
    assert fired(t) == {(4, 2): 3}
    t = tr.copy()
    t[4 + 1, 5] = (int(t[4 + 1, 5]) + 3) % P
    assert fired(t) == {(5, 1): P - 3}


def test_limbs_reduced_into_the_field():
    big = [P, P + 5, 2**64 - 1, 3]
    tr = O.nsec3_trace(8, [(big, big)] * 8)
    assert [int(v) for v in tr[:4, 0]] == [0, 5, (2**64 - 1) - P, 3]


def test_merge_is_zero_by_degree():
    """degree-1 constraints: the quotient is zero for a valid chain and for a
    broken one alike, with phi computed (n0 >= 4096) or not"""
    n = 128
    owners = [[i, i * 3, i * 5, i * 7] for i in range(n)]
    good = O.nsec3_trace(n, [(owners[i], owners[(i + 1) % n]) for i in range(n)])
    bad = good.copy()
    bad[5, 9] = (int(bad[5, 9]) + 1) % P
    for tr in (good, bad):
        lde = O.lde_columns(tr, 32)
        for mode in (0, 1):
            assert not O.merge_nsec3(lde, n, 32, mode).any()


def make_chain(n):
    """the sorted closed chain of the reference's nsec3_completeness_bench.rs:34-52"""
    owners = sorted(hashlib.sha3_256(b"NSEC3-DEMO-OWNER" + i.to_bytes(8, "little")).digest() for i in range(n))
    return [(owners[i], owners[(i + 1) % n]) for i in range(n)]


def test_oracle_prover_recipes():
    salt, fsb = b"swarm-test-zone1", b"\xCA" * 32
    recs = make_chain(5)
    proof, pi, cr, rf, nt = O.prove_nsec3_completeness(recs, salt, fsb)
    assert nt == 8 and rf == proof[:32]
    h = hashlib.sha3_256(b"DNS-NSEC3-CHAIN-ROOT-V1" + salt + (5).to_bytes(8, "little"))
    for o, x in recs:
        h.update(o + x)
    assert cr == h.digest()
    assert pi == hashlib.sha3_256(b"DNS-NSEC3-PIHASH-V1" + salt + (5).to_bytes(8, "little") + cr + rf + fsb).digest()


def test_oracle_rejects_open_or_empty_chain():
    recs = make_chain(6)
    recs[-1] = (recs[-1][0], bytes(32))
    assert O.prove_nsec3_completeness(recs, b"s" * 16, bytes(32)) == -2
    assert O.prove_nsec3_completeness([], b"s" * 16, bytes(32)) == -1
