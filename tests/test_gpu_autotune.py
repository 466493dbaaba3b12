"""GPU checks of the launch-plan autotuner: every admissible plan (pass
split, LDS tile) of the NTT, iNTT and coset LDE computes the oracle's values
(each pinned with stark_set_ntt_split / stark_set_ntt_tile), and with
autotune on the tuned plans are recorded and the outputs and proofs stay the
same bytes."""
import sys

import numpy as np
import pytest

from tests import _oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def S():
    import stark_swarm_amd as S
    return S


@pytest.fixture
def ctx(S):
    c = S.Context(0)
    yield c
    c.close()


def two_pass_window(L):
    """admissible first-pass sizes: within 2 of ceil(L/2), both passes 2^3..2^12"""
    mid = (L + 1) // 2
    return list(range(max(mid - 2, L - 12, 3), min(mid + 2, 12, L - 3) + 1))


def two_pass_plans(L):
    """(split, tile) plans: every split with the 2^13 tile, and with the 2^14
    tile when both passes have 2^8 points or more"""
    out = []
    for s in two_pass_window(L):
        out.append((s, 13))
        if s >= 8 and L - s >= 8:
            out.append((s, 14))
    return out


def rand(shape, seed):
    return np.random.default_rng(seed).integers(0, O.P, size=shape, dtype=np.uint64)


@pytest.mark.parametrize("L", [13, 15, 17])
def test_every_ntt_plan_matches_oracle(S, ctx, L):
    a = rand((3, 1 << L), L)
    fwd = np.stack([O.ntt(r) for r in a])
    inv = np.stack([O.ntt(r, True) for r in a])
    for s, t in two_pass_plans(L):
        say("ntt 2^%d split %d/%d tile 2^%d" % (L, s, L - s, t))
        ctx.ntt_split(s)
        ctx.ntt_tile(t)
        assert np.array_equal(S.ntt(a, ctx=ctx), fwd), (s, t)
        assert np.array_equal(S.ntt(a, inverse=True, ctx=ctx), inv), (s, t)


@pytest.mark.parametrize("L,blowup", [(13, 8), (14, 2), (10, 4), (11, 32), (16, 8)])
def test_every_lde_plan_matches_oracle(S, ctx, L, blowup):
    """two-pass plans, and for 2^10 / 2^11 points also the one-pass form
    with the cosets as repetitions (split = L), on either tile"""
    a = rand((3, 1 << L), 100 + L)
    want = O.lde_columns(a, blowup)
    plans = two_pass_plans(L) + ([(L, 13), (L, 14)] if L <= 12 else [])
    for s, t in plans:
        say("lde 2^%d x %d split %d tile 2^%d" % (L, blowup, s, t))
        ctx.ntt_split(s)
        ctx.ntt_tile(t)
        assert np.array_equal(S.lde_trace_columns(a, 1 << L, blowup, ctx=ctx), want), (s, t)


def test_autotune_records_plans_and_keeps_bytes(S, ctx):
    ctx.autotune(True)
    a = rand((4, 1 << 13), 77)
    say("lde 2^13 x 8 autotuned")
    want = O.lde_columns(a, 8)
    assert np.array_equal(S.lde_trace_columns(a, 1 << 13, 8, ctx=ctx), want)
    rep = S.autotune_report()
    for key in (("lde", 13, 3, 2), ("intt", 13, 0, 2)):
        assert key in rep, sorted(rep)
        best, ms = rep[key]
        assert sorted(ms) == sorted(two_pass_plans(13)) and best in ms
        assert all(v > 0 for v in ms.values())
        assert ms[best] == min(ms.values())
    # the cached plan serves the next call of the shape
    assert np.array_equal(S.lde_trace_columns(a, 1 << 13, 8, ctx=ctx), want)
    b = rand((2, 1 << 10), 78)
    say("lde 2^10 x 4 autotuned (one-pass candidate)")
    assert np.array_equal(S.lde_trace_columns(b, 1 << 10, 4, ctx=ctx), O.lde_columns(b, 4))
    best, ms = S.autotune_report()[("lde", 10, 2, 1)]
    assert sorted(ms) == sorted(two_pass_plans(10) + [(10, 13), (10, 14)])


def test_autotuned_prover_matches_untuned(S, ctx):
    """a SHA-256 proof at 2^13 rows (two-pass odd-coset transform,
    interpolation and quotient transforms) with the tuned plans, against
    the same proof with the default splits"""
    m, f = b"t" * 500, b"\x09" * 32
    say("prove_sha_batch 2^13 default splits")
    plain = S.prove_sha_batch(256, [m], [f], heights=[8192], ctx=ctx)[0]
    ctx.autotune(True)
    say("prove_sha_batch 2^13 autotuned")
    tuned = S.prove_sha_batch(256, [m], [f], heights=[8192], ctx=ctx)[0]
    again = S.prove_sha_batch(256, [m], [f], heights=[8192], ctx=ctx)[0]
    assert tuned.proof_blob == plain.proof_blob == again.proof_blob
