"""GPU merges with arbitrary combination coefficients, through the C ABI
(stark_merge, stark_merge_sha256 / _sha512, stark_merge_rsa_exp,
stark_merge_ecdsa_p256), against the oracle's merges with the same
coefficients on random LDEs (every constraint non-zero); and the SHA merges'
device routes at the tall C4 heights: the provers' from-trace route
(stark_merge_sha_from_trace_dev) against the direct merge on the full device
LDE (stark_merge_sha_dev), with the LDE's trace-domain points checked against
the trace itself."""
import sys

import numpy as np
import pytest

from tests import _oracle as O

pytestmark = pytest.mark.gpu
P = O.P


@pytest.fixture(scope="module")
def S():
    import stark_swarm_amd as S
    S.default_context()
    return S


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rnd(shape, seed):
    return np.random.default_rng(seed).integers(0, P, size=shape, dtype=np.uint64)


def omega(n):
    return pow(7, (P - 1) // n, P)


@pytest.mark.parametrize("air", ["HashRollup", "Nsec3Chain"])
def test_general_merge_random_coeffs(S, air):
    a = getattr(S.AirType, air)
    n_trace, blowup = 512, 8
    n = n_trace * blowup
    lde = rnd((a.width(), n), 11)
    co = rnd(a.num_constraints(), 12)
    say("merge %s random coeffs, %d points" % (air, n))
    got, _ = S.deep_ali_merge_general(lde, co, a, omega(n), n_trace, blowup, phi_mode=0)
    ref = (O.merge_hash_rollup if air == "HashRollup" else O.merge_nsec3)(lde, n_trace, blowup, 0, coeffs=co)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("alg", [256, 512])
def test_sha_merge_random_coeffs(S, alg):
    w, k = (756, 766) if alg == 256 else (1510, 1526)
    n_trace, blowup = 128, 32
    n = n_trace * blowup
    lde = rnd((w, n), alg)
    co = rnd(k, alg + 1)
    say("merge sha%d random lde, random coeffs" % alg)
    f = S.deep_ali_merge_sha256 if alg == 256 else S.deep_ali_merge_sha512
    got, _ = f(lde, co, omega(n), n_trace, blowup, 1)
    ref = (O.merge_sha256 if alg == 256 else O.merge_sha512)(lde, n_trace, blowup, 1, 1, coeffs=co)
    assert np.array_equal(got, ref)


def test_rsa_merge_random_coeffs(S):
    n_trace, blowup = 32, 4
    n = n_trace * blowup
    lde = rnd((O.RSA_W, n), 21)
    co = rnd(O.RSA_K, 22)
    say("merge rsa random lde, random coeffs")
    got, _ = S.deep_ali_merge_rsa_exp_multirow_streaming(lde, co, omega(n), n_trace, blowup)
    assert np.array_equal(got, O.merge_rsa(lde, n_trace, blowup, 1, coeffs=co))


def test_p256_merge_random_coeffs(S):
    n_trace, blowup = 16, 4
    n = n_trace * blowup
    w = O.p256_width()
    lde = rnd((w, n), 31)
    co = rnd(O.p256_n_constraints(), 32)
    say("merge p256 random lde (%d x %d), random coeffs" % (w, n))
    got, _ = S.deep_ali_merge_ecdsa_double_multirow_streaming(lde, co, omega(n), n_trace, blowup)
    assert np.array_equal(got, O.merge_p256(lde, n_trace, blowup, coeffs=co))


def test_general_path_equals_fast_path_on_reference_coeffs(S):
    """lambda_j = j + 1 written as j + 1 + p (the same field elements, not the
    same words) takes the term-by-term path; its result equals the fast
    path's, on a valid SHA-256 trace LDE and a random RSA LDE"""
    tr, nb = S.build_sha256_trace_multi(bytes(range(100)))
    n_trace, blowup = tr.shape[1], 32
    n = n_trace * blowup
    lde = S.lde_trace_columns(tr, n_trace, blowup)
    j1 = np.arange(1, 767, dtype=np.uint64)
    say("sha256 merge: j + 1 vs j + 1 + p")
    fast, _ = S.deep_ali_merge_sha256(lde, j1, omega(n), n_trace, blowup, nb)
    gen, _ = S.deep_ali_merge_sha256(lde, j1 + np.uint64(P), omega(n), n_trace, blowup, nb)
    assert np.array_equal(fast, gen)
    lde = rnd((O.RSA_W, 128), 41)
    j1 = np.arange(1, O.RSA_K + 1, dtype=np.uint64)
    say("rsa merge: j + 1 vs j + 1 + p")
    fast, _ = S.deep_ali_merge_rsa_exp_multirow_streaming(lde, j1, omega(128), 32, 4)
    gen, _ = S.deep_ali_merge_rsa_exp_multirow_streaming(lde, j1 + np.uint64(P), omega(128), 32, 4)
    assert np.array_equal(fast, gen)


def padded_trace(S, alg, msg, h):
    tr, nb = (S.build_sha256_trace_multi if alg == 256 else S.build_sha512_trace_multi)(msg)
    useful = 128 * nb
    if h > tr.shape[1]:
        tr = np.concatenate([tr, np.repeat(tr[:, useful - 1:useful], h - tr.shape[1], axis=1)], axis=1)
    return np.ascontiguousarray(tr), nb


@pytest.mark.parametrize("alg,lg", [(256, 14), (256, 15), (256, 16), (256, 17), (512, 14), (512, 15), (512, 16),
                                    (512, 17)])
def test_tall_sha_from_trace_equals_direct_merge(S, alg, lg):
    """C4's tall shapes (2^14..2^17 rows, blowup 32): the provers' merge from
    the trace (blowup-2 gate groups, two-pass odd coset and interpolation)
    equals the direct merge on the full LDE, both on the device; the LDE's
    points i = 32 t are the trace rows; random coefficients at 2^15"""
    h, blowup = 1 << lg, 32
    n = h * blowup
    ctx = S.default_context()
    msg = bytes((7 * i + lg) & 255 for i in range(1500 if alg == 256 else 3000))
    tr, nb = padded_trace(S, alg, msg, h)
    w = tr.shape[0]
    k = 766 if alg == 256 else 1526
    say("sha%d 2^%d rows: trace %d x %d, LDE x32 on the device" % (alg, lg, w, h))
    d_tr = ctx.alloc(tr.nbytes).upload(tr)
    d_lde = ctx.alloc(tr.nbytes * blowup)
    ctx._check(S.lib().stark_lde_columns_dev(ctx.h, d_tr.ptr, w, h, blowup, 1, d_lde.ptr))
    # the LDE (no shift) at w_n0^(32 t) = w_h^t is the trace: three whole columns
    for col in (0, w // 2, w - 1):
        assert np.array_equal(d_lde.download((n,), offset=8 * col * n)[::blowup], tr[col]), col
    co = np.arange(1, k + 1, dtype=np.uint64) if lg != 15 else rnd(k, lg)
    d_a, d_b = ctx.alloc(8 * n), ctx.alloc(8 * n)
    say("sha%d 2^%d: direct merge" % (alg, lg))
    S.merge_sha_dev(alg, d_lde, h, blowup, nb, co, d_a, phi_mode=1, ctx=ctx)
    say("sha%d 2^%d: merge from the trace" % (alg, lg))
    S.merge_sha_from_trace_dev(alg, d_tr, h, blowup, nb, co, d_b, phi_mode=1, ctx=ctx)
    a, b = d_a.download((n,)), d_b.download((n,))
    assert a.any()
    assert np.array_equal(a, b)
    for x in (d_tr, d_lde, d_a, d_b):
        x.free()
