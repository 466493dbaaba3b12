"""CPU checks of the library's host-side P-256 witness (the limbs,
quotients and carries every mul gadget needs, with the fast reduction mod p
and the exact division by p): stark_host_build_ecdsa_p256_trace against the
oracle's fill, cell for cell, on small chains and on a full 256-row ECDSA
record trace."""
import numpy as np

from tests import _oracle as O
from tests.test_p256_air import GX, GY, N, ec_mul, ecdsa_sign


def S():
    import stark_swarm_amd as S
    return S


def test_host_trace_small_chains_match_oracle():
    s = S()
    g, q = (GX, GY), ec_mul(5, (GX, GY))
    pts = [g[0], g[1], 1, q[0], q[1], 1, q[0], q[1], 1, g[0], g[1], 1]
    a, b = [1, 0, 1, 1, 0, 1, 1], [0, 1, 1, 0, 1, 0, 0, 1]
    ch = s.P256Chains(tuple(pts[0:3]), tuple(pts[3:6]), tuple(pts[6:9]), tuple(pts[9:12]), a, b)
    tr, rp = s.fill_ecdsa_double_multirow(ch, 8)
    want = O.p256_trace(pts, list(rp), a, b, 8)
    assert np.array_equal(tr, want)


def test_host_trace_ecdsa_record_matches_oracle():
    """u1 = e / s, u2 = r / s chains of a signature, 256 rows each: every mul
    gadget's quotient runs up to p - 1 here"""
    s = S()
    qx, qy, r, sig, dg = ecdsa_sign(31337, b"host witness")
    w = pow(sig, -1, N)
    u1, u2 = int.from_bytes(dg, "big") % N * w % N, r * w % N
    bits = lambda u: [(u >> (255 - i)) & 1 for i in range(256)]
    ch = s.P256Chains((GX, GY, 1), (GX, GY, 1), (qx, qy, 1), (qx, qy, 1), bits(u1), bits(u2))
    tr, rp = s.fill_ecdsa_double_multirow(ch, 256)
    want, wrp = O.p256_ecdsa_trace(qx, qy, r, sig, dg)
    assert list(rp) == wrp
    assert np.array_equal(tr, want)
