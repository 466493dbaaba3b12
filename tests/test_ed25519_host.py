"""CPU checks of the library's Ed25519 host side (no GPU): the witness words
and column descriptors, expanded on the host, give the oracle's v16 trace cell
for cell; native verification and RFC 8032 signing. The device expansion of
the same words is compared in tests/test_gpu_ed25519.py."""
import hashlib

import numpy as np
import pytest

from tests import _oracle as O
from tests import ed25519_py as E


@pytest.fixture(scope="module")
def S():
    import stark_swarm_amd as S
    return S


def seeded(i, n):
    seed = hashlib.sha256(b"ed25519-host-%d" % i).digest()
    msg = bytes((7 * j + i) & 0xFF for j in range(n))
    pk, sig = E.sign(seed, msg)
    return pk, sig, msg


@pytest.mark.parametrize("case", [("rfc", 0, 256), ("rfc", 1, 8), ("rfc", 2, 256), ("seed", 47, 256),
                                  ("seed", 48, 16), ("seed", 200, 8)])
def test_host_trace_equals_oracle(S, case):
    kind, arg, k = case
    if kind == "rfc":
        _, pk, msg, sig = E.rfc_vector(arg)
    else:
        pk, sig, msg = seeded(arg, arg)  # 47 / 48 bytes: R || A || M at the one / two SHA-512 block boundary
    e = O.Ed25519(pk, sig, msg, k)
    shape = S.verify_air_layout_v16(64 + len(msg), k)
    assert (shape.width, shape.height, shape.n_constraints) == (e.width, e.height, e.n_constraints)
    tr = S.fill_verify_air_v16(pk, sig, msg, k)
    ref = e.trace()
    bad = np.argwhere(tr != ref)
    assert len(bad) == 0, bad[:8].tolist()


def test_sign_verify(S):
    for i in range(3):
        seed, pk, msg, sig = E.rfc_vector(i)
        assert S.ed25519_sign(seed, msg) == (pk, sig)
        assert S.ed25519_verify(pk, sig, msg)
        bad = bytearray(sig)
        bad[40] ^= 1
        assert not S.ed25519_verify(pk, bytes(bad), msg)
    for i in range(8):
        seed = hashlib.sha256(b"s%d" % i).digest()
        msg = bytes(range(i * 13))
        pk, sig = S.ed25519_sign(seed, msg)
        assert (pk, sig) == E.sign(seed, msg)
        assert S.ed25519_verify(pk, sig, msg) and O.ed25519_verify(pk, sig, msg)
    # s >= L is rejected (is_canonical, ed25519_verify.rs:67)
    _, pk, msg, sig = E.rfc_vector(0)
    s = int.from_bytes(sig[32:], "little") + E.L
    assert not S.ed25519_verify(pk, sig[:32] + s.to_bytes(32, "little"), msg)


def test_undecodable_key_is_an_error(S):
    _, pk, msg, sig = E.rfc_vector(0)
    bad = None
    for b in range(256):  # an encoding whose y has no square root
        cand = bytes([b]) + pk[1:]
        if E.decompress(cand) is None:
            bad = cand
            break
    assert bad is not None
    with pytest.raises(S.StarkError):
        S.fill_verify_air_v16(bad, sig, msg, 8)
