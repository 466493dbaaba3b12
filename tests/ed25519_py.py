"""Ed25519 in Python big integers (RFC 8032 sections 5.1.1-5.1.7), written from
the RFC to pin the oracle's and the library's arithmetic. Test infrastructure.
"""
import hashlib

P = 2**255 - 19
L = 2**252 + 27742317777372353535851937790883648493
D = -121665 * pow(121666, -1, P) % P
SQRT_M1 = pow(2, (P - 1) // 4, P)
BY = 4 * pow(5, -1, P) % P


def recover_x(y, sign):
    """x with x^2 = (y^2 - 1) / (d y^2 + 1) and the given parity, or None"""
    if y >= P:
        return None
    x2 = (y * y - 1) * pow(D * y * y + 1, -1, P) % P
    if x2 == 0:
        return None if sign else 0
    x = pow(x2, (P + 3) // 8, P)
    if (x * x - x2) % P:
        x = x * SQRT_M1 % P
    if (x * x - x2) % P:
        return None
    if x & 1 != sign:
        x = P - x
    return x


BX = recover_x(BY, 0)
B = (BX, BY)


def add(p, q):
    """affine twisted Edwards addition, a = -1"""
    (x1, y1), (x2, y2) = p, q
    t = D * x1 * x2 * y1 * y2 % P
    x3 = (x1 * y2 + x2 * y1) * pow(1 + t, -1, P) % P
    y3 = (y1 * y2 + x1 * x2) * pow(1 - t, -1, P) % P
    return x3, y3


IDENT = (0, 1)


def mul(k, p):
    acc = IDENT
    for bit in bin(k)[2:] if k else "":
        acc = add(acc, acc)
        if bit == "1":
            acc = add(acc, p)
    return acc


def neg(p):
    return (-p[0]) % P, p[1]


def compress(p):
    x, y = p
    return (y | (x & 1) << 255).to_bytes(32, "little")


def decompress(b):
    v = int.from_bytes(b, "little")
    y, sign = v & ((1 << 255) - 1), v >> 255
    x = recover_x(y, sign)
    return None if x is None else (x, y)


def h512(*parts):
    return hashlib.sha512(b"".join(parts)).digest()


def secret_expand(seed):
    h = h512(seed)
    a = int.from_bytes(h[:32], "little")
    a &= (1 << 254) - 8
    a |= 1 << 254
    return a, h[32:]


def public_key(seed):
    a, _ = secret_expand(seed)
    return compress(mul(a, B))


def sign(seed, msg):
    a, prefix = secret_expand(seed)
    pk = compress(mul(a, B))
    r = int.from_bytes(h512(prefix, msg), "little") % L
    rb = compress(mul(r, B))
    k = int.from_bytes(h512(rb, pk, msg), "little") % L
    s = (r + k * a) % L
    return pk, rb + s.to_bytes(32, "little")


def verify(pk, sig, msg):
    """cofactored verification [8]([s]B - R - [k]A) = O"""
    a, r = decompress(pk), decompress(sig[:32])
    s = int.from_bytes(sig[32:], "little")
    if a is None or r is None or s >= L:
        return False
    k = int.from_bytes(h512(sig[:32], pk, msg), "little") % L
    res = add(add(mul(s, B), neg(r)), neg(mul(k, a)))
    return mul(8, res) == IDENT


def proj_eq(xyz, aff):
    """projective (X, Y, Z) (Z != 0) represents the affine point"""
    x, y, z = xyz[0], xyz[1], xyz[2]
    return z % P != 0 and (x - aff[0] * z) % P == 0 and (y - aff[1] * z) % P == 0


# RFC 8032 section 7.1, TEST 1-3
RFC_VECTORS = [
    ("9d61b19deffd5a60ba844af492ec2cc44449c5697b326919703bac031cae7f60",
     "d75a980182b10ab7d54bfed3c964073a0ee172f3daa62325af021a68f707511a", "",
     "e5564300c360ac729086e2cc806e828a84877f1eb8e5d974d873e065224901555fb8821590a33bacc61e39701cf9b46bd25bf5f0595bbe24655141438e7a100b"),
    ("4ccd089b28ff96da9db6c346ec114e0f5b8a319f35aba624da8cf6ed4fb8a6fb",
     "3d4017c3e843895a92b70aa74d1b7ebc9c982ccf2ec4968cc0cd55f12af4660c", "72",
     "92a009a9f0d4cab8720e820b5f642540a2b27b5416503f8fb3762223ebdb69da085ac1e43e15996e458f3613d0f11d8c387b2eaeb4302aeeb00d291612bb0c00"),
    ("c5aa8df43f9f837bedb7442f31dcb7b166d38535076f094b85ce3a2e0b4458f7",
     "fc51cd8e6218a1a38da47ed00230f0580816ed13ba3303ac5deb911548908025", "af82",
     "6291d657deec24024827e69c3abe01a30ce548a284743a445e3680d7db5ac3ac18ff9b538d16f290ae67f760984dc6594a7c15e9716ed28dc027beceea1ec40a"),
]


def rfc_vector(i):
    seed, pk, msg, sig = RFC_VECTORS[i]
    return bytes.fromhex(seed), bytes.fromhex(pk), bytes.fromhex(msg), bytes.fromhex(sig)
