"""Synthetic DNS zones that drive the provers.

build_synthetic_zone / split_evenly restate the swarm controller's zone
generator (crates/swarm-ctrl/src/main.rs:1317-1357), so the golden
primary-state job (3072 records in 3 shards) can be regenerated.

synthetic_rrset_zone / ZoneSigner make the signed zone of the serving
workload: RRsets with seeded owner names and rdata, each signed with a
seeded Ed25519, ECDSA-P256 or RSA-2048 key over the record message of the
mixed-algorithm zone example (mixed_algorithm_zone_rollup.rs:120-123,
207-210, 335-338), generalised from A records to A / AAAA / TXT.
"""
from .api import DnsRecord


def build_synthetic_zone(n):
    out = []
    for i in range(n):
        k = i % 4
        if k == 0:
            out.append(DnsRecord.a("host-%08x.example.com" % i, 300, [10, (i >> 16) & 0xFF, (i >> 8) & 0xFF, i & 0xFF]))
        elif k == 1:
            b = i.to_bytes(8, "big") + (i ^ 0xCAFEBABEDEADBEEF).to_bytes(8, "big")
            out.append(DnsRecord.aaaa("v6-%08x.example.com" % i, 300, b))
        elif k == 2:
            out.append(DnsRecord.txt("txt-%08x.example.com" % i, 60, "v=auth1;rec=%d" % i))
        else:
            out.append(DnsRecord.mx("mail-%08x.example.com" % i, 300, i % 1000, "mx%d.example.com" % (i & 0xFF)))
    return out


def split_evenly(records, n_shards):
    if n_shards == 0:
        return []
    base, rem = divmod(len(records), n_shards)
    out, i = [], 0
    for s in range(n_shards):
        ln = base + (1 if s < rem else 0)
        out.append(records[i:i + ln])
        i += ln
    return out


# P-256 (FIPS 186-4 D.1.2.3): y^2 = x^3 - 3x + b over F_p, base point G of order n
_P256_P = 2**256 - 2**224 + 2**192 + 2**96 - 1
_P256_N = 0xFFFFFFFF00000000FFFFFFFFFFFFFFFFBCE6FAADA7179E84F3B9CAC2FC632551
_P256_G = (0x6B17D1F2E12C4247F8BCE6E563A440F277037D812DEB33A0F4A13945D898C296,
           0x4FE342E2FE1A7F9B8EE7EB4A7C0F9E162BCE33576B315ECECBB6406837BF51F5)


def _p256_add(p1, p2):
    if p1 is None:
        return p2
    if p2 is None:
        return p1
    (x1, y1), (x2, y2) = p1, p2
    P = _P256_P
    if x1 == x2 and (y1 + y2) % P == 0:
        return None
    if p1 == p2:
        lam = (3 * x1 * x1 - 3) * pow(2 * y1, -1, P) % P
    else:
        lam = (y2 - y1) * pow(x2 - x1, -1, P) % P
    x3 = (lam * lam - x1 - x2) % P
    return x3, (lam * (x1 - x3) - y1) % P


def _p256_mul(k, pt):
    acc = None
    for bit in bin(k)[2:]:
        acc = _p256_add(acc, acc)
        if bit == "1":
            acc = _p256_add(acc, pt)
    return acc


def _p256_sign(d, digest, rng):
    e = int.from_bytes(digest, "big") % _P256_N
    while True:
        k = rng.randrange(1, _P256_N)
        r = _p256_mul(k, _P256_G)[0] % _P256_N
        s = pow(k, -1, _P256_N) * (e + r * d) % _P256_N
        if r and s:
            return r, s


def synthetic_ecdsa_records(n, seed, names=None):
    """n ECDSA-P256 signature records over SHA-256 digests of
    "DNSSEC-RRSIG-V0|<owner>|A|<ip>|epoch-0|alg-13" messages (the record
    message of mixed_algorithm_zone_rollup.rs:205-208), from seeded keys and
    nonces; returns EcdsaRecord values. Synthetic data for benchmarks and the
    serving mix."""
    import hashlib
    import random

    from .api import EcdsaRecord
    rng = random.Random(seed)
    out = []
    for i in range(n):
        d = rng.randrange(1, _P256_N)
        q = _p256_mul(d, _P256_G)
        owner = names[i] if names else "host-%06d.example.com." % i
        msg = "DNSSEC-RRSIG-V0|%s|A|10.%d.%d.%d|epoch-0|alg-13" % (owner, (i >> 16) & 255, (i >> 8) & 255, i & 255)
        digest = hashlib.sha256(msg.encode()).digest()
        r, s = _p256_sign(d, digest, rng)
        out.append(EcdsaRecord(q[0], q[1], r, s, digest))
    return out


def synthetic_ed25519_signatures(n, seed, min_len=32, max_len=256):
    """n Ed25519 signatures (RFC 8032) from seeded 32-byte secrets over seeded
    random messages of min_len..max_len bytes: (public key, signature,
    message) triples. Synthetic data for benchmarks and the serving mix."""
    import hashlib
    import random

    from .api import ed25519_sign

    rng = random.Random(seed)
    out = []
    for _ in range(n):
        secret = bytes(rng.getrandbits(8) for _ in range(32))
        msg = hashlib.shake_256(secret + b"msg").digest(rng.randint(min_len, max_len))
        pk, sig = ed25519_sign(secret, msg)
        out.append((pk, sig, msg))
    return out


_SMALL_PRIMES = [p for p in range(3, 2000) if all(p % q for q in range(2, int(p ** 0.5) + 1))]


def _probable_prime(n, rng, rounds=24):
    """Miller-Rabin with `rounds` seeded bases after trial division"""
    for p in _SMALL_PRIMES:
        if n % p == 0:
            return n == p
    d, r = n - 1, 0
    while d % 2 == 0:
        d //= 2
        r += 1
    for _ in range(rounds):
        x = pow(rng.randrange(2, n - 2), d, n)
        if x in (1, n - 1):
            continue
        for _ in range(r - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def _rsa_prime(rng, bits):
    while True:
        c = rng.getrandbits(bits) | (3 << (bits - 2)) | 1  # top two bits set: p * q has 2 * bits bits
        if (c - 1) % 65537 and _probable_prime(c, rng):
            return c


class RsaKey:
    """RSA-2048 key with e = 65537 from a seeded prime search; signs
    PKCS#1 v1.5 SHA-256 (RFC 8017 section 8.2) with the CRT."""

    def __init__(self, rng):
        while True:
            p, q = _rsa_prime(rng, 1024), _rsa_prime(rng, 1024)
            if p != q and (p * q).bit_length() == 2048:
                break
        self.n = p * q
        d = pow(65537, -1, (p - 1) * (q - 1))
        self._crt = (p, q, d % (p - 1), d % (q - 1), pow(q, -1, p))

    def sign_em(self, em):
        p, q, dp, dq, qinv = self._crt
        m = int.from_bytes(em, "big")
        m1, m2 = pow(m, dp, p), pow(m, dq, q)
        return m2 + q * ((m1 - m2) * qinv % p)


RRSET_TYPES = ("A", "AAAA", "TXT")
ALG_ED25519, ALG_P256, ALG_RSA = 15, 13, 8


class RRset:
    __slots__ = ("owner", "rtype", "rdata", "alg", "key")

    def __init__(self, owner, rtype, rdata, alg, key):
        self.owner, self.rtype, self.rdata, self.alg, self.key = owner, rtype, rdata, alg, key

    def message(self):
        """the signed record message: DNSSEC-RRSIG-V0|owner|TYPE|rdata|epoch-0|alg-N"""
        return ("DNSSEC-RRSIG-V0|%s|%s|%s|epoch-0|alg-%d" % (self.owner, self.rtype, self.rdata, self.alg)).encode()


def synthetic_rrset_zone(n, seed, keys_per_alg=(16, 16, 4)):
    """n RRsets: owner names of 2-4 labels of 3-12 lowercase letters (ending
    in "."), types 60 % A / 25 % AAAA / 15 % TXT (TXT rdata 16-255 printable
    bytes), signing algorithms 50 % Ed25519 / 30 % ECDSA-P256 / 20 %
    RSA-2048, each RRset naming one of keys_per_alg seeded keys of its
    algorithm. Only names and rdata are made here; ZoneSigner signs."""
    import random
    rng = random.Random(seed)
    letters = "abcdefghijklmnopqrstuvwxyz"
    printable = letters + letters.upper() + "0123456789 =;:-_.,"
    out = []
    for _ in range(n):
        owner = ".".join("".join(rng.choice(letters) for _ in range(rng.randint(3, 12)))
                         for _ in range(rng.randint(2, 4))) + "."
        u = rng.random()
        if u < 0.60:
            rtype, rdata = "A", ".".join(str(rng.randrange(256)) for _ in range(4))
        elif u < 0.85:
            rtype, rdata = "AAAA", ":".join("%x" % rng.randrange(65536) for _ in range(8))
        else:
            rtype, rdata = "TXT", "".join(rng.choice(printable) for _ in range(rng.randint(16, 255)))
        v = rng.random()
        alg, nk = (ALG_ED25519, keys_per_alg[0]) if v < 0.5 else (ALG_P256, keys_per_alg[1]) if v < 0.8 else \
            (ALG_RSA, keys_per_alg[2])
        out.append(RRset(owner, rtype, rdata, alg, rng.randrange(nk)))
    return out


class ZoneSigner:
    """The zone's seeded keys (keys_per_alg Ed25519 seeds, P-256 scalars and
    RSA-2048 keys) and the proof request of one RRset's signature:
    ("ed25519", pk, sig, msg, pk_hash_pre), ("p256", EcdsaRecord) or
    ("rsa", RsaStackedRecord, pk_hash_pre) -- pk_hash_pre being the record's
    public-input hash of the mixed-algorithm zone example. An RRset's key
    index is taken modulo the signer's key count of its algorithm."""

    def __init__(self, seed, keys_per_alg=(16, 16, 4)):
        import hashlib
        import random
        rng = random.Random(seed)
        self.ed_seeds = [hashlib.sha256(b"zsk-ed25519-%d-%d" % (seed, i)).digest() for i in range(keys_per_alg[0])]
        self.p256 = []
        for _ in range(keys_per_alg[1]):
            d = rng.randrange(1, _P256_N)
            self.p256.append((d, _p256_mul(d, _P256_G)))
        self.rsa = [RsaKey(rng) for _ in range(keys_per_alg[2])]
        self._rng = rng

    def request(self, rrset):
        import hashlib

        from .api import (EcdsaRecord, RsaStackedRecord, ed25519_record_pk_hash, ed25519_sign,
                          emsa_pkcs1_v1_5_encode_sha256, rsa_record_pk_hash)
        msg = rrset.message()
        if rrset.alg == ALG_ED25519:
            pk, sig = ed25519_sign(self.ed_seeds[rrset.key % len(self.ed_seeds)], msg)
            return ("ed25519", pk, sig, msg, ed25519_record_pk_hash(pk, sig, msg))
        digest = hashlib.sha256(msg).digest()
        if rrset.alg == ALG_P256:
            d, q = self.p256[rrset.key % len(self.p256)]
            r, s = _p256_sign(d, digest, self._rng)
            return ("p256", EcdsaRecord(q[0], q[1], r, s, digest))
        key = self.rsa[rrset.key % len(self.rsa)]
        em = emsa_pkcs1_v1_5_encode_sha256(digest)
        s = key.sign_em(em)
        n_be, sig = key.n.to_bytes(256, "big"), s.to_bytes(256, "big")
        return ("rsa", RsaStackedRecord(key.n, s, int.from_bytes(em, "big")), rsa_record_pk_hash(n_be, sig, msg, em))
