"""Synthetic DNS zones that drive the provers.

build_synthetic_zone / split_evenly restate the swarm controller's zone
generator (crates/swarm-ctrl/src/main.rs:1317-1357), so the golden
primary-state job (3072 records in 3 shards) can be regenerated.
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
        e = int.from_bytes(digest, "big") % _P256_N
        while True:
            k = rng.randrange(1, _P256_N)
            r = _p256_mul(k, _P256_G)[0] % _P256_N
            s = pow(k, -1, _P256_N) * (e + r * d) % _P256_N
            if r and s:
                break
        out.append(EcdsaRecord(q[0], q[1], r, s, digest))
    return out
