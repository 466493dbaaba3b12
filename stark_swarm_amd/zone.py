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
