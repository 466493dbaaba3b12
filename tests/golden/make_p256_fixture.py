"""Writes tests/golden/p256_ecdsa_record.json: the oracle's per-record
ECDSA-P256 proof (prove_record_ecdsa of the mixed-algorithm zone example,
n_trace 256, blowup 32) for seeded signature records, as SHA3-256 of the
proof bytes, pi_hash and root_f0.

The oracle needs about 140 s per record on one core, too slow for a GPU test
to recompute, so the GPU test compares with these stored oracle outputs.
Inputs are signatures from a seeded key and nonce (tests/test_p256_air.py
ecdsa_sign); nothing here reads the reference.

    python tests/golden/make_p256_fixture.py
"""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from tests import _oracle as O  # noqa: E402
from tests.test_p256_air import ecdsa_sign  # noqa: E402

SEEDS = [(11, b"DNSSEC-RRSIG-V0|www.example.com.|A|192.0.2.11|epoch-0|alg-13")]


def main():
    out = []
    for seed, msg in SEEDS:
        qx, qy, r, s, digest = ecdsa_sign(seed, msg)
        proof, root_f0, pi_hash = O.prove_p256_ecdsa(qx, qy, r, s, digest)
        out.append({"seed": seed, "message": msg.decode(), "qx": "%064x" % qx, "qy": "%064x" % qy,
                    "r": "%064x" % r, "s": "%064x" % s, "digest": digest.hex(),
                    "proof_len": len(proof), "proof_sha3_256": hashlib.sha3_256(proof).hexdigest(),
                    "root_f0": root_f0.hex(), "pi_hash": pi_hash.hex()})
        print(seed, len(proof), out[-1]["proof_sha3_256"], flush=True)
    with open(os.path.join(ROOT, "tests", "golden", "p256_ecdsa_record.json"), "w") as f:
        json.dump({"records": out}, f, indent=1)


if __name__ == "__main__":
    main()
