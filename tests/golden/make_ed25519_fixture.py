"""Writes tests/golden/ed25519_zsk_ksk.json: the oracle's
prove_zsk_ksk_binding_v2 proof (k_scalar 256, blowup 32, 54 STIR queries;
crates/swarm-dns/src/prover.rs:856-951) of RFC 8032 TEST 1 and of a seeded
signature over a two-block SHA-512 input, as SHA3-256 of the proof bytes,
pi_hash and root_f0.

The oracle builds the 41 519 x 32 768 LDE (about 11 GB) and evaluates all
70 606 constraints at every point, a few minutes per proof on one core, too
slow for a GPU test to recompute, so the GPU test compares with these stored
oracle outputs. Inputs are the RFC vector and a signature from a seeded key
(tests/ed25519_py.py); nothing here reads the reference.

    python tests/golden/make_ed25519_fixture.py
"""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from tests import _oracle as O  # noqa: E402
from tests import ed25519_py as E  # noqa: E402

FS = bytes([0x5A] * 32)
ROOT32 = bytes(range(32))


def cases():
    _, pk, msg, sig = E.rfc_vector(0)
    yield "rfc8032-test1", pk, sig, msg
    seed = hashlib.sha256(b"ed25519-fixture-seed-1").digest()
    msg = b"example.com. 3600 IN DNSKEY 257 3 15 " + bytes(range(48))  # R || A || M spans two SHA-512 blocks
    pk, sig = E.sign(seed, msg)
    yield "seeded-two-block", pk, sig, msg


def main():
    out = []
    for name, pk, sig, msg in cases():
        r = O.prove_zsk_ksk_binding_v2(pk, sig, msg, FS, ROOT32, 256, 32)
        assert not isinstance(r, int), r
        proof, root_f0, pi_hash = r
        out.append({"name": name, "pk": pk.hex(), "sig": sig.hex(), "msg": msg.hex(), "fs_binding": FS.hex(),
                    "merkle_root": ROOT32.hex(), "k_scalar": 256, "blowup": 32, "proof_len": len(proof),
                    "proof_sha3_256": hashlib.sha3_256(proof).hexdigest(), "root_f0": root_f0.hex(),
                    "pi_hash": pi_hash.hex()})
        print(name, len(proof), out[-1]["proof_sha3_256"], flush=True)
    with open(os.path.join(ROOT, "tests", "golden", "ed25519_zsk_ksk.json"), "w") as f:
        json.dump({"proofs": out}, f, indent=1)


if __name__ == "__main__":
    main()
