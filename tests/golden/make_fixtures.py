"""Regenerate the golden fixtures under tests/golden/ from the reference tree.

Reads data files only (JSON bundles and ark-serialize proof blobs) from the
reference checkout; nothing there is imported or executed. Run in the build
container, where /root/reference exists:

    python tests/golden/make_fixtures.py

Writes:
  primary_state.json   zone salt, authority pk hash, job id, per-shard nonces,
                       pi_hashes, DNS Merkle roots, outer root_f0 and the
                       SHA3-256 digest + length of every golden proof blob
                       (crates/.. primary-state/zone_bundle.json)
  bench_runs.json      DNS Merkle roots of the 4096-record zone split into
                       1 / 2 / 4 shards (bench-runs/run-00{1,2,4}/zone_bundle.json)
  inner-shard-0001.bin, outer-rollup.bin   two golden proofs kept whole so a
                       mismatch can be located to the byte
"""
import hashlib
import json
import os
import shutil

REF = "/root/reference/stark-stir-swarm"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ps = os.path.join(REF, "primary-state")
    d = json.load(open(os.path.join(ps, "zone_bundle.json")))
    blobs = {}
    for name in sorted(os.listdir(os.path.join(ps, "proofs"))):
        b = open(os.path.join(ps, "proofs", name), "rb").read()
        blobs[name] = {"len": len(b), "sha3_256": hashlib.sha3_256(b).hexdigest()}
    out = {
        "zone_salt_hex": d["zone_salt_hex"],
        "record_count": d["record_count"],
        "authority_pk_hash_hex": d["authority_pk_hash_hex"],
        "job_id_hex": d["job_id_hex"],
        "outer_root_f0_hex": d["outer_root_f0_hex"],
        "outer_n_trace": d["outer_n_trace"],
        "outer_n0": d["outer_n0"],
        "inner_pi_hashes_hex": d["inner_pi_hashes_hex"],
        "shards": [
            {
                "shard_id": w["shard_id"],
                "shard_nonce_hex": w["shard_nonce_hex"],
                "pi_hash_hex": w["pi_hash_hex"],
                "merkle_root_hex": w["merkle_root_hex"],
                "record_count": w["record_count"],
                "n_trace": w["n_trace"],
                "proof": os.path.basename(w["proof_path"]),
            }
            for w in d["inner_workers"]
        ],
        "blobs": blobs,
    }
    json.dump(out, open(os.path.join(HERE, "primary_state.json"), "w"), indent=1)
    for name in ("inner-shard-0001.bin", "outer-rollup.bin"):
        shutil.copyfile(os.path.join(ps, "proofs", name), os.path.join(HERE, name))

    runs = []
    br = os.path.join(REF, "bench-runs")
    for run in sorted(os.listdir(br)):
        f = os.path.join(br, run, "zone_bundle.json")
        if not os.path.exists(f):
            continue
        b = json.load(open(f))
        runs.append({
            "run": run,
            "zone_salt_hex": b["zone_salt_hex"],
            "record_count": b["record_count"],
            "shards": [{"shard_id": w["shard_id"], "record_count": w["record_count"],
                        "merkle_root_hex": w["merkle_root_hex"]} for w in b["inner_workers"]],
        })
    json.dump(runs, open(os.path.join(HERE, "bench_runs.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
