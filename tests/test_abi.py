"""CPU checks of the C ABI library: it loads, exports every symbol that
include/stark_gpu.h declares, its host-side helpers match the oracle, and the
Python mirror's pure-host logic follows the reference. No device work."""
import ctypes
import hashlib
import os
import re

import numpy as np

from tests import _oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "stark_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(stark_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    import stark_swarm_amd as S
    L = S.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s


def test_host_sha3_and_rng_match_oracle():
    import stark_swarm_amd as S
    L = S.lib()
    for n in [0, 5, 135, 136, 300]:
        m = bytes(range(256))[:n] * 1 if n <= 256 else bytes(n)
        out = ctypes.create_string_buffer(32)
        L.stark_host_sha3_256(m, len(m), out)
        assert out.raw == hashlib.sha3_256(m).digest()
    for seed in [0, 1, 12345, O.P - 1, 2**63 + 5]:
        seed %= O.P
        assert L.stark_host_stdrng_first_u64(seed) == O.lib().ref_stdrng_first_u64(seed)


def test_dns_leaf_hash_matches_oracle():
    from stark_swarm_amd.zone import build_synthetic_zone
    salt = b"swarm-test-zone1"
    recs = build_synthetic_zone(12)
    ref = O.zone_leaf_hashes(0, 12, salt)
    for i, r in enumerate(recs):
        assert r.leaf_hash(salt) == ref[32 * i:32 * (i + 1)]
        canon = ctypes.create_string_buffer(256)
        n = O.lib().ref_zone_record_canonical(i, canon)
        assert n > 13


def test_schedule_and_params():
    import stark_swarm_amd as S
    assert S.make_schedule(131072) == [8, 8, 8, 8, 8, 4]
    assert S.make_schedule(512) == [8, 8, 8]
    assert S.make_schedule(8192) == [8, 8, 8, 8, 2]
    assert S.make_schedule(16, S.LdtMode.Fri) == [2, 2, 2, 2]
    p = S.build_params(131072, b"\x07" * 32)
    assert p.r == 54 and p.s0 == 54 and p.coeff_commit_final and p.stir and p.d_final == 1
    assert S.FriDomain.new_radix2(131072).omega == 12380578893860276750


def test_context_without_gpu_reports_error_not_crash():
    import torch
    import stark_swarm_amd as S
    if torch.cuda.is_available():
        return
    h = ctypes.c_void_p()
    rc = S.lib().stark_ctx_create(0, ctypes.byref(h))
    assert rc != 0 and not h.value


def test_shard_fs_binding_matches_oracle():
    import stark_swarm_amd as S
    a, b, c = b"\x01" * 32, b"\x02" * 32, b"\x03" * 32
    assert S.shard_fs_binding(a, b, 7, c) == O.shard_fs_binding(a, b, 7, c)


def test_autotune_switches_without_a_device():
    """the process-wide default and the plan report need no device: with
    nothing tuned the report is empty, and a null context sets the default"""
    import stark_swarm_amd as S
    S.set_autotune_default(True)
    S.set_autotune_default(False)
    assert S.autotune_report() == {}
    assert S.lib().stark_set_ntt_split(None, 10) != 0  # a split needs a context
