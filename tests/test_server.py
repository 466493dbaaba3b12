"""Proving-server scheduling on the CPU (dry-run mode: the workers sleep
instead of proving, so admission, batching and ordering are tested without
a GPU). The proof results of the same server are checked on the GPU in
test_gpu_parity.py."""
import hashlib
import threading
import time

import pytest

import stark_swarm_amd as S
from stark_swarm_amd.server import ProvingServer
from stark_swarm_amd.zone import build_synthetic_zone

KSK_NEED_1BLK = 8 * 756 * 4 * 128 + 8 * 5 * 2 * 128 + 8 * 24 * 4096  # Server::need_bytes for a one-block key


def _keys(n, start=0):
    return [bytes([(i * 31 + j) & 255 for j in range(20 + (i % 3) * 40)]) for i in range(start, start + n)]


def test_budget_is_never_exceeded_and_all_complete():
    budget = 5 * KSK_NEED_1BLK
    with ProvingServer(workers=4, hbm_budget=budget, max_batch=8, dry_run=True) as srv:
        ids = [srv.submit_ds_ksk(k, b"\0" * 32, b"\1" * 32) for k in _keys(120)]
        res = [srv.wait(i, timeout=60) for i in ids]
        st = srv.stats()
    assert st["completed"] == 120
    assert st["peak_bytes_admitted"] <= budget
    assert all(r.batch_size >= 1 for r in res)
    assert max(r.batch_size for r in res) > 1, "queued requests must be batched"


def test_batch_heads_are_admitted_in_submission_order():
    """The queue head is never overtaken: batches start in the order of their
    oldest request. Younger requests may join an older head's batch when they
    fit the budget (backfilling), never start a batch of their own early."""
    with ProvingServer(workers=3, hbm_budget=4 * KSK_NEED_1BLK, max_batch=4, dry_run=True) as srv:
        ids = [srv.submit_ds_ksk(k, b"\0" * 32, b"\1" * 32) for k in _keys(60)]
        res = [srv.wait(i, timeout=60) for i in ids]
    batches = {}
    for r in res:
        batches.setdefault(r.batch_seq, []).append(r)
    assert all(len(v) == v[0].batch_size for v in batches.values())
    heads = sorted((min(r.id for r in v), seq) for seq, v in batches.items())
    assert all(a[1] < b[1] for a, b in zip(heads, heads[1:]))


def test_oversized_request_runs_alone():
    with ProvingServer(workers=2, hbm_budget=KSK_NEED_1BLK // 2, max_batch=8, dry_run=True) as srv:
        ids = [srv.submit_ds_ksk(k, b"\0" * 32, b"\1" * 32) for k in _keys(6)]
        res = [srv.wait(i, timeout=60) for i in ids]
    assert all(r.batch_size == 1 for r in res)
    # no two of them overlapped in time
    spans = sorted((r.t_start, r.t_done) for r in res)
    assert all(b[0] >= a[1] for a, b in zip(spans, spans[1:]))


def test_mixed_kinds_and_concurrent_submitters():
    zone = build_synthetic_zone(500)
    out = {}

    def submitter(srv, base):
        for j in range(12):
            if j % 4 == 1:
                rid = srv.submit_inner_shard(b"swarm-test-zone1", zone[base + 10 * j: base + 10 * j + 10], b"\2" * 32)
            elif j % 4 == 2:
                rid = srv.submit_rsa(S.RsaStackedRecord((1 << 2047) | 1, 5 + j, 7), b"\4" * 32)
            elif j % 4 == 3:  # dry run: the record is scheduled, never proven
                rid = srv.submit_p256(S.EcdsaRecord(1, 2, 3, 4, bytes(32)))
            else:
                rid = srv.submit_ds_ksk(_keys(1, base + j)[0], b"\0" * 32, b"\3" * 32)
            out[rid] = srv.wait(rid, timeout=60)

    with ProvingServer(workers=4, hbm_budget=8 << 30, max_batch=16, dry_run=True) as srv:
        th = [threading.Thread(target=submitter, args=(srv, 100 * t)) for t in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert srv.stats()["completed"] == 48
    kinds = {r.kind for r in out.values()}
    assert kinds == {0, 1, 2, 3}
    # a batch holds one kind only
    by_batch = {}
    for r in out.values():
        by_batch.setdefault(r.batch_seq, set()).add(r.kind)
    assert all(len(k) == 1 for k in by_batch.values())


def test_wait_timeout_and_bad_requests():
    with ProvingServer(workers=1, hbm_budget=1 << 30, dry_run=True) as srv:
        with pytest.raises(TimeoutError):
            srv.wait(12345, timeout=0.05)
        with pytest.raises(S.StarkError):
            srv.submit_inner_shard(b"s" * 16, S.PackedShards([[]]), b"\0" * 32)
