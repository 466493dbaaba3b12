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
from stark_swarm_amd.zone import build_synthetic_zone, synthetic_ecdsa_records, synthetic_ed25519_signatures

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


RSA_N = (1 << 2047) | 0x1D  # an odd 2048-bit modulus; em = s^65537 mod n makes a valid record


def _rsa(s):
    return S.RsaStackedRecord(RSA_N, s, pow(s, 65537, RSA_N))


ECDSA = synthetic_ecdsa_records(2, 0x5EED0101)
EDS = synthetic_ed25519_signatures(3, 0x5EED0102)


def test_mixed_kinds_and_concurrent_submitters():
    zone = build_synthetic_zone(500)
    out = {}

    def submitter(srv, base):
        for j in range(15):
            if j % 5 == 1:
                rid = srv.submit_inner_shard(b"swarm-test-zone1", zone[base + 10 * j: base + 10 * j + 10], b"\2" * 32)
            elif j % 5 == 2:
                rid = srv.submit_rsa(_rsa(5 + j), b"\4" * 32)
            elif j % 5 == 3:
                rid = srv.submit_p256(ECDSA[j & 1])
            elif j % 5 == 4:
                rid = srv.submit_ed25519(*EDS[j % 3], b"\5" * 32, b"\6" * 32)
            else:
                rid = srv.submit_ds_ksk(_keys(1, base + j)[0], b"\0" * 32, b"\3" * 32)
            out[rid] = srv.wait(rid, timeout=60)

    with ProvingServer(workers=4, hbm_budget=8 << 30, max_batch=16, dry_run=True) as srv:
        th = [threading.Thread(target=submitter, args=(srv, 100 * t)) for t in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert srv.stats()["completed"] == 60
    kinds = {r.kind for r in out.values()}
    assert kinds == {0, 1, 2, 3, 4}
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


def _bad_requests():
    """one wrong record of each signature kind, each with the reason the
    submit-time check gives"""
    pk, sig, msg = EDS[0]
    q = ECDSA[0]
    return [
        ("rsa", lambda srv: srv.submit_rsa(S.RsaStackedRecord(RSA_N, 5, 7), b"\4" * 32), "RSA"),
        ("rsa s >= n", lambda srv: srv.submit_rsa(S.RsaStackedRecord(RSA_N, RSA_N + 2, 7), b"\4" * 32), "RSA"),
        ("p256", lambda srv: srv.submit_p256(S.EcdsaRecord(q.qx, q.qy, q.r, (q.s + 1), q.digest)), "ecdsa"),
        ("ed25519", lambda srv: srv.submit_ed25519(pk, sig, msg + b"x", b"\5" * 32, b"\6" * 32), "Ed25519"),
    ]


def test_submit_rejects_signatures_that_do_not_verify():
    with ProvingServer(workers=1, hbm_budget=1 << 30, dry_run=True) as srv:
        for name, submit, why in _bad_requests():
            with pytest.raises(S.StarkError) as ei:
                submit(srv)
            assert ei.value.code == -1 and why in str(ei.value), name
        # valid ones of each kind are accepted
        for rid in [srv.submit_rsa(_rsa(3), b"\4" * 32), srv.submit_p256(ECDSA[1]),
                    srv.submit_ed25519(*EDS[1], b"\5" * 32, b"\6" * 32)]:
            assert srv.wait(rid, timeout=30).status == 0
    # with the checks off they are scheduled (the batch fallback is tested on the GPU)
    with ProvingServer(workers=1, hbm_budget=1 << 30, dry_run=True, validate=False) as srv:
        ids = [submit(srv) for _, submit, _ in _bad_requests()]
        assert all(srv.wait(i, timeout=30).status == 0 for i in ids)


def test_deadline_expires_in_the_queue():
    """one worker, each request alone: requests behind a long queue whose
    deadline passes before admission complete with E_TIMEOUT and no proof;
    the rest (no deadline, or a generous one) complete normally"""
    big = KSK_NEED_1BLK * 200  # the dry-run worker sleeps ~need / 64 MiB us: a few ms per request
    with ProvingServer(workers=1, hbm_budget=big, max_batch=1, dry_run=True) as srv:
        keys = _keys(30)
        plain = [srv.submit_ds_ksk(k, b"\0" * 32, b"\1" * 32) for k in keys[:20]]
        late = [srv.submit_ds_ksk(k, b"\0" * 32, b"\1" * 32, deadline=1e-4) for k in keys[20:25]]
        roomy = [srv.submit_ds_ksk(k, b"\0" * 32, b"\1" * 32, deadline=60) for k in keys[25:]]
        for i in plain + roomy:
            assert srv.wait(i, timeout=60).status == 0
        for i in late:
            with pytest.raises(S.StarkError) as ei:
                srv.wait(i, timeout=60)
            assert ei.value.code == -6 and "deadline" in str(ei.value)
        assert srv.stats()["completed"] == 30
        with pytest.raises(S.StarkError):
            srv.submit_ds_ksk(keys[0], b"\0" * 32, b"\1" * 32, deadline=-1)


def test_submit_errors_stay_with_their_thread():
    """two threads submitting different bad requests at once each read their
    own error text"""
    bad = _bad_requests()
    errs = {0: [], 1: []}

    def go(t, srv):
        _, submit, why = bad[0] if t == 0 else bad[3]
        for _ in range(200):
            try:
                submit(srv)
            except S.StarkError as e:
                errs[t].append(why in str(e))

    with ProvingServer(workers=1, hbm_budget=1 << 30, dry_run=True) as srv:
        th = [threading.Thread(target=go, args=(t, srv)) for t in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join()
    assert len(errs[0]) == len(errs[1]) == 200
    assert all(errs[0]) and all(errs[1])


def test_signed_zone_requests_pass_the_submit_checks():
    """the serving workload's zone: its type / algorithm mix, and the
    requests ZoneSigner makes are accepted by the submit-time checks"""
    from collections import Counter

    from stark_swarm_amd.zone import ZoneSigner, synthetic_rrset_zone
    zone = synthetic_rrset_zone(2000, 11)
    t = Counter(r.rtype for r in zone)
    a = Counter(r.alg for r in zone)
    assert abs(t["A"] / 2000 - 0.60) < 0.05 and abs(t["AAAA"] / 2000 - 0.25) < 0.05
    assert abs(a[15] / 2000 - 0.50) < 0.05 and abs(a[13] / 2000 - 0.30) < 0.05
    for r in zone:
        labels = r.owner.rstrip(".").split(".")
        assert 2 <= len(labels) <= 4 and all(3 <= len(x) <= 12 and x.islower() for x in labels)
        if r.rtype == "TXT":
            assert 16 <= len(r.rdata) <= 255
    signer = ZoneSigner(12, keys_per_alg=(2, 2, 1))
    picks = [next(r for r in zone if r.alg == alg) for alg in (15, 13, 8)]
    with ProvingServer(workers=1, hbm_budget=1 << 30, dry_run=True) as srv:
        for r in picks:
            q = signer.request(r)
            if q[0] == "ed25519":
                assert q[4] == S.ed25519_record_pk_hash(q[1], q[2], r.message())
                rid = srv.submit_ed25519(q[1], q[2], q[3], q[4], bytes(32))
            elif q[0] == "p256":
                rid = srv.submit_p256(q[1])
            else:
                assert pow(q[1].s, 65537, q[1].n) == q[1].em
                rid = srv.submit_rsa(q[1], q[2])
            assert srv.wait(rid, timeout=30).status == 0
