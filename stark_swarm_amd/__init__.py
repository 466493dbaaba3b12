"""stark_swarm_amd -- MI355X (gfx950) proving backend for stark-swarm's prover core.

Host-side mirror of the reference's prover API for the data-parallel path,
calling the hand-written HIP kernels in libstark_gpu.so through its C ABI
(include/stark_gpu.h). Names, argument meaning and error behaviour follow the
Rust functions they replace:

  lde_trace_columns        crates/deep_ali/src/trace_import.rs:251-281
  deep_ali_merge_general   crates/deep_ali/src/lib.rs:343-437
  deep_fri_prove           crates/deep_ali/src/fri.rs:1900-2055
  DeepFriParams, FriDomain crates/deep_ali/src/fri.rs:160-175,1303-1349
  make_schedule, build_params, prove_inner_shard, prove_outer_rollup
                           crates/swarm-dns/src/prover.rs:53-304
  prove_ds_ksk_binding     crates/swarm-dns/src/prover.rs:446-585
  prove_nsec3_completeness, Nsec3Record, Nsec3Output
                           crates/swarm-dns/src/prover.rs:321-446
  build_sha256_trace_multi crates/deep_ali/src/sha256_air.rs:429-479
  deep_ali_merge_sha256    crates/deep_ali/src/lib.rs:453-545
  build_sha512_trace_multi crates/deep_ali/src/sha512_air.rs:551-573
  deep_ali_merge_sha512    crates/deep_ali/src/lib.rs:561-653
  fill_rsa_exp_multirow, fill_rsa_stacked
                           crates/deep_ali/src/rsa2048_exp_air.rs:148-202,
                           rsa2048_stacked_air.rs:64-78
  deep_ali_merge_rsa_exp_multirow_streaming, deep_ali_merge_rsa_stacked_streaming
                           crates/deep_ali/src/lib.rs:1827-2049
  emsa_pkcs1_v1_5_encode_sha256
                           crates/deep_ali/src/rsa2048.rs
  prove_rsa_exp_batch      crates/swarm-dns/examples/mixed_algorithm_zone_rollup.rs:324-395
  fill_ecdsa_double_multirow, P256Chains
                           crates/deep_ali/src/p256_ecdsa_double_multirow_air.rs:134-222
  deep_ali_merge_ecdsa_double_multirow_streaming
                           crates/deep_ali/src/lib.rs:1421-1540
  prove_p256_chains_batch  crates/swarm-dns/examples/p256_full_ecdsa_stark_bench.rs:133-240
  prove_record_ecdsa_batch, EcdsaRecord
                           crates/swarm-dns/examples/mixed_algorithm_zone_rollup.rs:193-322
  DnsRecord                crates/swarm-dns/src/dns.rs:24-89
  deep_fri_verify          crates/deep_ali/src/fri.rs:2136-2710 (host code)
  MerkleChannelCfg, merkle_open, verify_opening, compute_leaf_hash
                           crates/merkle/src/lib.rs:117-128,478-512,754-812
  shard_fs_binding         crates/swarm-proto/src/messages.rs:162-176

There is no CPU fallback: importing works without a GPU, but every call that
computes needs the built library and a visible gfx950 device, and raises
otherwise.
"""
from .api import (  # noqa: F401
    AirType,
    CompositionInfo,
    Context,
    DeepFriParams,
    DnsRecord,
    FriDomain,
    InnerShardOutput,
    LdtMode,
    OuterRollupOutput,
    PackedShards,
    DsKskOutput,
    StarkError,
    BLOWUP,
    NUM_QUERIES,
    SEED_Z,
    PHI_RAYON,
    PHI_ONE_THREAD,
    PHI_SEQUENTIAL,
    build_params,
    deep_fri_verify,
    deep_fri_verify_batch,
    MerkleChannelCfg,
    merkle_open,
    verify_opening,
    compute_leaf_hash,
    set_autotune_default,
    autotune_report,
    deep_ali_merge_general,
    deep_ali_merge_sha256,
    build_sha256_trace_multi,
    deep_ali_merge_sha512,
    build_sha512_trace_multi,
    merge_sha_dev,
    merge_sha_from_trace_dev,
    prove_sha_batch,
    Nsec3Record,
    Nsec3Output,
    prove_nsec3_completeness,
    prove_nsec3_completeness_batch,
    RsaStackedRecord,
    RsaProof,
    fill_rsa_exp_multirow,
    fill_rsa_stacked,
    deep_ali_merge_rsa_exp_multirow_streaming,
    deep_ali_merge_rsa_stacked_streaming,
    emsa_pkcs1_v1_5_encode_sha256,
    rsa_record_pk_hash,
    prove_rsa_exp_batch,
    P256Chains,
    P256Proof,
    EcdsaRecord,
    fill_ecdsa_double_multirow,
    deep_ali_merge_ecdsa_double_multirow_streaming,
    prove_p256_chains_batch,
    prove_record_ecdsa_batch,
    ecdsa_record_pk_hash,
    ED25519_K_SCALAR,
    VerifyAirShapeV16,
    ZskKskOutput,
    verify_air_layout_v16,
    ed25519_verify,
    ed25519_sign,
    fill_verify_air_v16,
    deep_ali_merge_ed25519_verify_streaming,
    prove_zsk_ksk_binding_v2,
    prove_zsk_ksk_binding_v2_batch,
    ed25519_record_pk_hash,
    record_pi_hash,
    deep_fri_prove,
    deep_fri_prove_batch,
    default_context,
    lde_trace_columns,
    lib,
    make_schedule,
    merkle_commit,
    ntt,
    prove_inner_shard,
    prove_inner_shard_batch,
    prove_outer_rollup,
    prove_ds_ksk_binding,
    prove_ds_ksk_binding_batch,
    prover_core,
    shard_fs_binding,
)
