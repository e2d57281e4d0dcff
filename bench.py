#!/usr/bin/env python3
"""Batched KEM throughput on one MI355X.

Headline workload (the metric's single-GPU configuration): Saber l=3, batches
of 4096 instances; one step = 4096 keygens + 4096 encapsulations + 4096
decapsulations, inputs resident in HBM before the timed region, outputs left
in HBM.  value = KEM operations per second = 3 * 4096 * steps / seconds.

The steps run a rolling keygen -> encaps -> decaps pipeline over three batch
slots: step r generates the keys of slot r mod 3, encapsulates to the keys of
slot r - 1 (generated by the previous step) and decapsulates the ciphertexts
of slot r - 2 (encapsulated by the previous step, with the secret keys of two
steps before).  Each batch therefore goes through all three ops, one per
step, and every op consumes what an earlier op produced; within a step the
three ops are independent and run as two launches (dpkem_roll_*): all hash
chains of the three ops plus the closing hashes of the previous steps'
batches (keygen's H(pk), encaps' K) in one, all products in the other.
The autotuner also tries the overlapped form over six slots (encaps to the
keys of two steps earlier, decaps of the ciphertexts of two steps earlier),
in which each step is one launch holding its hash chains and the previous
step's products.  Every op recomputes everything from its own inputs
(encaps expands the public matrix from pk, decaps from sk), as separate
parties would.

Reported beside it in the same JSON line:
  saber_chained  the same ops with the keygen -> encaps -> decaps chain of ONE
                 batch inside each step (pipelined launch sequence)
  frodo976       FrodoKEM-976-SHAKE, batch 8192, keygen + encaps + decaps,
                 rolling steps as above (six launches: decaps' mu', every
                 serial SHAKE chain of the three ops in one launch, the three
                 fused products, decaps' last ss block); `chained` the plan
                 with one batch's chain inside each step
  frodo640 / frodo1344 / frodo976_aes / lightsaber / firesaber
                 the other parameter sets, rolling steps (chained_ops_per_s:
                 the chained form)
  plain          the one-workgroup-per-KEM comparator form on the same inputs
  latency        p50 / p99 of the per-step wall time (ms), and of a batch's
                 three steps from keygen to decaps
  roofline       dominant kernel of the headline step: the plan's own kernels
                 launched without a graph (same launches, streams and
                 kernels), each bracketed by HIP events on its stream;
                 kernels_ms lists them all, eager_kernels_ms the op-by-op
                 launches
  mixed          16,384 instances in random Frodo / Saber batches of U[1,4096],
                 ops 1:4:4, one autotuned graph vs plain per-batch launches
  serving        Poisson arrivals into the continuous-batching server
                 (70% Saber, ops 1:4:4) vs one request per launch
  saber_unpipelined
                 the chained step as three dependent plan tasks
  saber_shared_matrix_chain
                 the chained steps with the plan's opt-in matrix sharing
                 (encaps / decaps reuse keygen's A); not the headline value
  cpu_baseline   the CPU oracle on the host cores, on a bounded sample
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VALU_PEAK = 256 * 4 * 16 * 2.4e9  # lane-ops / s: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz
KECCAK_LANE_OPS = 4176            # int32 VALU lane-ops per Keccak-f[1600] (gfx950 floor, bitop3)
HBM_PEAK = 8.0e12

# Algorithmic work per instance for each kernel: (Keccak permutations, MACs).
# Only the MACs of the kernels in VALU_MACS_PER_OP run on the VALU (2 per
# v_dot2 lane-op in sab_dec_inner, 1 per multiply-add in frodo_dec_mu) and
# count against the VALU roofline; every other product (the Saber products
# and the chain's inner products, Frodo's A S and S' A) runs on int8 MFMA and
# is not counted.
KERNEL_WORK = {
    "sab_keygen_hash1": (1 + 23 + 5, 0),
    "sab_poly<kSabKeygen>": (0, 9 * 65536),
    "sab_keygen_hash2": (8, 0),
    "sab_enc_hash1": (23 + 1 + 8 + 1 + 5, 0),
    "sab_poly<kSabEnc>": (0, 12 * 65536),
    "sab_hash2": (9 + 1, 0),
    "sab_dec_inner": (0, 3 * 65536),
    "sab_dec_hash1": (23 + 1 + 5, 0),
    "sab_poly<kSabReenc>": (0, 12 * 65536),
    "frodo_keygen_pre": (1 + 230, 0),
    "frodo_as_fused": (976 * 12, 0),
    "frodo_keygen_post": (115, 0),
    "frodo_encdec_pre<false>": (115 + 1 + 231, 0),
    "frodo_encdec_pre<true>": (1 + 231, 0),
    "frodo_sa_fused<false>": (976 * 12, 0),
    "frodo_sa_fused<true>": (976 * 12, 0),
    "frodo_ss": (116, 0),
    "frodo_dec_mu": (0, 64 * 976),
    # Frodo-640 (SHAKE128 chains, rate 168 B) and Frodo-1344 (SHAKE256, 136 B):
    # A rows take 8 / 16 SHAKE128 blocks; noise streams 32n + 128 bytes
    # pipelined keygen -> encaps -> decaps chains (the plan kernels the
    # headline replays): stages of different ops sharing a launch
    "sab_pipe_x0": (1 + 23 + 5 + 23 + 23, 0),          # keygen seedA, GenMatrix, GenSecret; encaps' / decaps' GenMatrix
    "sab_pipe_x1": (8 + 1 + 1 + 5 + 8, 0),             # encaps' H(pk), m, G, GenSecret; keygen's H(pk)
    "sab_pipe_x2": (10 + 9 + 1 + 5 + 2, 3 * 65536),    # encaps' K; decaps' H(ct), G, GenSecret, both K; inner product
    "sab_reenc": (0, 12 * 65536),                        # re-encryption product, compare, select of decaps' K
    # independent keygen, encaps and decaps tasks in one mixed launch sequence
    # (per instance of each op): all hash chains, all products, closing hashes
    "sab_mix_x0": ((1 + 23 + 5) + (23 + 1 + 8 + 1 + 5) + (23 + 9 + 2 + 1 + 5), 3 * 65536),
    "sab_mix_poly": (0, (9 + 12 + 12) * 65536),
    "sab_mix_x1": (8 + 9 + 1, 0),
    "sab_mix_x0,l2": ((1 + 10 + 4) + (10 + 1 + 5 + 1 + 4) + (10 + 6 + 2 + 1 + 4), 2 * 65536),
    "sab_mix_poly,l2": (0, (4 + 6 + 6) * 65536),
    "sab_mix_x1,l2": (5 + 6 + 1, 0),
    "sab_mix_x0,l4": ((1 + 40 + 5) + (40 + 1 + 10 + 1 + 5) + (40 + 11 + 2 + 1 + 5), 4 * 65536),
    "sab_mix_poly,l4": (0, (16 + 20 + 20) * 65536),
    "sab_mix_x1,l4": (10 + 11 + 1, 0),
    # rolling pipeline step: the mixed hash chains plus the previous batches'
    # closing hashes (keygen's H(pk), encaps' H(ct) and K); products as mixed
    "sab_roll_x0": ((1 + 23 + 5) + (23 + 1 + 8 + 1 + 5) + (23 + 9 + 2 + 1 + 5) + (8 + 9 + 1), 3 * 65536),
    "sab_roll_x0,l2": ((1 + 10 + 4) + (10 + 1 + 5 + 1 + 4) + (10 + 6 + 2 + 1 + 4) + (5 + 6 + 1), 2 * 65536),
    "sab_roll_x0,l4": ((1 + 40 + 5) + (40 + 1 + 10 + 1 + 5) + (40 + 11 + 2 + 1 + 5) + (10 + 11 + 1), 4 * 65536),
    # overlapped rolling step in one launch: this step's chains (as
    # sab_roll_x0) and the previous step's products (as sab_mix_poly)
    "sab_roll_fused": ((1 + 23 + 5) + (23 + 1 + 8 + 1 + 5) + (23 + 9 + 2 + 1 + 5) + (8 + 9 + 1),
                       (3 + 9 + 12 + 12) * 65536),
    "sab_roll_fused,l2": ((1 + 10 + 4) + (10 + 1 + 5 + 1 + 4) + (10 + 6 + 2 + 1 + 4) + (5 + 6 + 1),
                          (2 + 4 + 6 + 6) * 65536),
    "sab_roll_fused,l4": ((1 + 40 + 5) + (40 + 1 + 10 + 1 + 5) + (40 + 11 + 2 + 1 + 5) + (10 + 11 + 1),
                          (4 + 16 + 20 + 20) * 65536),
    "sab_pipe_x0,l2": (1 + 10 + 4 + 10 + 10, 0),
    "sab_pipe_x1,l2": (5 + 1 + 1 + 4 + 5, 0),
    "sab_pipe_x2,l2": (7 + 6 + 1 + 4 + 2, 2 * 65536),
    "sab_reenc,l2": (0, 6 * 65536),
    "sab_pipe_x0,l4": (1 + 40 + 5 + 40 + 40, 0),
    "sab_pipe_x1,l4": (10 + 1 + 1 + 5 + 10, 0),
    "sab_pipe_x2,l4": (12 + 11 + 1 + 5 + 2, 4 * 65536),
    "sab_reenc,l4": (0, 20 * 65536),
    "sab_poly<kSabKeygen>,l2": (0, 4 * 65536),
    "sab_poly<kSabEnc>,l2": (0, 6 * 65536),
    "sab_poly<kSabReenc>,l2": (0, 6 * 65536),
    "sab_poly<kSabKeygen>,l4": (0, 16 * 65536),
    "sab_poly<kSabEnc>,l4": (0, 20 * 65536),
    "sab_poly<kSabReenc>,l4": (0, 20 * 65536),
    "frodo_pipe_x1": (115 + 347, 0),                    # keygen's pkh beside encaps' chain
    "frodo_pipe_x2": (116 + 232 + 115, 0),              # encaps' ss, decaps' chain, decaps' ct absorb
    "frodo_pipe_ssd": (1, 0),
    # rolling step: keygen's seedA + noise chain, keygen's pkh (closing hash),
    # encaps' pkh + G2 + noise chain, encaps' ss (closing hash), decaps' G2 +
    # noise chain, decaps' ct absorb
    "frodo_roll_x0": (1 + 230 + 115 + 347 + 116 + 232 + 115, 0),
    "frodo_roll_x0,640": (1 + 122 + 58 + 182 + 58 + 124 + 57, 0),
    "frodo_roll_x0,1344": (1 + 317 + 159 + 478 + 160 + 319 + 159, 0),
    "frodo_pipe_x1,640": (58 + 182, 0),
    "frodo_pipe_x2,640": (58 + 124 + 57, 0),
    "frodo_pipe_ssd,640": (1, 0),
    "frodo_pipe_x1,1344": (159 + 478, 0),
    "frodo_pipe_x2,1344": (160 + 319 + 159, 0),
    "frodo_pipe_ssd,1344": (1, 0),
    "frodo_keygen_pre,640": (1 + 122, 0),
    "frodo_as_fused,640": (640 * 8, 0),
    "frodo_keygen_post,640": (58, 0),
    "frodo_encdec_pre<false>,640": (58 + 1 + 123, 0),
    "frodo_encdec_pre<true>,640": (1 + 123, 0),
    "frodo_sa_fused<false>,640": (640 * 8, 0),
    "frodo_sa_fused<true>,640": (640 * 8, 0),
    "frodo_ss,640": (58, 0),
    "frodo_dec_mu,640": (0, 64 * 640),
    "frodo_keygen_pre,1344": (1 + 317, 0),
    "frodo_as_fused,1344": (1344 * 16, 0),
    "frodo_keygen_post,1344": (159, 0),
    "frodo_encdec_pre<false>,1344": (159 + 1 + 318, 0),
    "frodo_encdec_pre<true>,1344": (1 + 318, 0),
    "frodo_sa_fused<false>,1344": (1344 * 16, 0),
    "frodo_sa_fused<true>,1344": (1344 * 16, 0),
    "frodo_ss,1344": (160, 0),
    "frodo_dec_mu,1344": (0, 64 * 1344),
    # decaps' chain beside ss's ct absorb, then ss's last block
    "frodo_dec_pre": (1 + 231 + 115, 0),
    "frodo_dec_pre,640": (1 + 123 + 57, 0),
    "frodo_dec_pre,1344": (1 + 318 + 159, 0),
    # multi-task launch sequences (independent tasks of one op in one launch each)
    "frodo_mt_pre<keygen>": (1 + 230, 0),
    "frodo_mt_post<keygen>": (115, 0),
    "frodo_mt_pre<enc>": (115 + 1 + 231, 0),
    "frodo_mt_post<enc>": (116, 0),
    "frodo_mt_dec_pre": (1 + 231 + 115, 0),
    "frodo_mt_post<dec>": (1, 0),
}


VALU_MACS_PER_OP = {"sab_dec_inner": 2.0, "frodo_dec_mu": 1.0}


# HBM bytes (read, write) per launch at the profiled batch sizes (Saber family
# 4096, Frodo 8192 instances), from two rocprofv3 --pmc passes over
#   bench.py --schedule 1x1 --calibrate-copy --skip-mixed --skip-serving --skip-per-op --cpu-seconds 0
#            --frodo-steps 2 --plain-steps 1
# (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by what the same
# counter reads for an elementwise kernel over 256 MiB in / 256 MiB out:
# FETCH_SIZE x 1.99970, WRITE_SIZE x 1.00000.  Counters cannot be read live.
# The sab_roll_fused rows come from the same two passes over the Saber-family
# part of that command (--skip-frodo; scale 1.99975 / 1.00000), in which the
# unchanged Saber kernels (sab_roll_x0, sab_pipe_x0, sab_poly, ...) read
# within 4% of their rows here.
MEASURED_HBM_BYTES = {
    ('frodo_as_fused', 8192): (129247208, 256245767),
    ('frodo_as_fused,1344', 8192): (177458279, 353382438),
    ('frodo_as_fused,640', 8192): (85214814, 157548544),
    ('frodo_as_fused<aes>', 8192): (129303548, 256245941),
    ('frodo_dec_mu', 8192): (256060732, 262144),
    ('frodo_dec_mu,1344', 8192): (358477781, 262144),
    ('frodo_dec_mu,640', 8192): (165536658, 262144),
    ('frodo_dec_pre', 8192): (130915869, 130395792),
    ('frodo_encdec_pre<false>', 8192): (75028256, 128713549),
    ('frodo_keygen_post', 8192): (129167291, 262144),
    ('frodo_keygen_pre', 8192): (1430538, 256964056),
    ('frodo_keygen_pre,1344', 8192): (1547183, 353243891),
    ('frodo_keygen_pre,640', 8192): (1336732, 168558668),
    ('frodo_pipe_ssd', 8192): (4392128, 196608),
    ('frodo_pipe_ssd,1344', 8192): (4518336, 262144),
    ('frodo_pipe_ssd,640', 8192): (5536955, 131072),
    ('frodo_pipe_x1', 8192): (130434257, 128975605),
    ('frodo_pipe_x1,1344', 8192): (178671232, 177341228),
    ('frodo_pipe_x1,640', 8192): (81071137, 84934656),
    ('frodo_pipe_x2', 8192): (131991284, 130652792),
    ('frodo_pipe_x2,1344', 8192): (180311381, 178919302),
    ('frodo_pipe_x2,640', 8192): (83639043, 86442995),
    ('frodo_roll_x0', 8192): (520002953, 523797063),
    ('frodo_roll_x0,1344', 8192): (714049588, 714828370),
    ('frodo_roll_x0,640', 8192): (324252262, 341495698),
    ('frodo_sa_fused<false,aes>', 8192): (259476028, 128974848),
    ('frodo_sa_fused<false>', 8192): (260449633, 128974848),
    ('frodo_sa_fused<false>,1344', 8192): (356734065, 177209344),
    ('frodo_sa_fused<false>,640', 8192): (166974056, 79822848),
    ('frodo_sa_fused<true,aes>', 8192): (388306616, 262144),
    ('frodo_sa_fused<true>', 8192): (389398035, 262144),
    ('frodo_sa_fused<true>,1344', 8192): (533916936, 265416),
    ('frodo_sa_fused<true>,640', 8192): (247572911, 262285),
    ('frodo_sa_grouped<false>', 8192): (129748323, 128974848),
    ('frodo_sa_grouped<true>', 8192): (258629834, 65536),
    ('frodo_ss', 8192): (129686234, 196608),
    ('plain::copy_k', 8192): (55916302, 55462532),
    ('plain::fr_as', 8192): (16571141997, 127926272),
    ('plain::fr_bs_decode', 8192): (270769823, 393216),
    ('plain::fr_gen_a', 8192): (1385475, 19994088352),
    ('plain::fr_le_load', 8192): (128979031, 127926272),
    ('plain::fr_le_store', 8192): (127929462, 128189376),
    ('plain::fr_pack', 8192): (77188230, 77202233),
    ('plain::fr_sa', 8192): (18570791927, 131428000),
    ('plain::fr_sample', 8192): (256537649, 256551594),
    ('plain::fr_sb', 8192): (265728025, 1048576),
    ('plain::fr_select', 8192): (1487647, 393216),
    ('plain::fr_unpack', 8192): (96706278, 96206848),
    ('plain::sab_cbd', 8192): (3166753, 6291456),
    ('plain::sab_cm', 8192): (2642544, 524288),
    ('plain::sab_kdf', 8192): (724882, 131072),
    ('plain::sab_msg', 8192): (2900553, 131072),
    ('plain::sab_pack13', 8192): (6299719, 5111808),
    ('plain::sab_polymul', 8192): (18953423, 4194304),
    ('plain::sab_round_pack10', 8192): (6314053, 3932160),
    ('plain::sab_unpack10', 8192): (4443829, 6291456),
    ('plain::sab_unpack13', 8192): (13141261, 15728640),
    ('plain::sha3_256_k', 8192): (4186503, 131072),
    ('plain::sha3_512_64_k', 8192): (327630, 262144),
    ('plain::shake128_seed32', 8192): (469944, 131072),
    ('plain::shake256_dom_seed', 8192): (664859, 255852896),
    ('plain::shake256_k', 8192): (612643, 262144),
    ('plain::verify_k', 8192): (257935283, 262144),
    ('sab_dec_hash1', 4096): (1177522, 3407872),
    ('sab_dec_inner', 4096): (9888297, 131072),
    ('sab_enc_hash1', 4096): (4932886, 3538944),
    ('sab_hash2', 4096): (5234920, 131072),
    ('sab_keygen_hash1', 4096): (1289299, 19007307),
    ('sab_keygen_hash2', 4096): (4251005, 131072),
    ('sab_mix_poly', 4096): (72292904, 17590032),
    ('sab_mix_poly,l2', 4096): (39602419, 11801380),
    ('sab_mix_poly,l4', 4096): (115965620, 28766883),
    ('sab_mix_x0', 4096): (5775069, 19113673),
    ('sab_mix_x1', 4096): (4965184, 131072),
    ('sab_pipe_x0', 4096): (1288494, 49623287),
    ('sab_pipe_x0,l2', 4096): (1296772, 23486277),
    ('sab_pipe_x0,l4', 4096): (1304083, 85387969),
    ('sab_pipe_x1', 4096): (36087353, 34340864),
    ('sab_pipe_x1,l2', 4096): (17773024, 16777216),
    ('sab_pipe_x1,l4', 4096): (62046729, 58195968),
    ('sab_pipe_x2', 4096): (12584081, 3932160),
    ('sab_pipe_x2,l2', 4096): (9465449, 3407872),
    ('sab_pipe_x2,l4', 4096): (15914906, 3932160),
    ('sab_poly<kSabEnc>', 4096): (23843648, 4456448),
    ('sab_poly<kSabEnc>,l2', 4096): (13075527, 3014656),
    ('sab_poly<kSabEnc>,l4', 4096): (36688325, 6029312),
    ('sab_poly<kSabKeygen>', 4096): (18983120, 12979696),
    ('sab_poly<kSabKeygen>,l2', 4096): (9526879, 8650863),
    ('sab_poly<kSabKeygen>,l4', 4096): (30524921, 17328898),
    ('sab_poly<kSabReenc>', 4096): (28974514, 131072),
    ('sab_reenc', 4096): (29471737, 131072),
    ('sab_reenc,l2', 4096): (17010164, 131072),
    ('sab_reenc,l4', 4096): (48755660, 5111808),
    ('sab_roll_x0', 4096): (18634235, 57374980),
    ('sab_roll_x0,l2', 4096): (13111599, 30149308),
    ('sab_roll_x0,l4', 4096): (26774250, 93216428),
    ('sab_roll_fused', 4096): (92342679, 76169056),
    ('sab_roll_fused,l2', 4096): (53018414, 42397696),
    ('sab_roll_fused,l4', 4096): (140606620, 118882368),
}


# The table above holds only while the kernels are the ones it was read
# from: each family's device sources are hashed here and compared with the
# hash recorded beside the table; a mismatch reports traffic as unmeasured.
KERNEL_SOURCES = {
    "saber": ("saber.hip", "saber_poly.hpp", "keccak2.hpp", "keccak.hpp", "sponge.hpp", "common.hpp"),
    "frodo": ("frodo.hip", "frodo_sampler.hpp", "keccak2.hpp", "keccak.hpp", "sponge.hpp", "common.hpp"),
}
MEASURED_HBM_SOURCE_SHA = {"saber": "a3bbf5e3e97a1972", "frodo": "7a4099279b76c56d"}


def kernel_family(kernel: str) -> str:
    return "frodo" if kernel.startswith("frodo") else "saber"


def source_sha(family: str) -> str:
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES[family]:
        with open(os.path.join(ROOT, "dpcrypto_amd", "csrc", f), "rb") as fh:
            h.update(f.encode() + b"\0" + fh.read())
    return h.hexdigest()[:16]


def measured_hbm(kernel: str, n: int):
    """(read, write) HBM bytes per launch from the table, or None when the
    table has no entry or was read from other kernel sources."""
    fam = kernel_family(kernel)
    if MEASURED_HBM_SOURCE_SHA.get(fam) != source_sha(fam):
        return None
    return MEASURED_HBM_BYTES.get((kernel, n))


def lane_ops(kernel: str, n: int) -> float:
    perms, macs = KERNEL_WORK.get(kernel, (0, 0))
    per_op = VALU_MACS_PER_OP.get(kernel.split(",")[0])
    return n * (perms * KECCAK_LANE_OPS + (macs / per_op if per_op else 0.0))


def pct(vals, q):
    v = sorted(vals)
    if not v:
        return None
    k = (len(v) - 1) * q
    lo = int(k)
    hi = min(lo + 1, len(v) - 1)
    return v[lo] + (v[hi] - v[lo]) * (k - lo)


def run_workload(torch, eng, scheme, n, steps, warmup, seed, profile=False):
    """Time `steps` keygen+enc+dec steps over n device-resident instances."""
    import numpy as np
    import dpcrypto_amd as D
    s = D.SIZES[scheme]
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda:0")
    kc = torch.from_numpy(rng.integers(0, 256, (n, s["kcoins"]), dtype=np.uint8)).to(dev)
    ec = torch.from_numpy(rng.integers(0, 256, (n, s["ecoins"]), dtype=np.uint8)).to(dev)
    pk = torch.empty((n, s["pk"]), dtype=torch.uint8, device=dev)
    sk = torch.empty((n, s["sk"]), dtype=torch.uint8, device=dev)
    ct = torch.empty((n, s["ct"]), dtype=torch.uint8, device=dev)
    ss = torch.empty((n, s["ss"]), dtype=torch.uint8, device=dev)
    ss2 = torch.empty((n, s["ss"]), dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream(dev)
    eng.set_stream(stream.cuda_stream)

    def step():
        eng.keypair_into(scheme, n, pk, sk, kc)
        eng.enc_into(scheme, n, ct, ss, pk, ec)
        eng.dec_into(scheme, n, ss2, ct, sk)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    eng.kernel_stats(reset=True)
    eng.set_profiling(profile)
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    stops = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        starts[i].record(stream)
        step()
        stops[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    eng.set_profiling(False)
    stats = eng.kernel_stats(reset=True) if profile else {}
    per_step = [a.elapsed_time(b) for a, b in zip(starts, stops)]
    ok = bool(torch.equal(ss, ss2))
    return dict(wall_s=wall, per_step_ms=per_step, kernels=stats, roundtrip_ok=ok,
                ops_per_s=3 * n * steps / wall)


def chain_arrays(torch, scheme, n, seed):
    """Device-resident arrays of one keygen -> encaps -> decaps step over n
    instances: coins drawn from numpy's generator seeded with `seed` (as in
    every timed run), outputs left uninitialised."""
    import numpy as np
    import dpcrypto_amd as D
    s = D.SIZES[scheme]
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda:0")
    a = {"kc": torch.from_numpy(rng.integers(0, 256, (n, s["kcoins"]), dtype=np.uint8)).to(dev),
         "ec": torch.from_numpy(rng.integers(0, 256, (n, s["ecoins"]), dtype=np.uint8)).to(dev)}
    for name, w in (("pk", s["pk"]), ("sk", s["sk"]), ("ct", s["ct"]), ("ss", s["ss"]), ("ss2", s["ss"])):
        a[name] = torch.empty((n, w), dtype=torch.uint8, device=dev)
    return a


def build_chain_plan(eng, scheme, a, nstreams, chunks, knob=None, share_matrix=False, pipeline=True, graph=True):
    """The plan a timed step replays: keygen -> encaps -> decaps chained per
    instance over the arrays of chain_arrays.  knob is the scheme's
    launch-shape knob: the Saber product's instances per workgroup (1 / 2 /
    4), or the Frodo fused kernels' workgroup size (1024 / 512)."""
    import dpcrypto_amd as D
    saber_family = scheme in (D.SABER, D.LIGHTSABER, D.FIRESABER)
    eng.set_tuning(D.TUNE_SABER_POLY_IPW, knob if saber_family and knob else 1)
    eng.set_tuning(D.TUNE_FRODO_THREADS, knob if not saber_family and knob else 0)
    p = eng.plan(nstreams=nstreams, chunks=chunks, share_matrix=share_matrix, pipeline=pipeline, graph=graph)
    t0 = p.keypair(scheme, a["pk"], a["sk"], a["kc"])
    t1 = p.enc(scheme, a["ct"], a["ss"], a["pk"], a["ec"], after=t0)
    p.dec(scheme, a["ss2"], a["ct"], a["sk"], after=t1)
    return p.build()


def tune_chain_schedule(torch, eng, scheme, n, a, tune_budget_s=60.0, share_matrix=False, pipeline=True):
    """Autotuned (streams, chunks[, ipw]) schedule of the chain plan."""
    import dpcrypto_amd as D
    from dpcrypto_amd import autotune
    saber_family = scheme in (D.SABER, D.LIGHTSABER, D.FIRESABER)

    def build(ns, ch, knob=None):
        return build_chain_plan(eng, scheme, a, ns, ch, knob, share_matrix, pipeline)

    key = f"scheme{scheme}|kg-enc-dec|n={n}|share={int(share_matrix)}|pipe={int(pipeline)}|knob"
    if saber_family:
        return autotune.tune(torch, eng, key, build, use_cache=False, budget_s=tune_budget_s,
                             candidates=[c + (1,) for c in autotune.DEFAULT_CANDIDATES],
                             refine=lambda b: [(b[0], b[1], 2), (b[0], b[1], 4)])
    if scheme in (D.FRODO976, D.FRODO976_AES):
        return autotune.tune(torch, eng, key, build, use_cache=False, budget_s=tune_budget_s,
                             candidates=[c + (512,) for c in autotune.DEFAULT_CANDIDATES],
                             refine=lambda b: [(b[0], b[1], 1024)])
    return autotune.tune(torch, eng, key, build, use_cache=False, budget_s=tune_budget_s)


def run_plan_workload(torch, eng, scheme, n, steps, warmup, seed, tune_budget_s=60.0, share_matrix=False,
                      schedule=None, pipeline=True):
    """Time `steps` keygen+enc+dec steps, each one replay of a graph-captured
    plan (keygen -> enc -> dec chained per instance chunk); the plan's
    (streams, chunks) schedule is picked by the autotuner first."""
    a = chain_arrays(torch, scheme, n, seed)
    stream = torch.cuda.current_stream(torch.device("cuda:0"))
    eng.set_stream(stream.cuda_stream)
    if schedule:
        best, tried = tuple(schedule), {}
    else:
        best, tried = tune_chain_schedule(torch, eng, scheme, n, a, tune_budget_s, share_matrix, pipeline)
    plan = build_chain_plan(eng, scheme, a, *best, share_matrix=share_matrix, pipeline=pipeline)
    knob_name = "saber_poly_ipw" if scheme in (1, 3, 4) else "frodo_threads"
    for _ in range(warmup):
        plan.launch()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    stops = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        starts[i].record(stream)
        plan.launch()
        stops[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per_step = [x.elapsed_time(y) for x, y in zip(starts, stops)]
    ok = bool(torch.equal(a["ss"], a["ss2"]))
    plan.close()
    return dict(wall_s=wall, per_step_ms=per_step, roundtrip_ok=ok, ops_per_s=3 * n * steps / wall,
                schedule={"nstreams": best[0], "chunks": best[1],
                          **({knob_name: best[2]} if len(best) > 2 else {}),
                          "tuning_ms": {"x".join(map(str, c)): round(v, 4) for c, v in tried.items()}})


def rolling_slots(torch, scheme, n, seed, count=3):
    """`count` batch slots of device-resident arrays (chain_arrays) for the
    rolling keygen -> encaps -> decaps pipeline (3, or 6 for the overlapped
    form)."""
    return [chain_arrays(torch, scheme, n, seed + 7919 * k) for k in range(count)]


def build_roll(eng, scheme, slots, ipw=1, graph=True, threads=0, overlap=False):
    """The rolling pipeline (dpkem_roll_*) over the slots: step k generates
    the keys of slot k % 3, encapsulates to the keys of slot (k - 1) % 3
    (generated by step k - 1) and decapsulates the ciphertexts of slot
    (k - 2) % 3 (encapsulated by step k - 1), and finishes the previous
    steps' batches (keygen's H(pk), encaps' K) in the same launches.
    overlap: the six-slot form (slots k, k - 2, k - 4 mod 6; one launch per
    step with its chains and the previous step's products).  ipw: the Saber product's
    instances per workgroup; threads: the Frodo fused kernels' workgroup
    size (0 = the set's default)."""
    import dpcrypto_amd as D
    eng.set_tuning(D.TUNE_SABER_POLY_IPW, ipw)
    eng.set_tuning(D.TUNE_FRODO_THREADS, threads)
    r = eng.roll(scheme, slots[0]["kc"].shape[0])
    for k, sl in enumerate(slots[:6 if overlap else 3]):
        r.bind(k, sl["kc"], sl["ec"], sl["pk"], sl["sk"], sl["ct"], sl["ss"], sl["ss2"])
    return r.build(graph=graph and not overlap, overlap=overlap)


def roll_shapes(scheme):
    """Launch-shape candidates of a rolling pipeline: for the Saber family
    (overlapped form, product instances per workgroup), for FrodoKEM the
    fused kernels' workgroup size (FrodoKEM-976: the default 512 lanes in
    two passes over A's rows, or 1024 in one)."""
    import dpcrypto_amd as D
    if scheme in (D.SABER, D.LIGHTSABER, D.FIRESABER):
        return [(False, c) for c in (1, 2, 4)] + [(True, 4)]  # the overlapped form's products: 4 per workgroup
    return [0, 1024] if scheme in (D.FRODO976, D.FRODO976_AES) else [0]


def build_roll_shape(eng, scheme, slots, shape, graph=True):
    import dpcrypto_amd as D
    if scheme in (D.SABER, D.LIGHTSABER, D.FIRESABER):
        overlap, ipw = shape
        return build_roll(eng, scheme, slots, ipw, graph=graph, overlap=overlap)
    return build_roll(eng, scheme, slots, 1, graph=graph, threads=shape)


def run_rolling_workload(torch, eng, scheme, n, steps, warmup, seed, ipw=None):
    """Time `steps` steps of the rolling pipeline (graph replays, or direct
    launches, one per step in the overlapped form).  The launch shape
    (roll_shapes; `ipw` fixes the Saber product's and keeps the three-slot
    form) is picked first by timing two cycles of steps.  Afterwards the
    batches still in flight are finished and every slot must hold ss == ss2
    (encaps' and decaps' shared secrets)."""
    import statistics
    import dpcrypto_amd as D
    saber = scheme in (D.SABER, D.LIGHTSABER, D.FIRESABER)
    shapes = roll_shapes(scheme) if ipw is None else [(False, ipw)]
    slots = rolling_slots(torch, scheme, n, seed, 6 if any(isinstance(s, tuple) and s[0] for s in shapes) else 3)
    stream = torch.cuda.current_stream(torch.device("cuda:0"))
    eng.set_stream(stream.cuda_stream)

    def cycle_ms(r, reps, per):
        ts = []
        for _ in range(reps):
            x, y = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            x.record(stream)
            for _ in range(per):
                r.step()
            y.record(stream)
            torch.cuda.synchronize()
            ts.append(x.elapsed_time(y) / per)
        return statistics.median(ts)

    def lag(shape):  # steps from a batch's keygen step to the step that writes its last output
        return 5 if isinstance(shape, tuple) and shape[0] else 2

    tried = {}
    for c in shapes:
        r = build_roll_shape(eng, scheme, slots, c)
        if len(shapes) > 1:
            per = 6 if lag(c) == 4 else 3
            cycle_ms(r, 1 if not saber else 2, per)
            tried[c] = cycle_ms(r, 2 if not saber else 4, per)
        r.close()
    best = min(tried, key=tried.get) if tried else shapes[0]
    r = build_roll_shape(eng, scheme, slots, best)
    for _ in range(max(warmup, lag(best) + 1)):  # every slot through keygen, encaps and decaps before timing
        r.step()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    stops = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        starts[i].record(stream)
        r.step()
        stops[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    for _ in range(lag(best) + 1):
        r.step()
    torch.cuda.synchronize()
    per_step = [x.elapsed_time(y) for x, y in zip(starts, stops)]
    ok = all(bool(torch.equal(sl["ss"], sl["ss2"])) for sl in slots[:6 if lag(best) == 4 else 3])
    r.close()
    span = lag(best) + 1
    batch = [sum(per_step[i:i + span]) for i in range(max(1, len(per_step) - span + 1))]
    if saber:
        overlap, ipw_best = best
        sched = {"pipeline": ("dpkem_roll overlapped: six slots, one launch per step holding its chains and the "
                              "previous step's products" if overlap else "dpkem_roll: two launches per step"),
                 "saber_poly_ipw": ipw_best, "overlap": overlap,
                 "tuning_ms": {f"{'overlap' if o else 'plain'}-ipw{c}": round(v, 4) for (o, c), v in tried.items()}}
    else:
        sched = {"pipeline": "dpkem_roll: six launches per step (decaps' mu', every chain in one launch, three "
                             "fused products, decaps' last ss block)", "frodo_threads": best or None,
                 "tuning_ms": {f"threads{c or 'default'}": round(v, 3) for c, v in tried.items()}}
    return dict(wall_s=wall, per_step_ms=per_step, batch_ms=batch, batch_steps=span, roundtrip_ok=ok,
                ops_per_s=3 * n * steps / wall, schedule=sched)


def profile_rolling(torch, eng, scheme, n, steps, seed, shape=None):
    """Per-kernel HIP-event times of the rolling steps' kernels, launched
    without a graph.  shape: a roll_shapes candidate (default the first);
    for the Saber family an int is the three-slot form's product instances
    per workgroup.  The overlapped shape's step is one launch
    (sab_roll_fused); the three-slot form's launches each run alone."""
    import dpcrypto_amd as D
    saber = scheme in (D.SABER, D.LIGHTSABER, D.FIRESABER)
    if shape is None:
        shape = roll_shapes(scheme)[0]
    if saber and not isinstance(shape, tuple):
        shape = (False, shape)
    overlap = saber and bool(shape[0])
    slots = rolling_slots(torch, scheme, n, seed, 6 if overlap else 3)
    eng.set_stream(torch.cuda.current_stream(torch.device("cuda:0")).cuda_stream)
    r = build_roll_shape(eng, scheme, slots, shape, graph=False)
    for _ in range(6 if overlap else 3):
        r.step()
    torch.cuda.synchronize()
    eng.kernel_stats(reset=True)
    eng.set_profiling(True)
    for _ in range(steps):
        r.step()
    torch.cuda.synchronize()
    eng.set_profiling(False)
    stats = eng.kernel_stats(reset=True)
    for _ in range(6 if overlap else 2):  # every batch in flight finished
        r.step()
    torch.cuda.synchronize()
    ok = all(bool(torch.equal(sl["ss"], sl["ss2"])) for sl in slots)
    r.close()
    return dict(kernels=stats, roundtrip_ok=ok)


# Op-level ceilings (ops/s) from the algorithmic work of one op (Keccak
# permutations at 4,176 VALU lane-ops each plus MACs at 2 per lane-op) at the
# VALU peak; keygen+enc+dec at 1:1:1 is their harmonic combination.
# The serial-chain Keccak work sets them: permutations per op at 4,176 VALU
# lane-ops each over the VALU peak.  The polynomial and matrix products run
# on int8 MFMA beside the VALU (for Saber 9 / 12 / 12 x 65,536 13-bit MACs
# per op, two int8 limb MACs each: under 1% of a keygen's Keccak time at
# the 5 POP/s int8 peak), so they add no term.  Saber keygen: seedA 1,
# GenMatrix 23, GenSecret 5, H(pk) 8; encaps: GenMatrix 23, m 1, H(pk) 8,
# G 1, GenSecret 5, H(ct) 9, K 1; decaps: GenMatrix 23, H(ct) 9, both K 2,
# G 1, GenSecret 5.  FrodoKEM-976: A's 976 rows x 12 SHAKE128 blocks plus
# the SHAKE256 chains (keygen 1 + 230 + 115, encaps 115 + 1 + 231 + 116,
# decaps 1 + 231 + 116).
OP_PERMS = {
    ("saber", "keygen"): 37, ("saber", "enc"): 48, ("saber", "dec"): 40,
    ("frodo976", "keygen"): 976 * 12 + 346, ("frodo976", "enc"): 976 * 12 + 463, ("frodo976", "dec"): 976 * 12 + 348,
}
OP_CEILING = {k: VALU_PEAK / (v * KECCAK_LANE_OPS) for k, v in OP_PERMS.items()}
# The same ceilings with the Saber products' MACs also counted on the VALU
# (2 MACs per lane-op), as first estimated before the products moved to MFMA.
OP_CEILING_PRODUCTS_ON_VALU = {
    ("saber", "keygen"): 87.5e6, ("saber", "enc"): 66.2e6, ("saber", "dec"): 60.1e6,
}


def chain_ceiling(name, table=None):
    table = OP_CEILING if table is None else table
    return 3.0 / sum(1.0 / table[(name, op)] for op in ("keygen", "enc", "dec"))


def run_per_op(torch, eng, plain_eng, scheme, name, n, steps, seed, tune_budget_s=10.0):
    """keygen, encaps and decaps timed separately, each as its own
    graph-captured single-op plan over n device-resident instances (schedule
    autotuned per op), beside the plain one-workgroup-per-KEM form launched
    op by op on the same inputs.  Outputs of every form must agree."""
    import statistics
    import dpcrypto_amd as D
    from dpcrypto_amd import autotune
    a = chain_arrays(torch, scheme, n, seed)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    eng.set_stream(stream.cuda_stream)
    s = D.SIZES[scheme]
    # keys and ciphertexts the enc / dec plans read (untimed, eager)
    eng.keypair_into(scheme, n, a["pk"], a["sk"], a["kc"])
    eng.enc_into(scheme, n, a["ct"], a["ss"], a["pk"], a["ec"])
    eng.dec_into(scheme, n, a["ss2"], a["ct"], a["sk"])
    out = {op: [torch.empty((n, w), dtype=torch.uint8, device=dev) for w in ws]
           for op, ws in (("keygen", (s["pk"], s["sk"])), ("enc", (s["ct"], s["ss"])), ("dec", (s["ss"],)))}
    saber_family = scheme in (D.SABER, D.LIGHTSABER, D.FIRESABER)

    def build(op, ns, ch, knob=None):
        eng.set_tuning(D.TUNE_SABER_POLY_IPW, knob if saber_family and knob else 1)
        eng.set_tuning(D.TUNE_FRODO_THREADS, knob if not saber_family and knob else 0)
        p = eng.plan(nstreams=ns, chunks=ch)
        o = out[op]
        if op == "keygen":
            p.keypair(scheme, o[0], o[1], a["kc"])
        elif op == "enc":
            p.enc(scheme, o[0], o[1], a["pk"], a["ec"])
        else:
            p.dec(scheme, o[0], a["ct"], a["sk"])
        return p.build()

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            x, y = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            x.record(stream)
            fn()
            y.record(stream)
            y.synchronize()
            ts.append(x.elapsed_time(y))
        return statistics.median(ts)

    cands = ([(1, 1, 512), (2, 2, 512), (4, 4, 512), (8, 4, 512), (1, 1, 1024), (4, 4, 1024)] if not saber_family
             else [(1, 1, 1), (1, 1, 4), (2, 2, 1)])
    res = {}
    for op in ("keygen", "enc", "dec"):
        best, tried = autotune.tune(torch, eng, f"scheme{scheme}|{op}|n={n}", lambda *c: build(op, *c),
                                    candidates=cands, use_cache=False, budget_s=tune_budget_s)
        plan = build(op, *best)
        ms = timed(plan.launch, steps)
        plan.close()
        ceil = OP_CEILING[(name, op)]
        res[op] = {"ops_per_s": round(n / (ms / 1e3), 1), "ms_per_batch": round(ms, 4),
                   "ceiling_ops_per_s": ceil, "frac_of_ceiling": round(n / (ms / 1e3) / ceil, 4),
                   "schedule": "x".join(map(str, best))}
    ok = (torch.equal(out["keygen"][0], a["pk"]) and torch.equal(out["keygen"][1], a["sk"])
          and torch.equal(out["enc"][0], a["ct"]) and torch.equal(out["enc"][1], a["ss"])
          and torch.equal(out["dec"][0], a["ss"]) and torch.equal(a["ss2"], a["ss"]))
    if plain_eng is not None:
        plain_eng.set_stream(stream.cuda_stream)
        pout = {op: [torch.empty_like(t) for t in ts] for op, ts in out.items()}
        reps = max(1, min(steps, 3))
        pms = {
            "keygen": timed(lambda: plain_eng.keypair_into(scheme, n, pout["keygen"][0], pout["keygen"][1], a["kc"]), reps),
            "enc": timed(lambda: plain_eng.enc_into(scheme, n, pout["enc"][0], pout["enc"][1], a["pk"], a["ec"]), reps),
            "dec": timed(lambda: plain_eng.dec_into(scheme, n, pout["dec"][0], a["ct"], a["sk"]), reps)}
        for op in res:
            res[op]["plain_ops_per_s"] = round(n / (pms[op] / 1e3), 1)
            res[op]["batched_over_plain"] = round(pms[op] / res[op]["ms_per_batch"], 2)
        encdec = res["enc"]["ms_per_batch"] + res["dec"]["ms_per_batch"]
        res["enc_plus_dec"] = {"ops_per_s": round(2 * n / (encdec / 1e3), 1),
                               "plain_ops_per_s": round(2 * n / ((pms["enc"] + pms["dec"]) / 1e3), 1),
                               "batched_over_plain": round((pms["enc"] + pms["dec"]) / encdec, 2)}
        ok = ok and all(torch.equal(x, y) for op in out for x, y in zip(out[op], pout[op]))
    res["outputs_agree"] = bool(ok)
    res["workload"] = f"{name} batch {n}: each op as its own graph-captured plan over device-resident inputs"
    return res


def run_shared_key(torch, eng, scheme, n, steps, seed):
    """encaps + decaps of n instances against ONE key per step (serving
    pattern: many requests for a server's long-term key)."""
    import numpy as np
    import dpcrypto_amd as D
    s = D.SIZES[scheme]
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    eng.set_stream(stream.cuda_stream)
    kc = torch.from_numpy(rng.integers(0, 256, (1, s["kcoins"]), dtype=np.uint8)).to(dev)
    pk, sk = eng.keypair(scheme, kc)
    ec = torch.from_numpy(rng.integers(0, 256, (n, s["ecoins"]), dtype=np.uint8)).to(dev)
    ct, ss = eng.enc_shared_key(scheme, pk, ec)  # warm-up
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(steps):
        ct, ss = eng.enc_shared_key(scheme, pk, ec)
        ss2 = eng.dec_shared_key(scheme, ct, sk)
    b.record(stream)
    b.synchronize()
    ms = a.elapsed_time(b) / steps
    return {"workload": f"{'frodokem-976-shake' if scheme == 0 else 'saber-l3'}: {n} encaps + {n} decaps "
                        f"against one key per step", "ops_per_s": round(2 * n / (ms / 1e3), 1),
            "ms_per_step": round(ms, 3), "roundtrip_ok": bool(torch.equal(ss, ss2))}


def profile_plan(torch, eng, scheme, n, steps, seed):
    """Per-kernel HIP-event times of the plan the headline replays (one
    stream, one chunk, pipelined keygen -> encaps -> decaps chain), issued
    without a graph (DPKEM_PLAN_NO_GRAPH) so that the engine's profiler
    brackets every kernel with events on the stream it runs on."""
    a = chain_arrays(torch, scheme, n, seed)
    eng.set_stream(torch.cuda.current_stream(torch.device("cuda:0")).cuda_stream)
    p = build_chain_plan(eng, scheme, a, 1, 1, graph=False)
    p.launch()
    torch.cuda.synchronize()
    eng.kernel_stats(reset=True)
    eng.set_profiling(True)
    for _ in range(steps):
        p.launch()
    torch.cuda.synchronize()
    eng.set_profiling(False)
    stats = eng.kernel_stats(reset=True)
    ok = bool(torch.equal(a["ss"], a["ss2"]))
    p.close()
    return dict(kernels=stats, roundtrip_ok=ok)


# Serial lane-pair permutations on the critical path of the pipelined chain
# launches (the longest chain inside each launch), and the issue floor of
# one such permutation on a lone wave: 24 rounds x 108 VALU ops, one VALU op
# per 4 clocks per wave at 2.4 GHz.  A latency-bound launch cannot beat
# CRITICAL_PERMS x this floor, whatever its VALU throughput.
CRITICAL_PERMS = {"sab_pipe_x0": 24, "sab_pipe_x1": 14, "sab_pipe_x2": 10,
                  "sab_pipe_x0,l2": 11, "sab_pipe_x1,l2": 10, "sab_pipe_x2,l2": 7,
                  "sab_pipe_x0,l4": 41, "sab_pipe_x1,l4": 16, "sab_pipe_x2,l4": 12,
                  "sab_mix_x0": 24, "sab_mix_x1": 10, "sab_mix_x0,l2": 11, "sab_mix_x1,l2": 7,
                  "sab_mix_x0,l4": 41, "sab_mix_x1,l4": 12,
                  "sab_roll_x0": 24, "sab_roll_x0,l2": 11, "sab_roll_x0,l4": 41,
                  "sab_roll_fused": 24, "sab_roll_fused,l2": 11, "sab_roll_fused,l4": 41,
                  "frodo_roll_x0": 347, "frodo_roll_x0,640": 182, "frodo_roll_x0,1344": 478}
LANE_PAIR_PERM_FLOOR_S = 24 * 108 * 4 / 2.4e9


def measure_perm_latency(eng, short=64, long=576):
    """Measured latency of one lane-pair permutation on a lone wave (s): two
    launches of the engine's probe kernel (one wave, back-to-back
    permutations in the chains' unrolled form), their difference over the
    extra permutations, so the launch overhead cancels.  Median of 3."""
    import statistics
    eng.probe_perm_latency(short)  # warm-up
    vals = []
    for _ in range(3):
        t_short = eng.probe_perm_latency(short)
        t_long = eng.probe_perm_latency(long)
        vals.append((t_long - t_short) / 1e3 / (long - short))
    return statistics.median(vals)


def step_stats(run, prof):
    """Kernel stats for the roofline: in the overlapped form a step is one
    launch (sab_roll_fused), so the timed steps' own HIP events (recorded on
    the engine's stream right around each step's launch) time it without the
    profiler's per-launch event overhead; otherwise the profiled kernels."""
    names = list(prof["kernels"])
    if run["schedule"].get("overlap") and len(names) == 1:
        return {names[0]: (len(run["per_step_ms"]), float(sum(run["per_step_ms"])))}
    return prof["kernels"]


def roofline_from(stats, n, perm_latency_s=None):
    """Dominant kernel by total time -> achieved VALU lane-ops/s vs peak (and,
    for a serial-chain launch, its time against the chain's issue floor:
    nominal, and from the measured lone-wave permutation latency)."""
    if not stats:
        return None
    name, (count, total_ms) = max(stats.items(), key=lambda kv: kv[1][1])
    avg_s = total_ms / count / 1e3
    work = lane_ops(name, n)
    achieved = work / avg_s
    hbm = measured_hbm(name, n)
    chain = CRITICAL_PERMS.get(name)
    floor = None if chain is None else {
        "critical_perms": chain,
        "floor_ms": round(chain * LANE_PAIR_PERM_FLOOR_S * 1e3, 4),
        "frac": round(chain * LANE_PAIR_PERM_FLOOR_S / avg_s, 4),
        "what": "serial lane-pair permutations x 24 rounds x 108 VALU ops x 4 clocks at 2.4 GHz"}
    if floor is not None and perm_latency_s:
        floor.update({
            "measured_perm_us": round(perm_latency_s * 1e6, 3),
            "measured_floor_ms": round(chain * perm_latency_s * 1e3, 4),
            "measured_frac": round(chain * perm_latency_s / avg_s, 4),
            "measured_what": "critical permutations x the lone-wave permutation latency measured in this run "
                             "(dpkem_probe_perm_latency)"})
    return {
        "kernel": name,
        "bound": "valu",
        "achieved": round(achieved / 1e12, 3),
        "peak": round(VALU_PEAK / 1e12, 3),
        "unit": "Tlane-op/s",
        "frac": round(achieved / VALU_PEAK, 4),
        "traffic": None if hbm is None else hbm[0] + hbm[1],
        "traffic_read_write": None if hbm is None else list(hbm),
        "traffic_GBps": None if hbm is None else round((hbm[0] + hbm[1]) / avg_s / 1e9, 1),
        "traffic_source": ("rocprofv3 FETCH_SIZE / WRITE_SIZE of these kernel sources (sha "
                           f"{source_sha(kernel_family(name))}), scaled by a copy kernel"
                           if hbm is not None else "unmeasured for the current kernel sources"),
        "avg_launch_ms": round(avg_s * 1e3, 4),
        "algorithmic_lane_ops_per_launch": work,
        "share_of_step": round(total_ms / sum(v[1] for v in stats.values()), 3),
        "chain_issue_floor": floor,
    }


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(budget_s: float):
    """The oracle (a scalar C port of the path) on this host's cores, on a
    bounded sample: Saber l=3 and FrodoKEM-976 keygen+encaps+decaps, one
    instance per thread at a time, on all usable cores and on one core.  The
    oracle is compiled here with -O3 -march=native for this host first (the
    checked-in build targets a portable x86-64 level)."""
    import tempfile
    import numpy as np
    from oracle import pyoracle as O
    threads = max(1, min(len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count() or 1,
                         16))
    flags = "-O3 -march=native"
    tmp = tempfile.mkdtemp(prefix="dpkem_orc_")
    try:
        O.use_library(O.build_native(os.path.join(tmp, "liboracle_native.so")))
    except Exception as exc:  # no compiler on this host: time the portable build
        O.use_library(None)
        flags = f"-O3 -march=x86-64-v2 (native build failed: {exc.__class__.__name__})"

    def rate(scheme, th, per_thread, seconds):
        s = O.SIZES[scheme]
        n = th * per_thread
        rng = np.random.default_rng(1)
        kc = rng.integers(0, 256, (n, s["kcoins"]), dtype=np.uint8)
        ec = rng.integers(0, 256, (n, s["ecoins"]), dtype=np.uint8)
        ops = 0
        t_op = [0.0, 0.0, 0.0]  # keygen, encaps, decaps time within the same rounds
        t0 = time.perf_counter()
        while True:
            ta = time.perf_counter()
            pk, sk = O.keypair(scheme, kc, threads=th)
            tb = time.perf_counter()
            ct, ss = O.enc(scheme, pk, ec, threads=th)
            tc = time.perf_counter()
            O.dec(scheme, ct, sk, threads=th)
            td = time.perf_counter()
            t_op[0] += tb - ta
            t_op[1] += tc - tb
            t_op[2] += td - tc
            ops += 3 * n
            if td - t0 >= seconds:
                break
        dt = time.perf_counter() - t0
        rounds = ops // (3 * n)
        per_op = {name: round(n * rounds / t, 1) for name, t in zip(("keygen", "enc", "dec"), t_op)}
        return ops / dt, n, rounds, dt, per_op

    try:
        v, n, rounds, dt, per_op = rate(O.SABER, threads, 32, budget_s * 0.45)
        v1, _, _, dt1, per_op1 = rate(O.SABER, 1, 32, budget_s * 0.1)
        fv, fn, frounds, fdt, fper_op = rate(O.FRODO, threads, 2, budget_s * 0.35)
        fv1, _, _, fdt1, fper_op1 = rate(O.FRODO, 1, 2, budget_s * 0.1)
    finally:
        O.use_library(None)
    return {"value": round(v, 1), "unit": "KEM ops/s", "cores": threads, "kind": "port",
            "sample": f"Saber l=3 keygen+encaps+decaps, {n} instances per round, {rounds} rounds "
                      f"({dt:.1f} s) on {threads} threads of the CPU oracle (oracle/, {flags})",
            "single_core_value": round(v1, 1), "single_core_seconds": round(dt1, 1),
            "per_op": per_op, "single_core_per_op": per_op1,
            "frodo976": {"value": round(fv, 1), "unit": "KEM ops/s", "cores": threads,
                         "sample": f"FrodoKEM-976-SHAKE keygen+encaps+decaps, {fn} instances per round, {frounds} "
                                   f"rounds ({fdt:.1f} s) on {threads} threads",
                         "single_core_value": round(fv1, 1), "single_core_seconds": round(fdt1, 1),
                         "per_op": fper_op, "single_core_per_op": fper_op1},
            "compiler_flags": flags, "nproc": os.cpu_count(), "cpu_model": _cpu_model()}


# FrodoKEM rolling pipelines measured in a child process of their own (the
# parent reads one JSON line from it): keyed by the bench line's names.
FRODO_ROLLING = (("frodo976", 0, 0x44504B454D000002), ("frodo976_aes", 2, 0x44504B454D000007),
                 ("frodo640", 5, 0x44504B454D000010 + 5), ("frodo1344", 6, 0x44504B454D000010 + 6))


def frodo_rolling_child(args):
    """--frodo-rolling-child: every FrodoKEM set's rolling steps (graph
    replays, launch shape tuned first) and the same steps' kernels with HIP
    events; one JSON line on stdout."""
    import torch
    import dpcrypto_amd as D
    eng = D.Engine(0)
    out = {}
    for name, sc, seed in FRODO_ROLLING:
        print(f"[bench] {name} batch {args.frodo_batch}: rolling keygen -> encaps -> decaps steps, 3 warmup + "
              f"{args.frodo_steps} timed graph steps, then the steps' kernels with HIP events (child process)",
              file=sys.stderr, flush=True)
        r = run_rolling_workload(torch, eng, sc, args.frodo_batch, args.frodo_steps, 3, seed)
        prof = profile_rolling(torch, eng, sc, args.frodo_batch, 2 if name == "frodo976" else 1, seed,
                               r["schedule"]["frodo_threads"] or 0)
        out[name] = {"ops_per_s": r["ops_per_s"], "wall_s": r["wall_s"], "per_step_ms": r["per_step_ms"],
                     "batch_ms": r["batch_ms"], "roundtrip_ok": r["roundtrip_ok"] and prof["roundtrip_ok"],
                     "schedule": r["schedule"], "kernels": {k: list(v) for k, v in prof["kernels"].items()}}
        print(f"[bench] {name}: {r['ops_per_s']:.0f} ops/s rolling", file=sys.stderr, flush=True)
    eng.close()
    print(json.dumps(out), flush=True)


def run_frodo_rolling_child(args, progress):
    """Runs frodo_rolling_child in a fresh process; returns (results, None)
    or (None, what went wrong)."""
    import subprocess
    progress("frodo rolling pipelines (child process)")
    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--frodo-rolling-child", "--frodo-batch",
           str(args.frodo_batch), "--frodo-steps", str(args.frodo_steps)]
    try:
        r = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE, timeout=600)
    except subprocess.TimeoutExpired:
        return None, "child process timed out"
    if r.returncode != 0:
        return None, f"child process exited with status {r.returncode}"
    try:
        return json.loads(r.stdout.decode().strip().splitlines()[-1]), None
    except (ValueError, IndexError):
        return None, "child process printed no result"


def merge_rolling(entry, roll, steps, n, label):
    """A FrodoKEM parameter set's line: the rolling steps' rate as ops_per_s
    when the child measured them (the chained plan's under chained_ops_per_s),
    otherwise the chained plan's alone."""
    if not roll:
        entry["rolling_error"] = "not measured"
        return
    entry["chained_ops_per_s"] = entry["ops_per_s"]
    entry["chained_schedule"] = entry.get("schedule")
    entry["workload"] = f"{label} batch {n}: keygen+encaps+decaps per step (rolling, measured in a child process)"
    entry["ops_per_s"] = round(roll["ops_per_s"], 1)
    entry["ms_per_step"] = round(roll["wall_s"] * 1e3 / steps, 3)
    entry["roundtrip_ok"] = entry["roundtrip_ok"] and roll["roundtrip_ok"]
    entry["schedule"] = {k: v for k, v in roll["schedule"].items() if k != "tuning_ms"}
    entry["roofline"] = roofline_from(roll["kernels"], n)
    entry["kernels_ms"] = {k: round(v[1] / v[0], 4) for k, v in roll["kernels"].items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--saber-batch", type=int, default=4096)
    ap.add_argument("--frodo-batch", type=int, default=8192)
    ap.add_argument("--frodo-steps", type=int, default=6)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--skip-frodo", action="store_true")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-per-op", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--skip-mixed", action="store_true")
    ap.add_argument("--skip-serving", action="store_true")
    ap.add_argument("--serving-seconds", type=float, default=0.5, help="arrival window per rate")
    ap.add_argument("--serving-rates", default="10000,50000,100000,200000,500000,1000000,2000000")
    ap.add_argument("--schedule", default=None,
                    help="fixed plan schedule STREAMSxCHUNKS[xIPW] instead of autotuning (profiling runs; "
                         "IPW = Saber product instances per workgroup, default 1)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one elementwise kernel over 256 MiB in / 256 MiB out (HBM counter scale)")
    ap.add_argument("--frodo-rolling-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    import faulthandler
    faulthandler.enable()  # a fatal signal prints the Python stack of every thread
    if args.frodo_rolling_child:
        frodo_rolling_child(args)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: instances are independent, so N GPUs would run N disjoint "
              "shards of the batch with no communication; multi-GPU runs are not part of this build.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import dpcrypto_amd as D

    def progress(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    sched = tuple(int(x) for x in args.schedule.split("x")) if args.schedule else None
    if args.calibrate_copy:
        x = torch.ones(1 << 26, dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        torch.neg(x, out=y)
        torch.cuda.synchronize()
        del x, y
    eng = D.Engine(0)
    ipw_fixed = sched[2] if sched and len(sched) > 2 else None
    progress(f"saber batch {args.saber_batch}: rolling keygen -> encaps -> decaps steps, "
             f"{args.warmup} warmup + {args.steps} timed graph steps")
    sab = run_rolling_workload(torch, eng, D.SABER, args.saber_batch, args.steps, args.warmup, 0x44504B454D000001,
                               ipw=ipw_fixed)
    progress(f"saber: {sab['ops_per_s']:.0f} ops/s (product ipw {sab['schedule']['saber_poly_ipw']})")
    progress("saber: the rolling steps' kernels launched without a graph, with per-kernel HIP events")
    sab_shape = (sab["schedule"]["overlap"], sab["schedule"]["saber_poly_ipw"])
    sab_prof = profile_rolling(torch, eng, D.SABER, args.saber_batch, max(3, args.steps), 0x44504B454D000001,
                               sab_shape)
    # the same step's launches alone (three-slot form, chains then products)
    sab_alone = sab_prof if not sab_shape[0] else profile_rolling(
        torch, eng, D.SABER, args.saber_batch, max(3, args.steps), 0x44504B454D000001, (False, sab_shape[1]))
    progress(f"saber batch {args.saber_batch}: one batch's keygen -> encaps -> decaps chained inside each step")
    chained = run_plan_workload(torch, eng, D.SABER, args.saber_batch, args.steps, args.warmup, 0x44504B454D000001,
                                schedule=sched)
    chained_prof = profile_plan(torch, eng, D.SABER, args.saber_batch, args.steps, 0x44504B454D000001)
    progress("saber eager pass (op by op) with per-kernel HIP events")
    sab_eager = run_workload(torch, eng, D.SABER, args.saber_batch, args.steps, 1, 0x44504B454D000001, True)
    perm_s = measure_perm_latency(eng)
    progress(f"lone-wave lane-pair permutation latency: {perm_s * 1e6:.3f} us")
    result = {
        "metric": "KEM ops/s (keygen+enc+dec) Frodo-976 & Saber at 1 GPU; p99 latency",
        "value": round(sab["ops_per_s"], 1),
        "unit": "KEM ops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sab["wall_s"] * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16/u32 integer (mod 2^13 / 2^16)",
        "data": "synthetic seeded coins",
        "config": {"workload": f"saber-l3 batch {args.saber_batch}: {args.saber_batch} keygens + "
                               f"{args.saber_batch} encaps + {args.saber_batch} decaps per step; "
                               + ("rolling pipeline over six batch slots (encaps to the keys of two steps earlier, "
                                  "decaps of the ciphertexts of two steps earlier; each step is one launch holding "
                                  "its hash chains and the previous step's products; a batch is complete five "
                                  "steps after its keygen step)" if sab["schedule"].get("overlap") else
                                  "rolling pipeline over three batch slots (encaps to the keys of the previous "
                                  "step, decaps of the ciphertexts of the previous step; a batch is complete two "
                                  "steps after its keygen)"),
                   "global_batch": args.saber_batch, "parallelism": "instances (1 GPU)"},
        "latency_ms": {"p50": round(pct(sab["per_step_ms"], 0.5), 4), "p99": round(pct(sab["per_step_ms"], 0.99), 4),
                       "what": "per-step latency (one step = the three ops over 4096 instances each)",
                       "batch_p50": round(pct(sab["batch_ms"], 0.5), 4),
                       "batch_p99": round(pct(sab["batch_ms"], 0.99), 4),
                       "batch_what": f"one batch from keygen to decaps: {sab['batch_steps']} consecutive steps"},
        "roundtrip_ok": sab["roundtrip_ok"] and sab_prof["roundtrip_ok"] and sab_alone["roundtrip_ok"]
                        and sab_eager["roundtrip_ok"]
                        and chained["roundtrip_ok"] and chained_prof["roundtrip_ok"],
        "schedule": sab["schedule"],
        "eager_ops_per_s": round(sab_eager["ops_per_s"], 1),
        "roofline": roofline_from(step_stats(sab, sab_prof), args.saber_batch, perm_s),
        "kernels_ms": {k: round(v[1] / v[0], 4) for k, v in sab_prof["kernels"].items()},
        "alone_kernels_ms": {k: round(v[1] / v[0], 4) for k, v in sab_alone["kernels"].items()},
        "eager_kernels_ms": {k: round(v[1] / v[0], 4) for k, v in sab_eager["kernels"].items()},
        "saber_chained": {
            "workload": f"saber-l3 batch {args.saber_batch}: keygen -> encaps -> decaps of one batch chained "
                        "inside each step (pipelined launch sequence, six launches)",
            "ops_per_s": round(chained["ops_per_s"], 1),
            "ms_per_step": round(chained["wall_s"] * 1e3 / args.steps, 4),
            "latency_ms": {"p50": round(pct(chained["per_step_ms"], 0.5), 4),
                           "p99": round(pct(chained["per_step_ms"], 0.99), 4)},
            "schedule": chained["schedule"],
            "roofline": roofline_from(chained_prof["kernels"], args.saber_batch, perm_s),
            "kernels_ms": {k: round(v[1] / v[0], 4) for k, v in chained_prof["kernels"].items()}},
    }
    pe = None if args.skip_plain else D.Engine(0, form=D.FORM_PLAIN)
    if not args.skip_per_op:
        progress(f"saber batch {args.saber_batch}: keygen, encaps, decaps timed as separate plans, and in the plain form")
        result["per_op"] = {"saber": run_per_op(torch, eng, pe, D.SABER, "saber", args.saber_batch, args.steps,
                                                0x44504B454D000001)}
    result["ceiling_ops_per_s"] = {"saber_keygen_enc_dec": round(chain_ceiling("saber"), 1),
                                   "frodo976_keygen_enc_dec": round(chain_ceiling("frodo976"), 1),
                                   "basis": "serial-chain Keccak permutations per op at 4,176 VALU lane-ops each "
                                            "over the VALU peak; products on int8 MFMA alongside",
                                   "saber_with_products_on_valu": round(
                                       chain_ceiling("saber", OP_CEILING_PRODUCTS_ON_VALU), 1)}
    result["frac_of_ceiling"] = round(sab["ops_per_s"] / chain_ceiling("saber"), 4)
    if not args.skip_frodo:
        rolling, rolling_err = run_frodo_rolling_child(args, progress)
        progress(f"frodo976 batch {args.frodo_batch}: one batch's keygen -> encaps -> decaps chained inside each "
                 f"step, autotune, 2 warmup + {args.frodo_steps} timed graph steps")
        fr_ch = run_plan_workload(torch, eng, D.FRODO976, args.frodo_batch, args.frodo_steps, 2, 0x44504B454D000002,
                                  schedule=sched)
        progress(f"frodo976: {fr_ch['ops_per_s']:.0f} ops/s chained; plan kernels and an eager pass with per-kernel "
                 "HIP events")
        fr_ch_prof = profile_plan(torch, eng, D.FRODO976, args.frodo_batch, 1, 0x44504B454D000002)
        fr_eager = run_workload(torch, eng, D.FRODO976, args.frodo_batch, args.frodo_steps, 1, 0x44504B454D000002, True)
        chained_fr = {
            "workload": f"frodokem-976-shake batch {args.frodo_batch}: keygen -> encaps -> decaps of one batch "
                        "chained inside each step (pipelined plan, eight launches per chunk)",
            "ops_per_s": round(fr_ch["ops_per_s"], 1),
            "frac_of_ceiling": round(fr_ch["ops_per_s"] / chain_ceiling("frodo976"), 4),
            "ms_per_step": round(fr_ch["wall_s"] * 1e3 / args.frodo_steps, 3),
            "latency_ms": {"p50": round(pct(fr_ch["per_step_ms"], 0.5), 3),
                           "p99": round(pct(fr_ch["per_step_ms"], 0.99), 3)},
            "schedule": fr_ch["schedule"],
            "roofline": roofline_from(fr_ch_prof["kernels"], args.frodo_batch),
            "kernels_ms": {k: round(v[1] / v[0], 4) for k, v in fr_ch_prof["kernels"].items()}}
        fr = (rolling or {}).get("frodo976")
        if fr:
            result["frodo976"] = {
                "workload": f"frodokem-976-shake batch {args.frodo_batch}: {args.frodo_batch} keygens + "
                            f"{args.frodo_batch} encaps + {args.frodo_batch} decaps per step; rolling pipeline over "
                            "three batch slots (encaps to the keys of the previous step, decaps of the ciphertexts "
                            "of the previous step; a batch is complete two steps after its keygen); measured in a "
                            "child process",
                "ops_per_s": round(fr["ops_per_s"], 1),
                "frac_of_ceiling": round(fr["ops_per_s"] / chain_ceiling("frodo976"), 4),
                "ms_per_step": round(fr["wall_s"] * 1e3 / args.frodo_steps, 3),
                "latency_ms": {"p50": round(pct(fr["per_step_ms"], 0.5), 3),
                               "p99": round(pct(fr["per_step_ms"], 0.99), 3),
                               "batch_p50": round(pct(fr["batch_ms"], 0.5), 3),
                               "batch_p99": round(pct(fr["batch_ms"], 0.99), 3),
                               "batch_what": "one batch from keygen to decaps: three consecutive steps"},
                "steps": args.frodo_steps,
                "roundtrip_ok": (fr["roundtrip_ok"] and fr_ch["roundtrip_ok"] and fr_ch_prof["roundtrip_ok"]
                                 and fr_eager["roundtrip_ok"]),
                "schedule": fr["schedule"],
                "eager_ops_per_s": round(fr_eager["ops_per_s"], 1),
                "roofline": roofline_from(fr["kernels"], args.frodo_batch),
                "kernels_ms": {k: round(v[1] / v[0], 4) for k, v in fr["kernels"].items()},
                "eager_kernels_ms": {k: round(v[1] / v[0], 4) for k, v in fr_eager["kernels"].items()},
                "chained": chained_fr,
            }
        else:
            result["frodo976"] = {
                **chained_fr, "steps": args.frodo_steps,
                "roundtrip_ok": fr_ch["roundtrip_ok"] and fr_ch_prof["roundtrip_ok"] and fr_eager["roundtrip_ok"],
                "eager_ops_per_s": round(fr_eager["ops_per_s"], 1),
                "eager_kernels_ms": {k: round(v[1] / v[0], 4) for k, v in fr_eager["kernels"].items()},
                "rolling_error": rolling_err}
        if not args.skip_per_op:
            progress(f"frodo976 batch {args.frodo_batch}: keygen, encaps, decaps timed as separate plans, "
                     "and in the plain form")
            result["per_op"]["frodo976"] = run_per_op(torch, eng, pe, D.FRODO976, "frodo976", args.frodo_batch,
                                                      max(2, args.frodo_steps // 2), 0x44504B454D000002)
    if not args.skip_plain:
        progress("plain comparator forms")
        sp = run_workload(torch, pe, D.SABER, args.saber_batch, args.plain_steps, 1, 0x44504B454D000001)
        plain = {"saber_ops_per_s": round(sp["ops_per_s"], 1),
                 "saber_speedup_batched_over_plain": round(sab["ops_per_s"] / sp["ops_per_s"], 2)}
        if not args.skip_frodo:
            fp = run_workload(torch, pe, D.FRODO976, args.frodo_batch, 1, 0, 0x44504B454D000002)
            plain["frodo976_ops_per_s"] = round(fp["ops_per_s"], 1)
            plain["frodo976_speedup_batched_over_plain"] = round(result["frodo976"]["ops_per_s"] / fp["ops_per_s"], 2)
        result["plain"] = plain
    if not args.skip_frodo:
        progress(f"frodo976-aes batch {args.frodo_batch}: autotuned graph steps, then an eager pass with HIP events")
        fa = run_plan_workload(torch, eng, D.FRODO976_AES, args.frodo_batch, args.frodo_steps, 1, 0x44504B454D000007,
                               tune_budget_s=15.0, schedule=sched)
        fa_eager = run_workload(torch, eng, D.FRODO976_AES, args.frodo_batch, 1, 0, 0x44504B454D000007, True)
        result["frodo976_aes"] = {
            "workload": f"frodokem-976-aes batch {args.frodo_batch}: keygen+encaps+decaps per step",
            "ops_per_s": round(fa["ops_per_s"], 1), "ms_per_step": round(fa["wall_s"] * 1e3 / args.frodo_steps, 3),
            "roundtrip_ok": fa["roundtrip_ok"] and fa_eager["roundtrip_ok"],
            "schedule": {k: v for k, v in fa["schedule"].items() if k != "tuning_ms"},
            "kernels_ms": {k: round(v[1] / v[0], 4) for k, v in fa_eager["kernels"].items()}}
        merge_rolling(result["frodo976_aes"], (rolling or {}).get("frodo976_aes"), args.frodo_steps,
                      args.frodo_batch, "frodokem-976-aes")
        progress("frodo976 encaps + decaps against one key (A expanded once per 4 instances)")
        result["frodo976_shared_key"] = run_shared_key(torch, eng, D.FRODO976, args.frodo_batch, args.frodo_steps,
                                                       0x44504B454D000006)
        for name, sc in (("frodo640", D.FRODO640), ("frodo1344", D.FRODO1344)):
            progress(f"{name} batch {args.frodo_batch}: autotuned graph steps, then an eager pass with HIP events")
            fs = run_plan_workload(torch, eng, sc, args.frodo_batch, args.frodo_steps, 1, 0x44504B454D000010 + sc,
                                   tune_budget_s=15.0, schedule=sched)
            fs_prof = profile_plan(torch, eng, sc, args.frodo_batch, 1, 0x44504B454D000010 + sc)
            result[name] = {
                "workload": f"{name}-shake batch {args.frodo_batch}: keygen+encaps+decaps per step",
                "ops_per_s": round(fs["ops_per_s"], 1), "ms_per_step": round(fs["wall_s"] * 1e3 / args.frodo_steps, 3),
                "roundtrip_ok": fs["roundtrip_ok"] and fs_prof["roundtrip_ok"],
                "schedule": {k: v for k, v in fs["schedule"].items() if k != "tuning_ms"},
                "roofline": roofline_from(fs_prof["kernels"], args.frodo_batch),
                "kernels_ms": {k: round(v[1] / v[0], 4) for k, v in fs_prof["kernels"].items()}}
            merge_rolling(result[name], (rolling or {}).get(name), args.frodo_steps, args.frodo_batch, f"{name}-shake")
    for name, sc in (("lightsaber", D.LIGHTSABER), ("firesaber", D.FIRESABER)):
        progress(f"{name} batch {args.saber_batch}: rolling steps, then the chained form")
        fam = run_rolling_workload(torch, eng, sc, args.saber_batch, args.steps, args.warmup,
                                   0x44504B454D000008 + sc, ipw=ipw_fixed)
        fam_prof = profile_rolling(torch, eng, sc, args.saber_batch, 3, 0x44504B454D000008 + sc,
                                   (fam["schedule"]["overlap"], fam["schedule"]["saber_poly_ipw"]))
        fam_ch = run_plan_workload(torch, eng, sc, args.saber_batch, args.steps, args.warmup,
                                   0x44504B454D000008 + sc, tune_budget_s=20.0, schedule=sched)
        result[name] = {"workload": f"{name} batch {args.saber_batch}: keygen+encaps+decaps per step (rolling)",
                        "ops_per_s": round(fam["ops_per_s"], 1),
                        "ms_per_step": round(fam["wall_s"] * 1e3 / args.steps, 4),
                        "roundtrip_ok": fam["roundtrip_ok"] and fam_prof["roundtrip_ok"] and fam_ch["roundtrip_ok"],
                        "schedule": {k: v for k, v in fam["schedule"].items() if k != "tuning_ms"},
                        "roofline": roofline_from(step_stats(fam, fam_prof), args.saber_batch, perm_s),
                        "kernels_ms": {k: round(v[1] / v[0], 4) for k, v in fam_prof["kernels"].items()},
                        "chained_ops_per_s": round(fam_ch["ops_per_s"], 1),
                        "chained_schedule": {k: v for k, v in fam_ch["schedule"].items() if k != "tuning_ms"}}
    progress("saber with the chain run as three dependent tasks (no stage pipelining)")
    unpiped = run_plan_workload(torch, eng, D.SABER, args.saber_batch, args.steps, args.warmup, 0x44504B454D000001,
                                tune_budget_s=20.0, schedule=sched, pipeline=False)
    result["saber_unpipelined"] = {
        "ops_per_s": round(unpiped["ops_per_s"], 1), "roundtrip_ok": unpiped["roundtrip_ok"],
        "ms_per_step": round(unpiped["wall_s"] * 1e3 / args.steps, 4),
        "schedule": {k: v for k, v in unpiped["schedule"].items() if k != "tuning_ms"},
        "note": "keygen, encaps and decaps as three dependent plan tasks (ten launches per chunk); the headline "
                "runs the same work with stages of different ops sharing launches (six per chunk)"}
    progress("saber with the plan's opt-in matrix sharing across keygen -> enc -> dec")
    shared = run_plan_workload(torch, eng, D.SABER, args.saber_batch, args.steps, args.warmup, 0x44504B454D000001,
                               tune_budget_s=20.0, share_matrix=True, schedule=sched)
    result["saber_shared_matrix_chain"] = {
        "ops_per_s": round(shared["ops_per_s"], 1), "roundtrip_ok": shared["roundtrip_ok"],
        "schedule": {k: v for k, v in shared["schedule"].items() if k != "tuning_ms"},
        "note": "encaps/decaps reuse keygen's expanded matrix; valid only when one caller runs all three ops"}
    if not args.skip_mixed:
        from dpcrypto_amd import workloads as W
        mixed_plain = pe if pe is not None else D.Engine(0, form=D.FORM_PLAIN)
        result["mixed"] = W.run_mixed(torch, eng, mixed_plain, progress=progress)
        if pe is None:
            mixed_plain.close()
        progress(f"mixed: {result['mixed']['batched_ops_per_s']:.0f} ops/s batched, "
                 f"{result['mixed']['plain_ops_per_s']:.0f} plain")
    if not args.skip_serving:
        # each server runs in a fresh child process, as a serving process
        # would: its lanes' streams then get hardware queues of their own
        # instead of queues shared with the streams this process created
        import subprocess

        def serve(extra):
            cmd = [sys.executable, "-u", "-m", "dpcrypto_amd.workloads", "serving", "--rates", args.serving_rates,
                   "--seconds", str(args.serving_seconds)] + extra
            out = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE, timeout=300, check=True).stdout
            return json.loads(out.decode().strip().splitlines()[-1])

        progress("serving: continuous batching (child process)")
        bat = serve([])
        progress("serving: continuous batching without key grouping (child process)")
        nogroup = serve(["--no-key-grouping"])
        progress("serving: one request per launch (child process)")
        one = serve(["--one-per-launch"])
        result["serving"] = {
            "workload": f"Poisson arrivals, {args.serving_seconds} s per rate, 70% Saber / 30% Frodo-976, "
                        f"keygen:enc:dec = 1:4:4, requests drawn over a store of 64 Frodo / 256 Saber keys "
                        f"(Frodo batches grouped by key: A expanded once per 4 same-key members), keys and "
                        f"ciphertexts resident in HBM, budget 32 GiB; "
                        f"latency = completion - scheduled arrival; requests still queued 2 s after the last "
                        f"arrival are cancelled",
            "batched": bat["rows"], "batched_server": bat["server"],
            "batched_without_key_grouping": nogroup["rows"],
            "one_request_per_launch": one["rows"]}
    if args.cpu_seconds > 0:
        progress(f"cpu baseline ({args.cpu_seconds:.0f} s)")
        result["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    if pe is not None:
        pe.close()
    eng.close()
    # leave no process behind: children of this run (serving processes,
    # compiler drivers) must be gone before the line is printed
    try:
        import psutil
        left = psutil.Process().children(recursive=True)
        for c in left:
            c.terminate()
        psutil.wait_procs(left, timeout=5)
        result["child_processes_left"] = len(left)
    except ImportError:
        pass
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
