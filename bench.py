#!/usr/bin/env python3
"""Benchmark of the RNS-CKKS hot path on one MI355X.

Workload (the production single-op configuration): ring dimension N = 2^16, 25 Q limbs
(one 60-bit + 24 x 50-bit primes), dnum = 3 (9 special primes), a batch of 64 ciphertexts at the
top level, 16 keyed rotation steps. One step rotates every ciphertext of the batch by a step drawn
(seeded) from the 16 keyed steps through the batched hybrid key switch (fh_rotate_batch). Inputs
and keys are synthetic uniform residues generated on the device; the same batch is also run one
rotation at a time (fh_eval_rotate) and the two results are compared bit for bit.

Also reported: NTT GB/s (forward + inverse over all 64 x 2 x 25 limbs), EvalMult+Rescale ops/s,
a mixed-level op stream (256 ciphertexts at levels 1..24; rotations, multiplications with
relinearization and Rescale, additions) run one operation per call and as one fused batch (eager
and HIP-graph replayed), the roofline of the dominant kernel (HIP-event timing of its launches
inside the timed region) and the CPU oracle timed on a bounded sample of the same rotation
workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fideslib-fork_amd"))

METRIC = "HKS rotations/s N=2^16 L=24, batched vs per-op, bit-exact; NTT GB/s; p99 ms"
HBM_PEAK_GBS = 8000.0
LOGN, L1, K, DNUM = 16, 25, 9, 3
STEPS16 = [1 << i for i in range(15)] + [3]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU oracle sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stream-ops", type=int, default=256, help="ciphertexts of the mixed-level op stream (0 = skip)")
    ap.add_argument("--lt-count", type=int, default=32, help="linear transforms of the BSGS workload (0 = skip)")
    ap.add_argument("--serve-requests", type=int, default=200, help="requests per serving session (0 = skip)")
    ap.add_argument("--serve-rates", type=str, default="50,150,400", help="arrival rates (requests/s) of the serving sweep")
    ap.add_argument("--copy-mb", type=int, default=1024, help="device copy of this size first (counter calibration; 0 = skip)")
    ap.add_argument("--no-bootstrap", action="store_true", help="skip the full-slot bootstrap at N=2^16")
    ap.add_argument("--tune", action="append", default=[], metavar="NAME=VALUE", help="set a context tunable (repeatable)")
    ap.add_argument("--rotation-only", action="store_true",
                    help="time the batched rotation step only (for profiler runs: every dispatch after set-up belongs to it)")
    ap.add_argument("--stream-only", action="store_true",
                    help="run the mixed-level op stream only and print its result (for profiler runs)")
    ap.add_argument("--ntt-only", action="store_true",
                    help="time forward + inverse NTTs over the batch's 64 x 2 x 25 limbs only and print the result")
    ap.add_argument("--serve-only", action="store_true",
                    help="run the serving sessions only (synthetic keys, then bootstrap requests on real keys) and print them")
    ap.add_argument("--lt-only", action="store_true",
                    help="run the BSGS linear-transform workload only and print its result (for profiler runs)")
    ap.add_argument("--traffic-file", type=str, default=os.path.join(ROOT, "profiles", "traffic_rotation_n65536_l25.txt"),
                    help="rocprofv3 FETCH_SIZE + WRITE_SIZE bytes per launch of each kernel family of this workload")
    ap.add_argument("--autotune-polys", type=int, default=32,
                    help="run the NTT autotuner on this many synthetic polynomials before timing (0 = keep the defaults)")
    return ap.parse_args()


def algorithmic_bytes(family, N, lq, ext, beta, B, groups):
    """HBM bytes one launch of a kernel family must move at this shape (8-byte words)."""
    limb = 8 * N
    if family == "ntt_rows_ksk":
        # the fused row stages + key inner product: reads each digit's target limbs after the
        # column pass, the own limbs, the keys (once per group), c0; writes both extended outputs
        return limb * (B * (beta * ext - lq) + B * lq + groups * 2 * beta * ext + B * lq + 2 * B * ext)
    if family.startswith("ntt_"):
        return None  # depends on the launch; handled by the caller
    if family == "ksk_batch":
        digits = B * beta * ext
        keys = groups * 2 * beta * ext
        return limb * (digits + keys + B * lq + 2 * B * ext)
    if family == "modup_conv":
        return limb * B * (2 * lq + beta * ext - lq)
    if family == "moddown_conv":
        return limb * 2 * B * (K + lq)
    return None


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1", "n_gpus": a.gpus}))
        sys.exit(2)
    import fides_hip as fh

    ctx = fh.Context(LOGN, L1, K, DNUM)
    if a.lt_only:
        calibration_copy(fh, ctx, a.copy_mb)
        print(json.dumps({"linear_transform": lt_workload(fh, ctx, a), "accumulate": accumulate_run(fh, ctx, a)}))
        ctx.close()
        return
    if a.stream_only:
        keys, gals = [], []
        kb = ctx.array((DNUM * 2 * (L1 + K), ctx.N))
        for i, r in enumerate(STEPS16):
            ctx.fill_uniform(kb, list(range(L1 + K)), num_polys=DNUM * 2, seed=a.seed + 100 + i)
            keys.append(ctx.key_from_device(kb))
            gals.append(ctx.galois_index(r))
        kb.free()
        print(json.dumps({"uneven_stream": uneven_stream(fh, ctx, keys, gals, a)}))
        ctx.close()
        return
    if a.serve_only:
        keys, gals = [], []
        kb = ctx.array((DNUM * 2 * (L1 + K), ctx.N))
        for i, r in enumerate(STEPS16[:8]):
            ctx.fill_uniform(kb, list(range(L1 + K)), num_polys=DNUM * 2, seed=a.seed + 100 + i)
            keys.append(ctx.key_from_device(kb))
            gals.append(ctx.galois_index(r))
        kb.free()
        serving = serving_sweep(ctx, keys, gals, keys[0], a)
        if not a.no_bootstrap:
            serving["with_bootstrap"] = serving_bootstrap_run(a)
        print(json.dumps({"serving": serving}))
        ctx.close()
        return
    if a.ntt_only:
        for kv in a.tune:
            name, value = kv.split("=", 1)
            ctx.tune_set(name, int(value))
        print(json.dumps({"ntt": ntt_run(fh, ctx, a)}))
        ctx.close()
        return
    autotune = ctx.autotune(a.autotune_polys) if a.autotune_polys > 0 else None
    for kv in a.tune:
        name, value = kv.split("=", 1)
        ctx.tune_set(name, int(value))
    tunables = {name: ctx.tune_get(name) for name in ("ntt_split", "ntt_fwd_cols_threads", "ntt_fwd_rows_threads", "ntt_inv_rows_threads",
                                                      "ntt_inv_cols_threads", "ntt_chunk_mb", "ntt_mix", "rot_streams", "ntt_streams",
                                                      "rot_fused")}
    N, B, lvl = ctx.N, a.batch, L1 - 1
    lq, ext, beta = lvl + 1, lvl + 1 + K, ctx.num_digits(lvl)
    ct_words = 2 * lq * N
    rng = np.random.default_rng(a.seed)

    # ---- a plain device copy of known size: the reference point that scales hardware byte counters ----
    copy_gbs = copy_dma_gbs = None
    if a.copy_mb > 0:
        words = a.copy_mb * (1 << 17)
        src, dst = ctx.array((words,)), ctx.array((words,))
        lib = fh.load()
        lib.fh_copy_d2d(ctx.h, dst.ptr, src.ptr, words * 8)
        ctx.sync()
        t0 = time.perf_counter()
        lib.fh_copy_d2d(ctx.h, dst.ptr, src.ptr, words * 8)
        ctx.sync()
        copy_dma_gbs = round(2 * words * 8 / (time.perf_counter() - t0) / 1e9, 1)
        # STREAM copy kernel: read + write of the buffer, the kernel timed with HIP events on the
        # context stream (average of 5 launches after one untimed)
        fh._check(lib.fh_stream_copy(ctx.h, dst.ptr, src.ptr, words * 8))
        ctx.sync()
        ctx.profile(True)
        ctx.profile_reset()
        for _ in range(5):
            fh._check(lib.fh_stream_copy(ctx.h, dst.ptr, src.ptr, words * 8))
        ctx.sync()
        nl, tms = ctx.profile_read("stream_copy")
        ctx.profile(False)
        copy_gbs = round(2 * words * 8 / (tms / max(nl, 1) * 1e-3) / 1e9, 1)
        src.free()
        dst.free()

    # ---- synthetic inputs, resident in HBM ----
    qids = list(range(lq))
    cts = ctx.array((B * 2 * lq, N))
    ctx.fill_uniform(cts, qids, num_polys=2 * B, seed=a.seed)
    full = list(range(L1 + K))
    gal, dev_keys = [], []
    kbuf = ctx.array((DNUM * 2 * (L1 + K), N))
    for i, r in enumerate(STEPS16):
        ctx.fill_uniform(kbuf, full, num_polys=DNUM * 2, seed=a.seed + 100 + i)
        dev_keys.append(ctx.key_from_device(kbuf))
        gal.append(ctx.galois_index(r))
    kbuf.free()
    out_b = ctx.array((B * 2 * lq, N))
    out_p = ctx.array((B * 2 * lq, N))

    def choose(step):
        r = np.random.default_rng(a.seed * 1000 + step).integers(0, len(STEPS16), B)
        return [dev_keys[i] for i in r], [gal[i] for i in r], len(set(r.tolist()))

    # ---- bit-exactness: batched == per-op on the same inputs ----
    bit_exact = bit_exact_fused = None
    if not a.rotation_only:
        ks, gs, _ = choose(0)
        ctx.rotate_batch(cts, B, lvl, ks, gs, out_b)
        for b in range(B):
            ctx_ct = _view(fh, ctx, cts, b * ct_words, (2 * lq, N))
            ctx_out = _view(fh, ctx, out_p, b * ct_words, (2 * lq, N))
            ctx.eval_rotate(ctx_ct, lvl, ks[b], gs[b], ctx_out)
        ctx.sync()
        ref = out_b.download()
        bit_exact = bool(np.array_equal(ref, out_p.download()))   # all B ciphertexts, every limb
        # the same rotations with the targets' row stages fused into the key inner product
        ctx.tune_set("rot_fused", 1)
        try:
            ctx.rotate_batch(cts, B, lvl, ks, gs, out_p)
            ctx.sync()
        finally:
            ctx.tune_set("rot_fused", 0)
        bit_exact_fused = bool(np.array_equal(ref, out_p.download()))
        del ref

    # ---- timed: batched rotations ----
    for w in range(a.warmup):
        ks, gs, _ = choose(w + 1)
        ctx.rotate_batch(cts, B, lvl, ks, gs, out_b)
    ctx.sync()
    ctx.profile(True)
    ctx.profile_reset()
    groups_total = 0
    t0 = time.perf_counter()
    for s in range(a.steps):
        ks, gs, ng = choose(1000 + s)
        groups_total += ng
        ctx.rotate_batch(cts, B, lvl, ks, gs, out_b)
    ctx.sync()
    t_batched = time.perf_counter() - t0
    fams = ["ntt_inv_rows", "ntt_inv_cols", "ntt_fwd_cols", "ntt_fwd_rows", "modup_conv", "ksk_batch",
            "moddown_conv", "ntt_fwd_rows_moddown", "ntt_rows_ksk"]
    prof = {f: ctx.profile_read(f) for f in fams}
    ctx.profile(False)
    ms_per_step = 1000.0 * t_batched / a.steps
    rot_per_s = B * a.steps / t_batched

    # ---- the same timed step with the key inner product fused into the targets' row stages ----
    fused = None
    if not a.rotation_only:
        ctx.tune_set("rot_fused", 1)
        try:
            ks, gs, _ = choose(1)
            ctx.rotate_batch(cts, B, lvl, ks, gs, out_b)
            ctx.sync()
            ctx.profile(True)
            ctx.profile_reset()
            t0 = time.perf_counter()
            for s in range(a.steps):
                ks, gs, _ = choose(1000 + s)
                ctx.rotate_batch(cts, B, lvl, ks, gs, out_b)
            ctx.sync()
            t_fused = time.perf_counter() - t0
            fused_ms = ctx.profile_read("ntt_rows_ksk")[1] / a.steps
            ctx.profile(False)
        finally:
            ctx.tune_set("rot_fused", 0)
        fused = {"ms_per_step": round(1000.0 * t_fused / a.steps, 3), "rotations_per_s": round(B * a.steps / t_fused, 1),
                 "fused_kernel_ms_per_step": round(fused_ms, 3),
                 "separate_row_pass_and_ksk_ms_per_step": round((prof["ntt_fwd_rows"][1] + prof["ksk_batch"][1]) / a.steps, 3),
                 "bit_exact_vs_separate": bit_exact_fused,
                 "note": "tunable rot_fused (off by default): the ModUp targets' row stages and the key inner product in one kernel"}

    per_op_rot_s = ntt_gbs = t_ntt = mult_rescale_s = mult_rescale_batched_s = None
    pcie = None
    nprof = {}
    if not a.rotation_only:
        # ---- the same step with the ciphertexts handed over in host memory (PCIe both ways) ----
        host_in = cts.download()
        ks, gs, _ = choose(4000)
        ctx.sync()
        t0 = time.perf_counter()
        cts.upload(host_in)
        ctx.rotate_batch(cts, B, lvl, ks, gs, out_b)
        host_out = out_b.download()
        t_pcie = time.perf_counter() - t0
        pcie = {"rotations_per_s": round(B / t_pcie, 1), "ms_per_step": round(1000 * t_pcie, 3),
                "bytes_each_way": int(host_in.nbytes), "host_memory": "pageable numpy arrays"}
        del host_in, host_out
        # ---- per-op rotations (same inputs, one call per ciphertext) ----
        per_steps = max(1, min(a.steps, 3))
        for s in range(1):
            ks, gs, _ = choose(2000)
            for b in range(B):
                ctx.eval_rotate(_view(fh, ctx, cts, b * ct_words, (2 * lq, N)), lvl, ks[b], gs[b],
                                _view(fh, ctx, out_p, b * ct_words, (2 * lq, N)))
        ctx.sync()
        t0 = time.perf_counter()
        for s in range(per_steps):
            ks, gs, _ = choose(3000 + s)
            for b in range(B):
                ctx.eval_rotate(_view(fh, ctx, cts, b * ct_words, (2 * lq, N)), lvl, ks[b], gs[b],
                                _view(fh, ctx, out_p, b * ct_words, (2 * lq, N)))
        ctx.sync()
        per_op_rot_s = B * per_steps / (time.perf_counter() - t0)

        # ---- NTT GB/s: forward + inverse over all 64 x 2 x 25 limbs ----
        ctx.ntt_forward(out_b, qids, num_polys=2 * B)
        ctx.ntt_inverse(out_b, qids, num_polys=2 * B)
        ctx.sync()
        nrep = 5
        ctx.profile(True)
        ctx.profile_reset()
        t0 = time.perf_counter()
        for _ in range(nrep):
            ctx.ntt_forward(out_b, qids, num_polys=2 * B)
            ctx.ntt_inverse(out_b, qids, num_polys=2 * B)
        ctx.sync()
        t_ntt = (time.perf_counter() - t0) / nrep
        nprof = {f: ctx.profile_read(f) for f in ("ntt_fwd_cols", "ntt_fwd_rows", "ntt_inv_rows", "ntt_inv_cols")}
        ctx.profile(False)
        ntt_limbs = 2 * B * lq
        ntt_bytes = 2 * ntt_limbs * 16 * N  # forward + inverse, 16N bytes per limb-transform
        ntt_gbs = ntt_bytes / t_ntt / 1e9

        # ---- EvalMult + Rescale (relinearization key = first rotation key's buffer shape) ----
        prod = ctx.array((2 * lq, N))
        rs = ctx.array((2 * (lq - 1), N))
        ctx.eval_mult(_view(fh, ctx, cts, 0, (2 * lq, N)), _view(fh, ctx, cts, ct_words, (2 * lq, N)), lvl, dev_keys[0], prod)
        ctx.sync()
        nmul = 16
        t0 = time.perf_counter()
        for i in range(nmul):
            x = _view(fh, ctx, cts, (2 * i % B) * ct_words, (2 * lq, N))
            y = _view(fh, ctx, cts, ((2 * i + 1) % B) * ct_words, (2 * lq, N))
            ctx.eval_mult(x, y, lvl, dev_keys[0], prod)
            ctx.rescale(prod, lvl, 2, rs)
        ctx.sync()
        mult_rescale_s = nmul / (time.perf_counter() - t0)
        # the same products as one batch: fh_eval_mult_many (one tensor launch, one ModUp / key inner
        # product / ModDown for all) followed by one Rescale over every component
        nb = min(B // 2, 32)
        prods = ctx.array((nb * 2 * lq, N))
        rss = ctx.array((nb * 2 * (lq - 1), N))
        lo = _view(fh, ctx, cts, 0, (nb * 2 * lq, N))
        hi = _view(fh, ctx, cts, nb * ct_words, (nb * 2 * lq, N))
        def mult_batch():
            ctx.eval_mult_many(nb, lo, hi, lvl, dev_keys[0], prods)
            for j in range(0, 2 * nb, 8):   # Rescale takes up to 8 components per call
                nc = min(8, 2 * nb - j)
                ctx.rescale(_view(fh, ctx, prods, j * lq * N, (nc * lq, N)), lvl, nc,
                            _view(fh, ctx, rss, j * (lq - 1) * N, (nc * (lq - 1), N)))

        mult_batch()
        ctx.sync()
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            mult_batch()
        ctx.sync()
        mult_rescale_batched_s = reps * nb / (time.perf_counter() - t0)
        prods.free()
        rss.free()

    # ---- roofline of the dominant kernel of the timed rotation step ----
    dom = max(prof, key=lambda f: prof[f][1])
    launches, tot = prof[dom]
    avg_ms = tot / max(launches, 1)
    groups_avg = groups_total / a.steps
    pass_gbs = None
    if dom.startswith("ntt_") and dom != "ntt_rows_ksk":
        # Algorithmic bytes of a transform: 16N per limb-transform (read 8N + write 8N), of which
        # each of the two passes completes half: 8N per limb per launch. The ModDown rows launch
        # also reads x_i for its fused epilogue (8N per limb; its write is the transform's write).
        limbs_per_step = {"ntt_inv_rows": B * lq + 2 * B * K, "ntt_inv_cols": B * lq + 2 * B * K,
                          "ntt_fwd_cols": B * (beta * ext - lq) + 2 * B * lq, "ntt_fwd_rows": B * (beta * ext - lq),
                          "ntt_fwd_rows_moddown": 2 * B * lq}
        per_limb = 16 * N if dom == "ntt_fwd_rows_moddown" else 8 * N
        limbs = limbs_per_step[dom] * a.steps / max(launches, 1)
        alg = per_limb * limbs
        # each pass itself reads and writes every limb once: its own bandwidth, reported beside
        pass_bytes = (24 if dom == "ntt_fwd_rows_moddown" else 16) * N * limbs
        pass_gbs = pass_bytes / (avg_ms * 1e-3) / 1e9
    else:
        alg = algorithmic_bytes(dom, N, lq, ext, beta, B, groups_avg)
    achieved = alg / (avg_ms * 1e-3) / 1e9 if alg else None
    roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None, "traffic": None,
            "algorithmic_bytes_per_launch": int(alg) if alg else None, "avg_launch_ms": round(avg_ms, 4), "launches": launches,
            "kernel_ms_per_step": {f: round(prof[f][1] / a.steps, 3) for f in fams}}
    if pass_gbs is not None:
        roof["pass_gbps"] = round(pass_gbs, 1)
        roof["pass_note"] = "the launch's own reads and writes (every limb once each) per second; achieved counts 8N of a limb's 16N transform bytes per pass"
    ntt_kernel_ms = {f: round(nprof[f][1] / max(nprof[f][0], 1), 4) for f in nprof}
    # HBM bytes per launch of the dominant kernel, from the rocprofv3 counter runs of this workload
    traffic = read_traffic(a.traffic_file).get(dom)
    roof["traffic"] = traffic
    if traffic:
        roof["traffic_source"] = os.path.relpath(a.traffic_file, ROOT)

    # ---- mixed-level op stream: per-op vs one batch (eager and graph-replayed) ----
    full = not a.rotation_only
    stream = uneven_stream(fh, ctx, dev_keys, gal, a) if full and a.stream_ops > 0 else None
    lt = lt_workload(fh, ctx, a) if full and a.lt_count > 0 else None
    serving = serving_sweep(ctx, dev_keys[:8], gal[:8], dev_keys[0], a) if full and a.serve_requests > 0 else None
    enc = encode_run(fh, ctx) if full else None
    accum = accumulate_run(fh, ctx, a) if full else None
    boot = bootstrap_run() if full and not a.no_bootstrap else None
    if boot is not None and serving is not None:
        serving["with_bootstrap"] = serving_bootstrap_run(a)

    # ---- CPU baseline: the oracle on a bounded sample of the same rotation ----
    cpu = None
    if full and not a.no_cpu:
        cpu = cpu_baseline(a)

    line = {
        "metric": METRIC,
        "value": round(rot_per_s, 1),
        "unit": "rotations/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (uniform residues and keys generated on the device)",
        "seed": a.seed,
        "config": {"workload": "batched hybrid key-switch rotation, N=2^16, 25 limbs, dnum=3, K=9, 64 ciphertexts, 16 keyed steps",
                   "ring_dim": N, "limbs": lq, "dnum": DNUM, "special_primes": K, "batch": B, "keyed_steps": len(STEPS16),
                   "tunables": tunables, "autotune": autotune},
        "roofline": roof,
        "cpu_baseline": cpu,
        "per_op_rotations_per_s": round(per_op_rot_s, 1) if per_op_rot_s else None,
        "batched_over_per_op": round(rot_per_s / per_op_rot_s, 3) if per_op_rot_s else None,
        "bit_exact_batched_vs_per_op": bit_exact,
        "fused_rows_ksk": fused,
        "pcie_inclusive": pcie,
        "ntt_gbps": round(ntt_gbs, 1) if ntt_gbs else None,
        "ntt_roofline_frac": round(ntt_gbs / HBM_PEAK_GBS, 4) if ntt_gbs else None,
        "ntt_ms_fwd_inv": round(1000 * t_ntt, 3) if t_ntt else None,
        "ntt_avg_launch_ms": ntt_kernel_ms,
        "eval_mult_rescale_per_s": round(mult_rescale_s, 1) if mult_rescale_s else None,
        "eval_mult_rescale_batched_per_s": round(mult_rescale_batched_s, 1) if mult_rescale_batched_s else None,
        "device_copy_gbps": copy_gbs,
        "device_copy_source": "STREAM copy kernel (fh_stream_copy), bytes read + written, HIP-event kernel time, mean of 5" if copy_gbs else None,
        "device_copy_dma_gbps": copy_dma_gbs,
        "uneven_stream": stream,
        "linear_transform": lt,
        "serving": serving,
        "bootstrap": boot,
        "encode": enc,
        "accumulate": accum,
    }
    print(json.dumps(line))
    ctx.close()


def ntt_run(fh, ctx, a):
    """Forward + inverse NTT over 64 x 2 x 25 limbs (16N algorithmic bytes per limb-transform)."""
    N, B, lq = ctx.N, a.batch, L1
    qids = list(range(lq))
    buf = ctx.array((B * 2 * lq, N))
    ctx.fill_uniform(buf, qids, num_polys=2 * B, seed=a.seed)
    ctx.ntt_forward(buf, qids, num_polys=2 * B)
    ctx.ntt_inverse(buf, qids, num_polys=2 * B)
    ctx.sync()
    nrep = max(1, a.steps)
    ctx.profile(True)
    ctx.profile_reset()
    t0 = time.perf_counter()
    for _ in range(nrep):
        ctx.ntt_forward(buf, qids, num_polys=2 * B)
        ctx.ntt_inverse(buf, qids, num_polys=2 * B)
    ctx.sync()
    t = (time.perf_counter() - t0) / nrep
    fams = ("ntt_fwd_cols", "ntt_fwd_rows", "ntt_inv_rows", "ntt_inv_cols", "ntt_fwd_team", "ntt_inv_team")
    prof = {f: ctx.profile_read(f) for f in fams}
    ctx.profile(False)
    buf.free()
    limbs = 2 * B * lq
    gbs = 2 * limbs * 16 * N / t / 1e9
    return {"limb_transforms": 2 * limbs, "ms_fwd_inv": round(1000 * t, 3), "gbps": round(gbs, 1),
            "roofline_frac": round(gbs / HBM_PEAK_GBS, 4), "team": ctx.tune_get("ntt_team"),
            "team_per_xcd": ctx.tune_get("ntt_team_per_xcd"), "teams_spanning_xcds": ctx.tune_get("ntt_team_split_teams"),
            "avg_launch_ms": {f: round(v[1] / v[0], 4) for f, v in prof.items() if v[0]}}


def calibration_copy(fh, ctx, mb):
    """One STREAM copy kernel (k_stream_copy) of mb MiB: in a counter run, the reference whose
    FETCH_SIZE / WRITE_SIZE readings scale the other kernels' (it reads and writes mb MiB)."""
    if mb <= 0:
        return
    words = mb * (1 << 17)
    src, dst = ctx.array((words,)), ctx.array((words,))
    fh._check(fh.load().fh_stream_copy(ctx.h, dst.ptr, src.ptr, words * 8))
    ctx.sync()
    src.free()
    dst.free()


def read_traffic(path):
    """kernel family -> HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, scaled by a device copy of
    known size), from a profile summary: lines 'family<TAB>bytes', '#' comments."""
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        line = line.strip()
        if not line or line.startswith("#"):
            continue
        fam, val = line.split()[:2]
        out[fam] = int(val)
    return out


def uneven_stream(fh, ctx, rot_keys, rot_gal, a):
    """A seeded stream of ciphertexts at levels uniform in [1, 24]: 40% EvalRotate (16 keyed steps),
    40% EvalMult + relinearization + Rescale, 20% EvalAdd. Timed one operation per call and as one
    mixed-level batch (eager, and replayed from a captured HIP graph); outputs compared bit for bit."""
    N, n = ctx.N, a.stream_ops
    rng = np.random.default_rng(a.seed + 7)
    kinds = rng.choice(3, size=n, p=[0.4, 0.4, 0.2])          # 0 rotate, 1 mult+rescale, 2 add
    levels = rng.integers(1, L1, size=n)
    keyi = rng.integers(0, len(rot_keys), size=n)
    relin = rot_keys[0]  # a key of the relinearization shape (synthetic residues)
    ins, outs_p, outs_b = [], [], []
    for i in range(n):
        lq = int(levels[i]) + 1
        x = ctx.array((2 * lq, N))
        y = ctx.array((2 * lq, N))
        ctx.fill_uniform(x, list(range(lq)) * 2, num_limbs=2 * lq, seed=a.seed + 10000 + 2 * i)
        ctx.fill_uniform(y, list(range(lq)) * 2, num_limbs=2 * lq, seed=a.seed + 10001 + 2 * i)
        olimbs = 2 * (lq - 1) if kinds[i] == 1 else 2 * lq
        ins.append((x, y))
        outs_p.append(ctx.array((olimbs, N)))
        outs_b.append(ctx.array((olimbs, N)))
    tmp = ctx.array((2 * L1, N))
    lib = fh.load()

    def per_op():
        for i in range(n):
            l = int(levels[i])
            x, y = ins[i]
            o = outs_p[i]
            if kinds[i] == 0:
                ctx.eval_rotate(x, l, rot_keys[keyi[i]], rot_gal[keyi[i]], o)
            elif kinds[i] == 1:
                ctx.eval_mult(x, y, l, relin, tmp)
                ctx.rescale(tmp, l, 2, o)
            else:
                ids = list(range(l + 1)) * 2
                ctx.eltwise(0, x, y, o, prime_ids=ids)

    b = fh.Batch(ctx)
    for i in range(n):
        l = int(levels[i])
        x, y = ins[i]
        if kinds[i] == 0:
            b.rotate(x.ptr, l, rot_keys[keyi[i]], rot_gal[keyi[i]], outs_b[i].ptr)
        elif kinds[i] == 1:
            b.mult(x.ptr, y.ptr, l, relin, True, outs_b[i].ptr)
        else:
            b.add(x.ptr, y.ptr, l, 2, outs_b[i].ptr)
    b.finalize()
    _, scratch = b.info()

    def timed(fn, reps):
        fn()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        ctx.sync()
        return (time.perf_counter() - t0) / reps

    t_per = timed(per_op, 2)
    t_eager = timed(lambda: b.run(graph=False), 3)
    t_graph = timed(lambda: b.run(graph=True), 3)
    same = all(np.array_equal(outs_p[i].download(), outs_b[i].download()) for i in range(n))
    for x, y in ins:
        x.free()
        y.free()
    for o in outs_p + outs_b:
        o.free()
    tmp.free()
    b.free()
    counts = {k: int((kinds == j).sum()) for j, k in enumerate(("rotate", "mult_relin_rescale", "add"))}
    return {"ops": n, "mix": counts, "levels": "uniform in [1, 24]",
            "per_op_ops_per_s": round(n / t_per, 1), "batched_ops_per_s": round(n / t_eager, 1),
            "batched_graph_ops_per_s": round(n / t_graph, 1), "batched_over_per_op": round(t_per / t_eager, 3),
            "graph_over_per_op": round(t_per / t_graph, 3), "bit_exact_batched_vs_per_op": bool(same),
            "batch_scratch_gib": round(scratch / 2 ** 30, 2)}


def lt_workload(fh, ctx, a):
    """Hoisted BSGS plaintext-matrix x ciphertext product: level 20 (21 limbs), 64 nonzero
    diagonals resident over Q_20 ∪ P (960 MiB), bStep = gStep = 8, stride 1; lt_count ciphertexts
    against the one resident diagonal set, one call per ciphertext and as one batched call
    (bit-exact comparison). Synthetic residues for keys, diagonals and ciphertexts (the LT
    arithmetic is data-independent)."""
    N, lvl, row, bstep = ctx.N, 20, 64, 8
    lq, ext = lvl + 1, lvl + 1 + K
    full = list(range(L1 + K))
    kbuf = ctx.array((DNUM * 2 * (L1 + K), N))
    keys = {}
    gsteps = [j * bstep for j in range(1, row // bstep)]  # the flat form's giant multiples
    for i, r in enumerate(list(range(1, bstep)) + gsteps):
        ctx.fill_uniform(kbuf, full, num_polys=DNUM * 2, seed=a.seed + 500 + i)
        keys[r] = ctx.key_from_device(kbuf)
    kbuf.free()
    ext_ids = list(range(lq)) + [L1 + k for k in range(K)]
    diags = ctx.array((row * ext, N))
    ctx.fill_uniform(diags, ext_ids, num_polys=row, seed=a.seed + 600)
    cts = ctx.array((a.lt_count * 2 * lq, N))
    ctx.fill_uniform(cts, list(range(lq)), num_polys=2 * a.lt_count, seed=a.seed + 700)
    out = ctx.array((2 * lq, N))
    bk = [None] + [keys[i] for i in range(1, bstep)]
    bg = [1] + [ctx.galois_index(i) for i in range(1, bstep)]
    gk, gg = keys[bstep], ctx.galois_index(bstep)

    def one(i):
        ctx.linear_transform(_view(fh, ctx, cts, i * 2 * lq * N, (2 * lq, N)), lvl, row, bstep, 1, diags, bk, bg,
                             _view(fh, ctx, outs, i * 2 * lq * N, (2 * lq, N)), giant_key=gk, giant_galois=gg)

    outs = ctx.array((a.lt_count * 2 * lq, N))
    outb = ctx.array((a.lt_count * 2 * lq, N))
    one(0)
    ctx.sync()
    t0 = time.perf_counter()
    for i in range(a.lt_count):
        one(i)
    ctx.sync()
    dt = time.perf_counter() - t0

    def batched():
        ctx.linear_transform_batch(cts, a.lt_count, lvl, row, bstep, 1, diags, bk, bg, outb, giant_key=gk, giant_galois=gg)

    batched()
    ctx.sync()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        batched()
    ctx.sync()
    dtb = (time.perf_counter() - t0) / reps
    same = bool(np.array_equal(outs.download(), outb.download()))

    # the flat giant-step form (each multiple of the giant step rotated by its own key, one batched
    # key switch for all 7), one call per ciphertext: the same map with its own rounding
    def flat(i):
        ctx.linear_transform_flat(_view(fh, ctx, cts, i * 2 * lq * N, (2 * lq, N)), lvl, row, bstep, 1, diags, bk, bg,
                                  _view(fh, ctx, outs, i * 2 * lq * N, (2 * lq, N)), giant_keys=[keys[r] for r in gsteps],
                                  giant_galois=[ctx.galois_index(r) for r in gsteps])

    flat(0)
    ctx.sync()
    t0 = time.perf_counter()
    for i in range(a.lt_count):
        flat(i)
    ctx.sync()
    dtf = time.perf_counter() - t0
    for x in [diags, cts, out, outs, outb] + list(keys.values()):
        x.free()
    # algorithmic bytes per transform: 15 key switches' keys, 64 diagonals, ciphertext in/out
    beta = ctx.num_digits(lvl)
    alg = 8 * N * (15 * 2 * beta * ext + row * ext + 4 * lq)
    # batched: keys and diagonals read once for the whole batch
    alg_b = 8 * N * (15 * 2 * beta * ext + row * ext + a.lt_count * 4 * lq)
    return {"transforms": a.lt_count, "level": lvl, "diagonals": row, "b_step": bstep, "g_step": row // bstep,
            "per_call_transforms_per_s": round(a.lt_count / dt, 2), "per_call_ms_per_transform": round(1000 * dt / a.lt_count, 3),
            "batched_transforms_per_s": round(a.lt_count / dtb, 2), "batched_ms_per_transform": round(1000 * dtb / a.lt_count, 3),
            "batched_over_per_call": round(dt / dtb, 3), "bit_exact_batched_vs_per_call": same,
            "flat_per_call_ms_per_transform": round(1000 * dtf / a.lt_count, 3),
            "flat_note": "fh_linear_transform_flat per call: 7 giant-step keys instead of 1, the 7 giant rotations as one batch",
            "algorithmic_gb_per_transform": round(alg / 1e9, 3), "algorithmic_gb_per_batch": round(alg_b / 1e9, 3)}


def accumulate_run(fh, ctx, a, reps=5):
    """Rotation-fold sum of one ciphertext at level 24 over 256 summands (the dot product of a
    256-feature request): fh_accumulate_sum, the lazy fold at radix 4 (12 rotations sharing 4 ModUps,
    c1 taking one ModDown per fold level and c0 one at the end) against the doubling loop of 8
    rotations and additions.
    Synthetic keys and residues (the arithmetic is data-independent)."""
    N, lvl, size = ctx.N, 24, 256
    lq = lvl + 1
    full = list(range(L1 + K))
    idx4 = fh.Context.accumulate_indices(4, 1, size)
    steps = sorted(set(idx4) | {1 << i for i in range(8)})
    kbuf = ctx.array((DNUM * 2 * (L1 + K), N))
    keys = {}
    for i, r in enumerate(steps):
        ctx.fill_uniform(kbuf, full, num_polys=DNUM * 2, seed=a.seed + 900 + i)
        keys[r] = ctx.key_from_device(kbuf)
    kbuf.free()
    ct = ctx.array((2 * lq, N))
    ctx.fill_uniform(ct, list(range(lq)), num_polys=2, seed=a.seed + 950)
    out, tmp, acc = ctx.array((2 * lq, N)), ctx.array((2 * lq, N)), ctx.array((2 * lq, N))
    ks, gs = [keys[r] for r in idx4], [ctx.galois_index(r) for r in idx4]
    pids = list(range(lq)) * 2

    def fold():
        ctx.accumulate_sum(ct, lvl, 4, 1, size, ks, gs, out)

    def doubling():
        src = ct
        for i in range(8):
            r = 1 << i
            ctx.eval_rotate(src, lvl, keys[r], ctx.galois_index(r), tmp)
            ctx.eltwise(0, src, tmp, acc, pids)
            src = acc

    res = {}
    for name, fn in (("fold_radix4", fold), ("doubling", doubling)):
        fn()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        ctx.sync()
        res[name] = 1000 * (time.perf_counter() - t0) / reps
    for x in [ct, out, tmp, acc] + list(keys.values()):
        x.free()
    return {"summands": size, "level": lvl, "radix4_rotations": len(idx4), "doubling_rotations": 8,
            "fold_radix4_ms": round(res["fold_radix4"], 3), "doubling_ms": round(res["doubling"], 3),
            "doubling_over_fold": round(res["doubling"] / res["fold_radix4"], 3)}


def encode_run(fh, ctx, count=64, reps=5):
    """Device CKKS encoding (fh_encode_slots) of 64 full-slot plaintexts (32768 complex slots) at
    N=2^16 over 25 limbs: one batched call and one call per plaintext, best of 5. Algorithmic bytes
    per plaintext: 16 B per slot uploaded + 25 limbs written and transformed (the NTT's 16N B per
    limb)."""
    import time
    N, lq = ctx.N, L1
    rng = np.random.default_rng(11)
    z = rng.uniform(-1, 1, (count, N // 2)) + 1j * rng.uniform(-1, 1, (count, N // 2))
    out = ctx.array((count * lq, N))
    pids = list(range(lq))
    one = [_view(fh, ctx, out, b * lq * N, (lq, N)) for b in range(count)]

    def batched():
        ctx.encode_slots(z, 2.0 ** 50, pids, out)

    def per_call():
        for b in range(count):
            ctx.encode_slots(z[b:b + 1], 2.0 ** 50, pids, one[b])

    res = {}
    for name, fn in (("batched", batched), ("per_call", per_call)):
        fn()
        ctx.sync()
        best = 1e30
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ctx.sync()
            best = min(best, time.perf_counter() - t0)
        res[name] = best
    # device time of one batched call: the inverse FFT, rounding and residues ("encode") and the NTT
    ctx.profile(True)
    ctx.profile_reset()
    batched()
    dev = {f: ctx.profile_read(f)[1] for f in ("encode", "ntt_fwd_cols", "ntt_fwd_rows")}
    ctx.profile(False)
    out.free()
    per_pt = 16 * (N // 2) + lq * 16 * N
    return {"plaintexts": count, "slots": N // 2, "limbs": lq,
            "device_ms": {"ifft_round_residues": round(dev["encode"], 3), "ntt": round(dev["ntt_fwd_cols"] + dev["ntt_fwd_rows"], 3)},
            "note": "batched_ms includes the host check and the 16 B/slot upload of the values (33.5 MB here)",
            "batched_plaintexts_per_s": round(count / res["batched"], 1), "batched_ms": round(1000 * res["batched"], 3),
            "per_call_plaintexts_per_s": round(count / res["per_call"], 1),
            "batched_gbps_algorithmic": round(count * per_pt / res["batched"] / 1e9, 1),
            "upload_bytes_per_plaintext": 16 * (N // 2), "host_encoded_upload_bytes_per_plaintext": lq * 8 * N}


def bootstrap_run():
    """EvalBootstrap through the C++ facade (tests/cpp/bootstrap_check.cpp) at N=2^16 with all
    32768 slots, level budget {3, 3}, 59-bit scaling under a 60-bit q_0 (correction factor 9),
    sparse ternary secret h=32 (one double angle), 25 towers: median latency of three bootstraps
    and the decryption error of the bootstrapped slots; the same for the two-iteration bootstrap
    (Meta-BTS) at precision 12."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "fideslib-fork_amd", "_build", "bootstrap_check")
    if not os.path.exists(exe):
        return {"error": "bootstrap_check is not built"}
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([exe, d, "16", "32768", "7", "3", "3", "24", "59", "32", "0", "12"], capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        return {"error": (r.stdout + r.stderr)[-500:]}
    res = json.loads(r.stdout.strip().splitlines()[-1])
    out = {"workload": "N=2^16, 32768 slots, level budget {3,3}, 59-bit scaling, sparse ternary h=32", "ms": res["boot_ms"],
           "bits": res["boot_bits"], "towers_in": res["towers_in"], "towers_out": res["towers_out"], "depth": res["depth"],
           "deterministic": res["deterministic"],
           "meta_bts": {"iterations": 2, "precision": res["meta_precision"], "bits": res["meta_bits"], "ms": res["meta_ms"]}}
    # the reference's own precision point (test/BootstrapPrecisionTests.cu): depth 25, 55-bit
    # scaling under a 56-bit q_0, budget {3,2}, dnum 6, FLEXIBLEAUTO, weight-192 secret, real slots
    # encrypted at level 21; the reference states 18.6 bits (OpenFHE), 11.5-15 (its GPU path) and
    # 29-32 (its GPU Meta-BTS) there
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([exe, d, "16", "32768", "7", "3", "2", "25", "55", "192", "1", "12", "56", "6", "1", "21"], capture_output=True,
                           text=True, timeout=600)
    if r.returncode != 0:
        out["reference_point"] = {"error": (r.stdout + r.stderr)[-500:]}
    else:
        res = json.loads(r.stdout.strip().splitlines()[-1])
        out["reference_point"] = {
            "workload": "N=2^16, 32768 real slots at level 21, budget {3,2}, depth 25, 55-bit scaling, q_0 56 bits, dnum 6, FLEXIBLEAUTO, h=192",
            "ms": res["boot_ms"], "bits": res["boot_bits"], "meta_bts_bits": res["meta_bits"], "meta_bts_ms": res["meta_ms"],
            "reference_states_bits": {"openfhe": 18.6, "gpu": [11.5, 15], "gpu_meta_bts": [29, 32]}}
    return out


BOOT_SERVE_REQUESTS = 200


def serving_bootstrap_run(a):
    """Serving sessions whose arrivals include bootstrap requests (tests/cpp/serve_check.cpp): the
    bootstrap context above under FLEXIBLEAUTO (N=2^16, 25 towers, full-slot bootstrapping, h=32),
    the mix of 70% inference, 25% polynomial and 5% bootstrap requests at 50 requests/s with the
    bootstraps in the tick, on the loop's worker thread and on its yielding worker thread, and the
    yielding placement again at 150 requests/s; batched and per-operation, the same checksum and
    every result bitwise the same alone and in the mix."""
    import subprocess
    exe = os.path.join(ROOT, "fideslib-fork_amd", "_build", "serve_check")
    if not os.path.exists(exe):
        return {"error": "serve_check is not built"}
    keep = ("p50_ms", "p99_ms", "p50_other_ms", "p99_other_ms", "mean_ms", "requests_per_s", "mean_batch_ops", "bootstraps", "inference",
            "polynomial")
    out = {"rate_per_s": 50.0, "requests": BOOT_SERVE_REQUESTS, "mix": "70% inference, 25% polynomial, 5% bootstrap",
           "technique": "FLEXIBLEAUTO", "p99_other": "p99 over the requests other than bootstraps"}
    # serial: each bootstrap runs in the tick that reaches it; concurrent: on the loop's worker
    # thread through a second context (same keys, its own stream) while the other requests tick;
    # yielding: the worker issues one bootstrap operation at a time, in the gaps between ticks.
    # The last session repeats the yielding placement at 150 requests/s.
    for name, conc, rate in (("serial_bootstrap", "0", 50), ("concurrent_bootstrap", "1", 50), ("yielding_bootstrap", "2", 50),
                             ("yielding_bootstrap_150", "2", 150)):
        args = ["16", "32768", "3", "3", "24", "59", str(rate), str(BOOT_SERVE_REQUESTS), "0.05", str(a.seed), "0.70", "32", "1", conc]
        r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            out[name] = {"error": (r.stdout + r.stderr)[-500:]}
            continue
        res = json.loads(r.stdout.strip().splitlines()[-1])
        out["towers"] = res["L1"]
        out[name] = {"rate_per_s": float(rate), "batched": {k: res["batched"][k] for k in keep}, "per_op": {k: res["per_op"][k] for k in keep},
                     "same_checksum": res["batched"]["checksum"] == res["per_op"]["checksum"],
                     "bitwise_same_alone_and_in_mix": res["bitwise_same_alone_and_in_mix"]}
    return out


def serving_sweep(ctx, rot_keys, rot_gal, relin, a):
    """Continuous-batching serving sessions (fh_serve_run) at N=2^16, requests arriving at level 24:
    74% inference (x*w + Rescale, rotate-and-sum over 8 doublings, degree-7 polynomial), 26%
    depth-6 polynomial; p50/p99 latency and requests/s per arrival rate, batched, plus the
    one-call-per-operation baseline at the middle rate. Keys are synthetic residues."""
    assert STEPS16[:8] == [1 << k for k in range(8)]
    rates = [float(r) for r in a.serve_rates.split(",") if r]
    out = {"requests": a.serve_requests, "mix": "74% inference (dot product + degree-7 polynomial), 26% depth-6 polynomial",
           "bootstrap_share": "0 in these sessions (synthetic keys); with_bootstrap: a session with bootstrap requests on real keys", "sessions": []}
    keep = ("p50_ms", "p99_ms", "mean_ms", "requests_per_s", "mean_batch_ops", "max_active", "ticks")
    out["ticks_in_flight"] = "ticks issued ahead of the GPU: 3 by default, 1 = the host waits for every tick"
    # a discarded warm-up session (batched, at a rate that fills batches) takes the first-use
    # allocations of the context pool and the batch's table staging out of the measured sessions
    ctx.serve(rot_keys, rot_gal, relin, 2000.0, 24, mode=0, seed=a.seed + 99)
    for i, rate in enumerate(rates):
        r = ctx.serve(rot_keys, rot_gal, relin, rate, a.serve_requests, mode=0, seed=a.seed + i)
        out["sessions"].append(dict({"rate_per_s": rate, "mode": "batched", "ticks_in_flight": 3, "batch_min_ops": 3},
                                    **{k: round(r[k], 3) for k in keep}))
    mid = rates[len(rates) // 2]
    for mode, tif, bmin in ((0, 1, 3), (0, 3, 2), (1, 3, 3), (1, 1, 3)):
        r = ctx.serve(rot_keys, rot_gal, relin, mid, a.serve_requests, mode=mode, seed=a.seed + len(rates) // 2, ticks_in_flight=tif,
                      batch_min_ops=bmin)
        s = {"rate_per_s": mid, "mode": ["batched", "per_op"][mode], "ticks_in_flight": tif}
        if mode == 0:
            s["batch_min_ops"] = bmin
        out["sessions"].append(dict(s, **{k: round(r[k], 3) for k in keep}))
    return out


def _view(fh, ctx, arr, word_offset, shape):
    """A non-owning DeviceArray aliasing part of `arr`."""
    v = fh.DeviceArray.__new__(fh.DeviceArray)
    v.ctx, v.shape = ctx, tuple(shape)
    v.words = int(np.prod(shape))
    v.ptr = arr.ptr + 8 * word_offset
    return v


def cpu_baseline(a):
    """The CPU oracle timed on the GPU host: hoisted-formulation rotations at the same shape."""
    from oracle import oracle as O

    primes = O.gen_primes(LOGN, L1, 60, 50, K)
    o = O.Ctx(LOGN, primes, L1, K, DNUM)
    lvl = L1 - 1
    ct = np.concatenate([o.random_poly(1, list(range(L1))), o.random_poly(2, list(range(L1)))])
    key = o.random_poly(3, list(range(L1 + K)) * (2 * DNUM))
    k = o.rot_index(1)
    o.eval_rotate(ct, lvl, key, k)  # warm-up
    n, t0 = 0, time.perf_counter()
    while n < 64 and time.perf_counter() - t0 < a.cpu_seconds:
        o.eval_rotate(ct, lvl, key, k)
        n += 1
    dt = time.perf_counter() - t0
    threads = O.lib().orc_num_threads()

    def rate(fn, budget, cap):
        fn()  # warm-up
        m, t = 0, time.perf_counter()
        while m < cap and time.perf_counter() - t < budget:
            fn()
            m += 1
        return m, time.perf_counter() - t

    # the other op classes of the single-op workload, each on a bounded sample
    ids = list(range(L1))
    poly2 = ct.copy()
    m_ntt, t_ntt = rate(lambda: o.intt(o.ntt(poly2, ids * 2), ids * 2), a.cpu_seconds / 3, 64)
    ct2 = np.concatenate([o.random_poly(4, ids), o.random_poly(5, ids)])
    m_mul, t_mul = rate(lambda: o.rescale_ct(o.eval_mult(ct, ct2, lvl, key), lvl), a.cpu_seconds / 3, 64)
    per_class = {
        "ntt_limb_transforms_per_s": round(2 * 2 * L1 * m_ntt / t_ntt, 2),
        "eval_mult_rescale_per_s": round(m_mul / t_mul, 3),
        "rotations_per_s": round(n / dt, 3),
        "samples": f"{m_ntt} forward+inverse NTTs of a 2 x 25-limb ciphertext ({t_ntt:.1f} s), {m_mul} EvalMult+Rescale at level 24 "
                   f"({t_mul:.1f} s), {n} rotations ({dt:.1f} s)",
    }
    return {"value": round(n / dt, 3), "unit": "rotations/s", "cores": threads, "kind": "port",
            "sample": f"{n} rotations of one N=2^16, 25-limb ciphertext (dnum=3, K=9) by the C oracle, {dt:.1f} s",
            "per_op_class": per_class}


if __name__ == "__main__":
    main()
