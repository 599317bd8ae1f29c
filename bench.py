"""Benchmark of the plain FP32 polar-FT correlation search on one MI355X.

Workload (2D-classification shape): 200 class references x 100,000 particles, 200-px box,
1.0 A/px, kfromto [2,60] (pftsz 308, 616 in-plane rotations), exhaustive search over all
references x rotations x a +-3 px shift grid at 0.5 px (169 shifts), CTF-weighted, all
particles resident in HBM as complex64 polar arrays.  Synthetic data: Gaussian-blob phantom
projections, random in-plane angle and shift, astigmatic CTF (defocus U(0.8,3.5) um), SNR 0.03.

One step = the search of one batch of particles (the device reduces each particle to its best
(reference, shift, rotation)).  The metric counts particle x reference x rotation x shift
correlation evaluations per second.

    python bench.py [--steps K] [--warmup W] [--batch B] [--nptcls N] [--no-cpu]

Prints one JSON line (see the keys below) with the dominant kernel's roofline position and the
reference CPU algorithm timed on this host's cores.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "simple3.0_amd"))

METRIC = ("ptcl×ref×rot corr evals/s: plain FP32 vs split-BF16 MFMA @ rel err ≤1e-4")
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0
# HBM bytes per launch of k_corr_fused<308,9,256,3,true> at 1000 particles x 200 refs x 169 shifts,
# from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of this script (--legs plain --copy-calibration):
# FETCH 411073084 KiB x 2.000 + WRITE 117375 KiB x 1.000, the factors read from k_copy moving
# 1 GiB in the same runs (profiles/k_corr_fused_308_plain_search.md).
PLAIN_TRAFFIC_BYTES_B1000 = (411073084.0 * 2.0 + 117375.0) * 1024.0
FP64_PEAK_TFLOPS = 78.6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def plain_flops_per_pair(n, K):
    """Algorithmic work of the reference FFT route per (ref, ptcl, shift) pair:
    K*(12.5 n log2 n + 25 n) (r2c_n + c2c_n + c2r_2n per shell at 5 N log2 N, plus products)."""
    return K * (12.5 * n * math.log2(n) + 25.0 * n)


def mfma_flops_per_pair(n, kcorr):
    """Algorithmic work of the per-angular-frequency contraction per (ref, ptcl) pair: one complex
    multiply-add (8 real flops) per shell for each of the n+1 angular frequencies."""
    return 8.0 * kcorr * (n + 1)


def precision_form(steps, nrefs=4684, nptcls=20000, plain_sample=2000, ntop=4):
    """Split-BF16 MFMA form at the 3D-refinement shape, next to the plain FP32 form on the same data."""
    import pftcc
    from pftcc import Params, PolarFTCorrCalc, PFTCC_PREC_FP16, PFTCC_PREC_SPLIT_BF16, PFTCC_PREC_SPLIT_FP16
    from pftcc import synth

    box, smpd, kfromto = 256, 1.06, (2, 80)
    pftsz = pftcc.magic_pftsz(126)
    w = synth.Workload(box=box, smpd=smpd, kfromto=kfromto, pftsz=pftsz, nrefs=nrefs, nptcls=nptcls, snr=0.03,
                       seed=20251203, dfrange=(0.8, 3.5), astig=0.1, max_shift=0.0, uniform_dirs=True)
    synth.make_params(w)
    lib = PolarFTCorrCalc().new(nrefs, (1, nptcls), Params(boxmatch=box, kfromto=kfromto, pftsz=pftsz, ctf=True))
    rbuf_refs = synth.DeviceBuffer(nrefs * w.nk * pftsz * 8)
    synth.gen_refs(w, rbuf_refs)
    t0 = time.perf_counter()
    synth.load_into(lib, w, chunk=5000)
    log(f"precision form setup: {nptcls} particles x {nrefs} refs in {time.perf_counter() - t0:.1f} s")
    S = synth.synth_lib()
    ids = np.arange(1, nptcls + 1, dtype=np.int32)
    nrots = 2 * pftsz
    kcorr = w.nk
    lib.search_mfma(ids[:256], prec=PFTCC_PREC_SPLIT_BF16, ntop=ntop)      # warm-up (code objects)
    S.synth_device_sync()
    lib.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        # new references every step, as every refinement iteration: the reference operand is rebuilt
        lib.set_ref_pfts_dev(1, nrefs, rbuf_refs.ptr, True)
        lib.set_ref_pfts_dev(1, nrefs, rbuf_refs.ptr, False)
        cc, ir, rot = lib.search_mfma(ids, prec=PFTCC_PREC_SPLIT_BF16, ntop=ntop)
    S.synth_device_sync()
    wall = (time.perf_counter() - t0) / steps
    ch = {name: lib.get_timing_channel(c) for c, name in ((1, "zgemm"), (2, "zifft"), (3, "pack_rescore"))}
    lib.set_timing(False)
    # the FP16 operand formats on the same search (one step each, reference operand rebuilt)
    others = {}
    for nm, pr in (("fp16", PFTCC_PREC_FP16), ("split_fp16", PFTCC_PREC_SPLIT_FP16)):
        lib.search_mfma(ids[:256], prec=pr, ntop=ntop)
        S.synth_device_sync()
        lib.set_timing(True)
        t0 = time.perf_counter()
        lib.set_ref_pfts_dev(1, nrefs, rbuf_refs.ptr, True)
        lib.set_ref_pfts_dev(1, nrefs, rbuf_refs.ptr, False)
        occ, oir, orot = lib.search_mfma(ids, prec=pr, ntop=ntop)
        S.synth_device_sync()
        owall = time.perf_counter() - t0
        others[nm] = {"value": nptcls * nrefs * nrots / owall, "unit": "evals/s", "ms_per_step": 1e3 * owall,
                      "contraction_ms": lib.get_timing_channel(1)[0],
                      "assignment_agreement_vs_split_bf16": float(np.mean((oir == ir) & (orot == rot)))}
        lib.set_timing(False)
    rbuf_refs.close()
    value = nptcls * nrefs * nrots / wall
    raw_cc, raw_ir, raw_rot = lib.search_mfma(ids[:plain_sample], prec=PFTCC_PREC_SPLIT_BF16, ntop=0)
    # the plain FP32 form on a sample of the same particles
    lib.search(ids[:64])
    S.synth_device_sync()
    t0 = time.perf_counter()
    pcc, pir, _, prot = lib.search(ids[:plain_sample])
    S.synth_device_sync()
    pwall = time.perf_counter() - t0
    plain_value = plain_sample * nrefs * nrots / pwall
    same = (pir == ir[:plain_sample]) & (prot == rot[:plain_sample])
    same_raw = (pir == raw_ir) & (prot == raw_rot)
    log("precision form: FP64 checker on a sample")
    cpu = precision_cpu_baseline(lib, nrefs, kcorr, PFTCC_PREC_SPLIT_BF16, ntop,
                                 (cc[:plain_sample], ir[:plain_sample], rot[:plain_sample]), (pir, prot),
                                 ids[:plain_sample], other_precs=(("fp16", PFTCC_PREC_FP16),
                                                                  ("split_fp16", PFTCC_PREC_SPLIT_FP16)))
    for nm in others:
        others[nm]["contraction_error_vs_fp64"] = cpu["error_vs_fp64"].pop(f"contraction_{nm}")
    lib.kill()
    zg_ms, zg_n = ch["zgemm"]
    zi_ms, zi_n = ch["zifft"]
    pairs = nptcls * nrefs
    zg_tflops = mfma_flops_per_pair(pftsz, kcorr) * pairs / (zg_ms / steps * 1e-3) / 1e12
    zbytes = pairs * (pftsz + 1) * 8.0
    return {
        "workload": "split-BF16 MFMA exhaustive search (no shifts), 3D-refinement shape, top-4 references "
                    "per particle rescored in FP64 with the exact per-pixel |CTF|",
        "value": value, "unit": "evals/s", "ms_per_step": 1e3 * wall, "steps": steps,
        "refs": nrefs, "ptcls": nptcls, "box": box, "smpd": smpd, "kfromto": list(kfromto), "pftsz": pftsz,
        "nrots": nrots, "ctf": "astigmatic <= 0.1 um, defocus U(0.8,3.5) um", "snr": 0.03,
        "ref_directions": "even spiral on the sphere (~3 deg)",
        "plain_fp32_same_shape": {"value": plain_value, "unit": "evals/s", "sample_ptcls": plain_sample},
        "speedup_vs_plain_fp32": value / plain_value,
        "assignment_agreement_vs_plain_fp32": float(np.mean(same)),
        "max_abs_best_cc_diff_vs_plain_fp32": float(np.max(np.abs(pcc - cc[:plain_sample]))),
        "unrescored": {"assignment_agreement_vs_plain_fp32": float(np.mean(same_raw)),
                       "max_abs_best_cc_diff_vs_plain_fp32": float(np.max(np.abs(pcc - raw_cc)))},
        "kernels": {
            "k_zgemm_as<3,2,5>": {"ms_per_step": zg_ms / steps, "launches_per_step": zg_n / steps,
                           "achieved": zg_tflops, "unit": "TFLOP/s (algorithmic complex contraction)",
                           "peak": BF16_PEAK_TFLOPS, "frac": zg_tflops / BF16_PEAK_TFLOPS,
                           "mfma_issued_frac": 3.0 * zg_tflops / BF16_PEAK_TFLOPS},
            "k_zifft_v2<400,4,256>": {"ms_per_step": zi_ms / steps, "launches_per_step": zi_n / steps,
                                "achieved": zbytes / (zi_ms / steps * 1e-3) / 1e9,
                                "unit": "GB/s (spectra read)", "peak": HBM_PEAK_GBS,
                                "frac": zbytes / (zi_ms / steps * 1e-3) / 1e9 / HBM_PEAK_GBS},
            "pack_topk_rescore": {"ms_per_step": ch["pack_rescore"][0] / steps},
        },
        "fp16_operands": {**others, "note": "FP16 (one product) and split FP16 (hi+lo, three products) operands "
                          "with per-shell and global power-of-two scales, FP32 accumulation, the same top-4 FP64 "
                          "rescoring; contraction error against the FP64 ring-mean formula on the checker's particles"},
        "cpu_baseline": cpu,
    }


def precision_cpu_baseline(lib, nrefs, kcorr, prec, ntop, best, plain, sample_ids, nfp64=6, other_precs=()):
    """CPU baseline of the precision form, used as its FP64 checker: the reference formula
    (gencorrs_cc_1, FFT route, FP64 numpy restatement of oracle/pftcc_oracle.py) of nfp64 particles
    against all references, timed on one host core, and the errors of the GPU's results against it.

    * contraction: the precision form's correlations (no rescoring) against the FP64 formula with the
      particle's ring-mean |CTF| (the quantity the per-m contraction computes): its arithmetic error;
    * ring_mean_ctf: the FP64 formula with the ring-mean |CTF| against the exact per-pixel |CTF|: the
      model error that the rescoring removes;
    * unrescored_total: the precision form's correlations against the exact FP64 formula;
    * assignments: best (reference, rotation) of the rescored and unrescored searches against the
      exact FP64 best, counted where the FP64 top-2 gap exceeds 1e-5;
    * on the plain-FP32 sample: every particle where the rescored precision form and the plain FP32
      search disagree is rescored at both answers with the exact FP64 formula;
    * other_precs (name, prec): the contraction error of those operand formats on the same particles.
    Errors are absolute, in units of cc (|cc| <= 1), and per vector relative to its FP64 peak."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pftcc_oracle as O

    refs = np.stack([lib.get_ref_pft(r + 1, True) for r in range(nrefs)])
    ids = np.asarray(sample_ids[:nfp64])
    keys = ["contraction", "ring_mean_ctf", "unrescored_total"] + [f"contraction_{nm}" for nm, _ in other_precs]
    err = {k: [] for k in keys}
    rel = {k: [] for k in keys}
    agree_res = agree_raw = counted = 0
    cpu_s = 0.0
    dev_rs = lib.search_mfma(ids, prec=prec, ntop=ntop)
    dev_raw = lib.search_mfma(ids, prec=prec, ntop=0)
    for q, i in enumerate(ids):
        pt = lib.get_ptcl_pft(int(i))
        ctf = lib.get_ctfmat(int(i)).astype(np.float64)
        ctf_rm = np.repeat(ctf.mean(axis=1, keepdims=True), ctf.shape[1], axis=1)
        t0 = time.perf_counter()
        exact = O.gencorrs_refs_fft(refs, pt, ctf, kcorr)
        cpu_s += time.perf_counter() - t0
        ringm = O.gencorrs_refs_fft(refs, pt, ctf_rm, kcorr)
        dev = lib.gencorrs_mfma_batch([int(i)], prec=prec)[0].astype(np.float64)
        pairs = [("contraction", dev, ringm), ("ring_mean_ctf", ringm, exact), ("unrescored_total", dev, exact)]
        for nm, pr in other_precs:
            pairs.append((f"contraction_{nm}", lib.gencorrs_mfma_batch([int(i)], prec=pr)[0].astype(np.float64), ringm))
        for key, a, b in pairs:
            d = np.abs(a - b)
            err[key].append(d.ravel())
            rel[key].append(np.max(d, axis=1) / np.maximum(np.max(np.abs(b), axis=1), 1e-30))
        flat = np.sort(exact.ravel())
        if flat[-1] - flat[-2] > 1e-5:
            counted += 1
            r, t = np.unravel_index(int(np.argmax(exact)), exact.shape)
            agree_res += int((dev_rs[1][q], dev_rs[2][q]) == (r + 1, t + 1))
            agree_raw += int((dev_raw[1][q], dev_raw[2][q]) == (r + 1, t + 1))

    def summary(key):
        a = np.concatenate(err[key])
        v = np.concatenate(rel[key])
        return {"max_abs": float(a.max()), "p99_abs": float(np.quantile(a, 0.99)),
                "max_peak_rel": float(v.max()), "p99_peak_rel": float(np.quantile(v, 0.99))}

    # disagreements between the rescored precision form and the plain FP32 search, settled in FP64
    (bcc, bir, brot), (pir, prot) = best, plain
    worse = ties = 0
    nd = 0
    for q in np.where((bir != pir) | (brot != prot))[0]:
        nd += 1
        i = int(sample_ids[q])
        pt = lib.get_ptcl_pft(i).astype(np.complex128)
        ctf = lib.get_ctfmat(i).astype(np.float64)

        def cc64(iref, irot):
            qr = refs[iref - 1, :kcorr].astype(np.complex128) * ctf[:kcorr]
            c = O.corrk_for_rot(qr, pt[:kcorr], irot).sum()
            return c / math.sqrt(np.sum(np.abs(qr) ** 2) * np.sum(np.abs(pt[:kcorr]) ** 2))
        a, b = cc64(int(bir[q]), int(brot[q])), cc64(int(pir[q]), int(prot[q]))
        if abs(a - b) <= 1e-6:
            ties += 1
        elif a < b:
            worse += 1
    return {
        "value": len(ids) * nrefs * 2 * refs.shape[-1] / cpu_s, "unit": "evals/s", "cores": 1, "kind": "port",
        "sample": f"{len(ids)} ptcls x {nrefs} refs x {2 * refs.shape[-1]} rots: the reference formula in FP64 "
                  f"(numpy FFT route, one core), used as the checker of the precision form",
        "error_vs_fp64": {k: summary(k) for k in err},
        "assignments_vs_fp64_best": {"particles_with_gap_gt_1e-5": counted,
                                     "rescored_identical": agree_res, "unrescored_identical": agree_raw},
        "plain_fp32_sample_disagreements": {"count": nd, "fp64_ties_within_1e-6": ties,
                                            "precision_form_worse_in_fp64": worse,
                                            "identical_or_fp64_tie_frac": 1.0 - worse / max(1, len(pir))},
    }


def coarse_fine_form(nrefs=10000, nptcls=100000, sample=1000, tops=(16, 32, 64, 128, 256), fp64_refs=8):
    """FP8 e4m3 coarse pass with top-T FP64 rescoring at the large coarse-to-fine shape (1e4
    references x 1e5 resident particles, box 256, kfromto [2,64]: pftsz 400, K 63), next to split
    BF16 with top-4 rescoring on the same data.

    Recall is measured against the FP64-best orientation of a sample of particles: the plain FP32
    form searches every reference at every rotation (per-reference maxima, strategy_srch greedy
    2-D), the fp64_refs best references are rescored in FP64 (numpy, the reference's gencorrs_cc_1
    with the exact per-pixel |CTF|) at their best rotation and both neighbours, and the best of
    those is the FP64 best.  A search's answer counts as recalled when it is that orientation or its
    FP64 cc is within 1e-6 of it (an FP64 tie).  The smallest T of `tops` with recall >= 99.9 % is
    then timed over all particles (one call, particles resident in HBM, rescoring included), and so
    is split BF16 with top-4 rescoring."""
    import pftcc
    from pftcc import Params, PolarFTCorrCalc, PFTCC_PREC_FP8, PFTCC_PREC_SPLIT_BF16, SRCH_GREEDY2D
    from pftcc import synth
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pftcc_oracle as O

    box, smpd, kfromto = 256, 1.06, (2, 64)
    pftsz = pftcc.magic_pftsz(126)
    w = synth.Workload(box=box, smpd=smpd, kfromto=kfromto, pftsz=pftsz, nrefs=nrefs, nptcls=nptcls, snr=0.03,
                       seed=20251204, dfrange=(1.0, 3.0), astig=0.1, max_shift=0.0, uniform_dirs=True)
    synth.make_params(w)
    lib = PolarFTCorrCalc().new(nrefs, (1, nptcls), Params(boxmatch=box, kfromto=kfromto, pftsz=pftsz, ctf=True))
    t0 = time.perf_counter()
    synth.load_into(lib, w, chunk=10000)
    log(f"coarse-to-fine setup: {nptcls} particles x {nrefs} refs in {time.perf_counter() - t0:.1f} s")
    S = synth.synth_lib()
    nrots = 2 * pftsz
    kcorr = w.nk
    ids = np.arange(1, nptcls + 1, dtype=np.int32)
    sids = ids[:sample]
    # FP64 best of the sample
    t0 = time.perf_counter()
    g = lib.strategy_srch(SRCH_GREEDY2D, sids, ref_corrs=True)
    plain_s = time.perf_counter() - t0
    refs = {}

    def ref_ring(r):
        if r not in refs:
            refs[r] = lib.get_ref_pft(r, True).astype(np.complex128)
        return refs[r]

    ptc = {}

    def cc64(i, iref, irot):
        if i not in ptc:
            pt = lib.get_ptcl_pft(int(i)).astype(np.complex128)[:kcorr]
            ctf = lib.get_ctfmat(int(i)).astype(np.float64)[:kcorr]
            ptc[i] = (pt, ctf, float(np.sum(np.abs(pt) ** 2)))
        pt, ctf, sp = ptc[i]
        q = ref_ring(iref)[:kcorr] * ctf
        return float(O.corrk_for_rot(q, pt, irot).sum() / math.sqrt(np.sum(np.abs(q) ** 2) * sp))

    t0 = time.perf_counter()
    best64 = []
    for q, i in enumerate(sids):
        rc, rr = g["ref_corrs"][q], g["ref_rots"][q]
        cands = np.argsort(-rc, kind="stable")[:fp64_refs]
        bv, bk = -np.inf, None
        for r0 in cands:
            for t in ((rr[r0] - 2) % nrots + 1, rr[r0], rr[r0] % nrots + 1):
                v = cc64(i, int(r0) + 1, int(t))
                if v > bv:
                    bv, bk = v, (int(r0) + 1, int(t))
        best64.append((bk, bv))
    fp64_s = time.perf_counter() - t0
    plain_recall = float(np.mean([(int(g["best_ref"][q]), int(g["best_rot"][q])) == best64[q][0]
                                  or cc64(sids[q], int(g["best_ref"][q]), int(g["best_rot"][q])) >= best64[q][1] - 1e-6
                                  for q in range(sample)]))

    def recall(ir, rot):
        hit = 0
        for q, i in enumerate(sids):
            (br, bt), bv = best64[q]
            if (int(ir[q]), int(rot[q])) == (br, bt) or (ir[q] > 0 and cc64(i, int(ir[q]), int(rot[q])) >= bv - 1e-6):
                hit += 1
        return hit / len(sids)

    lib.search_mfma(sids[:256], prec=PFTCC_PREC_FP8, ntop=16)          # warm-up (code objects)
    lib.search_mfma(sids[:256], prec=PFTCC_PREC_SPLIT_BF16, ntop=4)
    sweep = {}
    for T in tops:
        cc, ir, rot = lib.search_mfma(sids, prec=PFTCC_PREC_FP8, ntop=T)
        sweep[f"fp8_top{T}"] = recall(ir, rot)
    cc, ir, rot = lib.search_mfma(sids, prec=PFTCC_PREC_FP8, ntop=0)
    sweep["fp8_unrescored"] = recall(ir, rot)
    cc, ir, rot = lib.search_mfma(sids, prec=PFTCC_PREC_SPLIT_BF16, ntop=4)
    sweep["split_bf16_top4"] = recall(ir, rot)
    ok = [T for T in tops if sweep[f"fp8_top{T}"] >= 0.999]
    tsel = ok[0] if ok else tops[-1]
    log(f"coarse-to-fine: recall sweep {sweep}; timing FP8 top-{tsel} over {nptcls} particles")

    def timed(prec, ntop):
        S.synth_device_sync()
        lib.set_timing(True)
        t0 = time.perf_counter()
        cc, ir, rot = lib.search_mfma(ids, prec=prec, ntop=ntop)
        S.synth_device_sync()
        wall = time.perf_counter() - t0
        ch = {name: lib.get_timing_channel(c)[0] for c, name in ((1, "contraction"), (2, "inverse_fft"),
                                                                 (3, "select_rescore"))}
        lib.set_timing(False)
        return wall, ch, recall(ir[:sample], rot[:sample])

    wall8, ch8, rec8 = timed(PFTCC_PREC_FP8, tsel)
    wallb, chb, recb = timed(PFTCC_PREC_SPLIT_BF16, 4)
    lib.kill()
    evals = float(nptcls) * nrefs * nrots
    return {
        "workload": "FP8 e4m3 coarse pass (per-shell power-of-two scales, FP32 accumulation, spectra stored as FP16), the top-T references "
                    "of each particle rescored in FP64 at their coarse best rotation and both neighbours; "
                    "large coarse-to-fine shape, all particles resident in HBM, 256-particle operand chunks",
        "value": evals / wall8, "unit": "evals/s (end to end, one call over all particles)",
        "refs": nrefs, "ptcls": nptcls, "box": box, "smpd": smpd, "kfromto": list(kfromto), "pftsz": pftsz,
        "nrots": nrots, "ctf": "astigmatic <= 0.1 um, defocus U(1,3) um", "snr": 0.03,
        "ntop": tsel, "wall_s": wall8, "recall_vs_fp64_best": rec8, "kernel_ms": ch8,
        "hbm_resident_GB": nptcls * kcorr * (pftsz * 8 + pftsz * 4 + (pftsz + 1) * 8) / 1e9
                           + 2 * nrefs * kcorr * pftsz * 8 / 1e9,
        "recall_sweep": {"sample_ptcls": sample, "fp64_best": f"plain FP32 search of every reference, the "
                         f"{fp64_refs} best rescored in FP64 at their best rotation and both neighbours",
                         "plain_fp32": plain_recall, **sweep},
        "split_bf16_top4": {"value": evals / wallb, "wall_s": wallb, "recall_vs_fp64_best": recb, "kernel_ms": chb},
        "plain_fp32_sample": {"value": sample * nrefs * nrots / plain_s, "unit": "evals/s",
                              "note": "strategy_srch greedy 2-D with per-reference maxima, the FP64 checker's first stage"},
        "fp64_checker_s": fp64_s,
    }


def gradient_form(steps, nptcls=50000, nrefs=500, fd_sample=256, h=1e-3):
    """Batched adjoint (gencorr_grad_for_rot_8 + in-plane angle derivative) at the continuous-
    refinement shape: every particle paired with 1..8 references, shifts within +-4 px."""
    import pftcc
    from pftcc import Params, PolarFTCorrCalc
    from pftcc import synth

    box, smpd, kfromto = 200, 1.0, (2, 70)
    pftsz = pftcc.magic_pftsz(98)
    w = synth.Workload(box=box, smpd=smpd, kfromto=kfromto, pftsz=pftsz, nrefs=nrefs, nptcls=nptcls, snr=0.03,
                       seed=20251204, dfrange=(0.8, 3.5), astig=0.1, max_shift=4.0, uniform_dirs=True)
    synth.make_params(w)
    lib = PolarFTCorrCalc().new(nrefs, (1, nptcls), Params(boxmatch=box, kfromto=kfromto, pftsz=pftsz, ctf=True))
    synth.load_into(lib, w, chunk=10000)
    rng = np.random.default_rng(20251205)
    per = rng.integers(1, 9, nptcls)
    iptcls = np.repeat(np.arange(1, nptcls + 1, dtype=np.int32), per)
    irefs = rng.integers(1, nrefs + 1, iptcls.size).astype(np.int32)
    shv = rng.uniform(-4.0, 4.0, (iptcls.size, 2))
    irots = rng.integers(1, 2 * pftsz + 1, iptcls.size).astype(np.int32)
    npairs = iptcls.size
    S = synth.synth_lib()
    d_r = synth.DeviceBuffer.from_array(irefs)
    d_p = synth.DeviceBuffer.from_array(iptcls)
    d_s = synth.DeviceBuffer.from_array(shv)
    d_o = synth.DeviceBuffer.from_array(irots)
    d_f = synth.DeviceBuffer(8 * npairs)
    d_g = synth.DeviceBuffer(24 * npairs)
    lib.grad_batch_dev(npairs, d_r.ptr, d_p.ptr, d_s.ptr, d_o.ptr, d_f.ptr, d_g.ptr)     # warm-up
    S.synth_device_sync()
    lib.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        lib.grad_batch_dev(npairs, d_r.ptr, d_p.ptr, d_s.ptr, d_o.ptr, d_f.ptr, d_g.ptr)
    S.synth_device_sync()
    wall = (time.perf_counter() - t0) / steps
    k_ms, k_n = lib.get_timing_channel(4)
    lib.set_timing(False)
    g = d_g.to_array(np.float64, (npairs, 3))
    # central finite differences of the FP64 objective on a sample, h = 1e-3 px
    sel = rng.choice(npairs, fd_sample, replace=False)
    fd = np.zeros((fd_sample, 2))
    for d in range(2):
        e = np.zeros(2)
        e[d] = h
        fp = lib.grad_batch(irefs[sel], iptcls[sel], shv[sel] + e, irots[sel], with_grad=False)
        fm = lib.grad_batch(irefs[sel], iptcls[sel], shv[sel] - e, irots[sel], with_grad=False)
        fd[:, d] = (fp - fm) / (2 * h)
    scale = np.maximum(np.abs(fd), 1e-3 * np.max(np.abs(fd)))
    fd_err = np.abs(g[sel, :2] - fd) / scale
    for b in (d_r, d_p, d_s, d_o, d_f, d_g):
        b.close()
    lib.kill()
    K = w.nk
    flops_pair = 32.0 * K * pftsz
    avg_ms = k_ms / max(k_n, 1)
    achieved = flops_pair * npairs / (avg_ms * 1e-3) / 1e12
    return {
        "workload": "batched FP64 objective + adjoint (d/dx, d/dy, d/dpsi), continuous-refinement shape",
        "value": npairs / wall, "unit": "(f, grad f) per s", "ms_per_step": 1e3 * wall, "steps": steps,
        "ptcls": nptcls, "refs": nrefs, "pairs": int(npairs), "refs_per_ptcl": "U{1..8}", "box": box,
        "kfromto": list(kfromto), "pftsz": pftsz, "shift_range_px": 4.0,
        "fd_check": {"h_px": h, "sample_pairs": fd_sample, "max_rel_err": float(np.max(fd_err)),
                     "p99_rel_err": float(np.percentile(fd_err, 99))},
        "kernel": {"name": "k_rot8", "avg_launch_ms": avg_ms, "achieved": achieved, "unit": "TFLOP/s",
                   "flops_per_pair": flops_pair, "peak": FP64_PEAK_TFLOPS, "frac": achieved / FP64_PEAK_TFLOPS},
    }


def shift_search_form(nptcls=50000, nrefs=500, trs=5.0, cpu_sample=12):
    """Batched gradient shift search (pftcc_shsrch_grad%minimize with the in-plane callback) at the
    continuous-refinement shape: every particle against its source reference; the oracle's
    restatement (Python L-BFGS-B + numpy FP64 objective) on a few pairs is the CPU figure and the
    parity check at full size."""
    import pftcc
    from pftcc import Params, PolarFTCorrCalc
    from pftcc import synth

    box, smpd, kfromto = 200, 1.0, (2, 70)
    pftsz = pftcc.magic_pftsz(98)
    w = synth.Workload(box=box, smpd=smpd, kfromto=kfromto, pftsz=pftsz, nrefs=nrefs, nptcls=nptcls, snr=0.1,
                       seed=20251206, dfrange=(0.8, 3.5), astig=0.1, max_shift=4.0, uniform_dirs=True)
    synth.make_params(w)
    lib = PolarFTCorrCalc().new(nrefs, (1, nptcls), Params(boxmatch=box, kfromto=kfromto, pftsz=pftsz, ctf=True))
    synth.load_into(lib, w, chunk=10000)
    iptcls = np.arange(1, nptcls + 1, dtype=np.int32)
    irefs = (w.prm["ref"] + 1).astype(np.int32)
    S = synth.synth_lib()
    lib.shsrch_grad_batch(irefs[:64], iptcls[:64], trs=trs)          # warm-up (code objects)
    S.synth_device_sync()
    lib.set_timing(True)
    t0 = time.perf_counter()
    cxy, rot = lib.shsrch_grad_batch(irefs, iptcls, trs=trs)
    wall = time.perf_counter() - t0
    k_ms, k_n = lib.get_timing_channel(4)
    c_ms, c_n = lib.get_timing_channel(0)
    lib.set_timing(False)
    out = {
        "workload": "batched gradient shift search (L-BFGS-B over the shift, in-plane rotation callback), "
                    "every particle against its source reference, continuous-refinement shape",
        "value": nptcls / wall, "unit": "(reference, particle) searches per s", "wall_s": wall,
        "pairs": nptcls, "refs": nrefs, "box": box, "kfromto": list(kfromto), "pftsz": pftsz, "trs_px": trs,
        "snr": 0.1, "improved_frac": float(np.mean(rot > 0)),
        "objective_kernel": {"name": "k_rot8", "launches": int(k_n), "ms": k_ms},
        "callback_kernel": {"name": "k_corr_fused_items", "launches": int(c_n), "ms": c_ms},
        "state_machine_and_host_ms": 1e3 * wall - k_ms - c_ms,
    }
    # CPU: the oracle's restatement on a few of the same pairs (inputs read back from the library)
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import lbfgsb_oracle as LO
        import pftcc_oracle as O
        rng = np.random.default_rng(7)
        sel = rng.choice(nptcls, cpu_sample, replace=False)
        grid = O.make_grid(pftsz, kfromto[0], kfromto[1], box)
        refs = np.stack([lib.get_ref_pft(int(irefs[i]), True) for i in sel])
        pts = np.stack([lib.get_ptcl_pft(int(iptcls[i])) for i in sel])
        ctf = np.stack([lib.get_ctfmat(int(iptcls[i])) for i in sel])
        pf = O.Pftcc(grid=grid, kstop=kfromto[1], refs_even=refs, refs_odd=refs, ptcls=pts, ctfmats=ctf,
                     iseven=np.ones(cpu_sample, bool))
        t0 = time.perf_counter()
        agree, dsh = 0, 0.0
        for j, i in enumerate(sel):
            fdf = (lambda sh, r, j=j: pf.gencorr_grad_for_rot_8(j, j, sh, r))
            gen = (lambda sh, j=j: pf.gencorrs(j, j, tuple(float(v) for v in sh), method="fft"))
            ocxy, orot = LO.shsrch_minimize(fdf, gen, grid.angtab_deg, trs=trs, opt_angle=True)
            if orot == rot[i] and (orot == 0 or abs(float(ocxy[0]) - float(cxy[i, 0])) < 1e-5):
                agree += 1
            if orot and orot == rot[i]:
                dsh = max(dsh, float(np.max(np.abs(ocxy[1:] - cxy[i, 1:]))))
        cpu_wall = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cpu_sample / cpu_wall, "unit": "(reference, particle) searches per s",
                               "cores": 1, "kind": "port",
                               "sample": f"{cpu_sample} pairs of the same workload: the Python restatement "
                                         "(L-BFGS-B + numpy FP64 objective + FFT gencorrs callback)",
                               "agreement": agree / cpu_sample, "max_abs_shift_diff_px": dsh}
    except Exception as e:  # the GPU figure stands on its own
        out["cpu_baseline"] = {"error": repr(e)}
    lib.kill()
    return out


def polarizer_form(nimg=10000, box=200, reps=3):
    """polarize (Kaiser-Bessel gridding, 3 x 3 window) of device-resident Cartesian transforms into
    particle slots at the main workload's polar grid (box 200, kfromto [2,60], pftsz 308)."""
    import pftcc
    from pftcc import Params, PolarFTCorrCalc
    from pftcc import synth

    kfromto = (2, 60)
    pftsz = pftcc.magic_pftsz(98)
    lib = PolarFTCorrCalc().new(1, (1, nimg), Params(boxmatch=box, kfromto=kfromto, pftsz=pftsz, ctf=False))
    lib.init_polarizer()
    rng = np.random.default_rng(3)
    hc = box // 2 + 1
    chunk = 1000
    base = (rng.standard_normal((chunk, box, hc)) + 1j * rng.standard_normal((chunk, box, hc))).astype(np.complex64)
    d = synth.DeviceBuffer(nimg * box * hc * 8)
    S = synth.synth_lib()
    for c0 in range(0, nimg, chunk):
        S.synth_h2d(d.ptr + c0 * box * hc * 8, base.ctypes.data, base.nbytes)
    ids = np.arange(1, nimg + 1, dtype=np.int32)
    lib.polarize_dev(ids[:16], d.ptr, isptcl=True)                   # warm-up
    S.synth_device_sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        lib.polarize_dev(ids, d.ptr, isptcl=True)
    S.synth_device_sync()
    wall = (time.perf_counter() - t0) / reps
    d.close()
    lib.kill()
    # algorithmic bytes per image: the distinct half-plane components the windows touch (the
    # polar grid's coordinates k sin, -k cos, windows nint +- 1, Friedel-folded) + the polar output
    nk = kfromto[1] - kfromto[0] + 1
    th = np.arange(pftsz) * (np.pi / pftsz)
    ks = np.arange(kfromto[0], kfromto[1] + 1)[:, None]
    cx = np.rint(ks * np.sin(th)).astype(int)
    cy = np.rint(-ks * np.cos(th)).astype(int)
    pts = set()
    for dh in (-1, 0, 1):
        for dk in (-1, 0, 1):
            h = (cx + dh).ravel()
            k = (cy + dk).ravel()
            neg = h < 0
            pts.update(zip(np.where(neg, -h, h).tolist(), np.where(neg, -k, k).tolist()))
    bytes_img = 8.0 * len(pts) + 8.0 * nk * pftsz
    achieved = bytes_img * nimg / wall / 1e9
    return {
        "workload": f"polarize {nimg} device-resident {box}x{box} Fourier transforms into particle slots "
                    f"(kfromto {list(kfromto)}, pftsz {pftsz}, sqsums refreshed)",
        "value": nimg / wall, "unit": "images per s", "ms_per_call": 1e3 * wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "bytes_per_image": bytes_img,
                     "note": "whole call (k_polarize + k_sqsums) timed from the host"},
    }


def prep_form(nimg=10000, box=200, reps=3):
    """prepimg4align (noise_norm, fft, shift, CTF phase flip, ifft, clip, soft mask, fft) of
    device-resident 200 x 200 particle images, then polarize into particle slots (box 200,
    kfromto [2,60], pftsz 308)."""
    import pftcc
    from pftcc import Params, PolarFTCorrCalc
    from pftcc import synth

    kfromto = (2, 60)
    pftsz = pftcc.magic_pftsz(98)
    lib = PolarFTCorrCalc().new(1, (1, nimg), Params(boxmatch=box, kfromto=kfromto, pftsz=pftsz, ctf=False))
    lib.init_polarizer()
    rng = np.random.default_rng(4)
    chunk = 1000
    yy, xx = np.mgrid[0:box, 0:box].astype(np.float32)
    base = rng.standard_normal((chunk, box, box)).astype(np.float32)
    base += 3.0 * np.exp(-((xx - box / 2) ** 2 + (yy - box / 2) ** 2) / (2 * 20.0 ** 2)).astype(np.float32)
    d_img = synth.DeviceBuffer(nimg * box * box * 4)
    hc = box // 2 + 1
    d_cm = synth.DeviceBuffer(nimg * box * hc * 8)
    S = synth.synth_lib()
    for c0 in range(0, nimg, chunk):
        S.synth_h2d(d_img.ptr + c0 * box * box * 4, base.ctypes.data, base.nbytes)
    shifts = rng.uniform(-5.0, 5.0, (nimg, 2)).astype(np.float32)

    class _Ctf:
        smpd, kv, cs, fraca, angast, phshift, phaseplate = 1.0, 300.0, 2.7, 0.1, 30.0, 0.0, False

        def __init__(self, dfx, dfy):
            self.dfx, self.dfy = dfx, dfy

    df = rng.uniform(0.8, 3.5, nimg)
    ctfs = pftcc.ctf_params_array([_Ctf(float(d), float(d) + 0.05) for d in df])
    ids = np.arange(1, nimg + 1, dtype=np.int32)
    msk = 80.0
    lib.prep_imgs_dev(16, d_img.ptr, shifts[:16], box, msk, d_cm.ptr, ctfflag=1, ctfparams=ctfs)   # warm-up
    S.synth_device_sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        lib.prep_imgs_dev(nimg, d_img.ptr, shifts, box, msk, d_cm.ptr, ctfflag=1, ctfparams=ctfs)
    S.synth_device_sync()
    wall = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        lib.prep_imgs_dev(nimg, d_img.ptr, shifts, box, msk, d_cm.ptr, ctfflag=1, ctfparams=ctfs)
        lib.polarize_dev(ids, d_cm.ptr, isptcl=True)
    S.synth_device_sync()
    wall_chain = (time.perf_counter() - t0) / reps
    d_img.close()
    d_cm.close()
    lib.kill()
    bytes_img = 4.0 * box * box + 8.0 * box * hc        # the image in, its transform out
    achieved = bytes_img * nimg / wall / 1e9
    return {
        "workload": f"prepimg4align of {nimg} device-resident {box}x{box} particle images (noise_norm, fft, "
                    f"shift, CTF phase flip, ifft, clip, soft mask r {msk:.0f}, fft), then polarize",
        "value": nimg / wall, "unit": "images per s", "ms_per_call": 1e3 * wall,
        "with_polarize": {"value": nimg / wall_chain, "unit": "images per s", "ms_per_call": 1e3 * wall_chain},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "bytes_per_image": bytes_img,
                     "note": "algorithmic bytes (image in, transform out) over the whole call timed from the "
                             "host; the call streams the batch through HBM 8 times (8 kernels)"},
    }


def precision_same_workload(lib, w, shifts, plain_value, nptcls, sample=1000, ntop=16):
    """The split-BF16 MFMA form (with the 169-shift grid as extra operand rows and top-4 exact
    FP64 rescoring) on the main workload's references and particles, next to the plain form."""
    from pftcc import PFTCC_PREC_SPLIT_BF16
    from pftcc import synth

    S = synth.synth_lib()
    nptcls = min(nptcls, w.nptcls)
    ids = np.arange(1, nptcls + 1, dtype=np.int32)
    nrots = 2 * w.pftsz
    lib.search_mfma(ids[:64], prec=PFTCC_PREC_SPLIT_BF16, ntop=ntop, shifts=shifts)   # warm-up
    # rebuild the reference operand inside the timed region (as for every new reference set)
    refs_e = [lib.get_ref_pft(r + 1, True) for r in range(w.nrefs)]
    lib.set_ref_pfts(1, np.stack(refs_e), True)
    S.synth_device_sync()
    lib.set_timing(True)
    t0 = time.perf_counter()
    cc, ir, ish, rot = lib.search_mfma(ids, prec=PFTCC_PREC_SPLIT_BF16, ntop=ntop, shifts=shifts)
    S.synth_device_sync()
    wall = time.perf_counter() - t0
    ch = {name: lib.get_timing_channel(c) for c, name in ((1, "zgemm"), (2, "zifft"), (3, "pack_rescore"))}
    lib.set_timing(False)
    value = nptcls * w.nrefs * len(shifts) * nrots / wall
    pcc, pir, pish, prot = lib.search(ids[:sample], None, shifts)
    same = (pir == ir[:sample]) & (pish == ish[:sample]) & (prot == rot[:sample])
    return {
        "workload": f"split-BF16 MFMA exhaustive search with the +-3 px shift grid as operand rows, top-{ntop} "
                    "(reference, shift) rows rescored in FP64 with the exact per-pixel |CTF| and shift phases; "
                    "same references, particles and shifts as the plain line",
        "value": value, "unit": "evals/s", "ptcls": nptcls, "refs": w.nrefs, "shifts": len(shifts),
        "wall_s": wall, "speedup_vs_plain_fp32": value / plain_value,
        "assignment_agreement_vs_plain_fp32": float(np.mean(same)), "sample_ptcls": sample,
        "max_abs_best_cc_diff_vs_plain_fp32": float(np.max(np.abs(pcc - cc[:sample]))),
        "kernel_ms": {k: v[0] for k, v in ch.items()},
    }


def strategy_cpu_baseline(lib, w, nthreads):
    """The port (oracle/pftcc_cpu.cpp, OpenMP over particles) correlating 2 particles per thread
    with all references without shifts (gencorrs), timed on the host, then the greedy choice over
    references; checked against the device's greedy 3-D choice for the same particles."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cpu_port
    from pftcc import SRCH_GREEDY3D

    n, nk = w.pftsz, w.nk
    npt = 2 * nthreads
    pt = np.stack([lib.get_ptcl_pft(i + 1) for i in range(npt)])
    ctf = np.stack([lib.get_ctfmat(i + 1) for i in range(npt)])
    sq = np.array([lib.get_sqsum_ptcl(i + 1) for i in range(npt)], np.float32)
    pa, pb = cpu_port.memoize(pt, nthreads)
    refs = np.stack([lib.get_ref_pft(r + 1, True) for r in range(w.nrefs)])
    t0 = time.perf_counter()
    ccc, cir, _, crot = cpu_port.search_fp32(refs, ctf, sq, pa, pb, None, nk, nthreads)
    dt = time.perf_counter() - t0
    g = lib.strategy_srch(SRCH_GREEDY3D, np.arange(1, npt + 1, dtype=np.int32), ntop=1)
    same = (g["best_ref"] == cir) & (g["best_rot"] == crot)
    return {"value": npt * w.nrefs * 2 * n / dt, "unit": "evals/s", "cores": nthreads, "kind": "port",
            "sample": f"{npt} ptcls x {w.nrefs} refs x {2 * n} rots, no shift ({dt:.1f} s): the reference's FP32 "
                      "FFT route, OpenMP over particles, then the greedy choice",
            "gpu_assignment_agreement": float(np.mean(same)),
            "max_abs_best_cc_diff": float(np.max(np.abs(g["best_corr"] - ccc)))}


def strategy_form(lib, w, nptcls=10000, reps=2, with_cpu=True):
    """The reference-loop search strategies (pftcc_strategy_srch) over the main line's references
    and particles: every particle against all references at shift 0 (gencorrs), then the greedy /
    snhc 2-D and greedy / shc / snhc 3-D choices.  The stochastic kinds get a random search order
    with the previous class last (put_last) and prev_corr at 0.9 of the greedy best; snhc 3-D
    searches the first fifth of the order."""
    from pftcc import SRCH_GREEDY2D, SRCH_GREEDY3D, SRCH_SHC3D, SRCH_SNHC2D, SRCH_SNHC3D

    n = min(nptcls, w.nptcls)
    ids = np.arange(1, n + 1, dtype=np.int32)
    lib.strategy_srch(SRCH_GREEDY2D, ids[:64])                 # warm-up
    base = lib.strategy_srch(SRCH_GREEDY2D, ids)
    rng = np.random.default_rng(7)
    order = np.empty((n, w.nrefs), np.int32)
    for p in range(n):
        o = rng.permutation(w.nrefs) + 1
        o = np.concatenate([o[o != base["best_ref"][p]], [base["best_ref"][p]]])
        order[p] = o
    prev = (0.9 * base["best_corr"]).astype(np.float32)
    rnd = rng.random(n).astype(np.float32)
    kinds = {"greedy2d": dict(kind=SRCH_GREEDY2D),
             "snhc2d": dict(kind=SRCH_SNHC2D, srch_order=order, prev_corr=prev, rnd=rnd, stoch_bound=0.0),
             "greedy3d": dict(kind=SRCH_GREEDY3D, ntop=10),
             "shc3d": dict(kind=SRCH_SHC3D, srch_order=order, prev_corr=prev, ntop=10),
             "snhc3d": dict(kind=SRCH_SNHC3D, srch_order=order, ntop=10, szsn=w.nrefs // 5)}
    nrots = 2 * w.pftsz
    res = {}
    for name, kw in kinds.items():
        lib.set_timing(True)
        t0 = time.perf_counter()
        args = {k: v for k, v in kw.items() if k != "kind"}
        for _ in range(reps):
            out = lib.strategy_srch(kw["kind"], ids, **args)
        wall = (time.perf_counter() - t0) / reps
        kern_ms, launches = lib.get_timing()
        lib.set_timing(False)
        if name == "greedy3d":
            _, ir0, _, rot0 = lib.search(ids)
            same = bool(np.array_equal(out["best_ref"], ir0) and np.array_equal(out["best_rot"], rot0))
        # positions the device correlated: the greedy kinds all references; the stochastic kinds
        # whole rounds of 32, 64, 128, ... positions up to the one that ended each particle's scan
        if name in ("snhc2d", "shc3d"):
            ends = np.cumsum(32 * 2 ** np.arange(16))
            pos = np.minimum(ends[np.searchsorted(ends, out["nrefs_eval"])], w.nrefs)
        elif name == "snhc3d":
            pos = out["nrefs_eval"].astype(np.int64)
        else:
            pos = np.full(n, w.nrefs)
        res[name] = {"ms_per_call": 1e3 * wall, "evals_per_s": float(pos.sum()) * nrots / wall,
                     "full_scan_equivalent_evals_per_s": n * w.nrefs * nrots / wall,
                     "correlation_kernel_ms": kern_ms / reps, "mean_positions_correlated": float(pos.mean()),
                     "mean_nrefs_eval": float(np.mean(out["nrefs_eval"])),
                     "found_better_frac": float(np.mean(out["found_better"]))}
    cpu = None
    if with_cpu:
        try:
            cpu = strategy_cpu_baseline(lib, w, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        except Exception as e:  # the GPU figure stands on its own
            cpu = {"error": repr(e)}
    kern = res["greedy2d"]["correlation_kernel_ms"]
    achieved = plain_flops_per_pair(w.pftsz, w.nk) * n * w.nrefs / (kern * 1e-3) / 1e12
    return {"workload": f"search strategies: {n} particles x {w.nrefs} references, all rotations at shift 0 "
                        "(gencorrs), then the strategy's choice on the device",
            "value": res["greedy2d"]["evals_per_s"], "unit": "evals/s (greedy 2-D, whole call)",
            "note": "snhc / shc scan their search order in rounds of 32, 64, 128, ... positions; evals_per_s counts "
                    "the evaluations made, full_scan_equivalent_evals_per_s the greedy scan's",
            "ptcls": n, "refs": w.nrefs, "nrots": nrots, "kinds": res,
            "greedy3d_best_equals_pftcc_search": same,
            "cpu_baseline": cpu,
            "roofline": {"bound": "valu", "pipe": "VALU fp32", "kernel": "k_corr_fused<308,9,256,3,false>",
                         "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": achieved / FP32_PEAK_TFLOPS,
                         "note": "unshifted search: one pair per reference, so the staged-reference tile does not apply"}}


def host_cpu():
    """CPU model and core counts of this host (the CPU baseline's hardware)."""
    model, phys = "unknown", set()
    try:
        cur = {}
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    if "physical id" in cur and "core id" in cur:
                        phys.add((cur["physical id"], cur["core id"]))
                    cur = {}
                    continue
                k, v = (x.strip() for x in line.split(":", 1))
                cur[k] = v
                if k == "model name":
                    model = v
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(), "physical_cores": len(phys) or None,
            "affinity_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def cpu_baseline(pftcc, w, shifts, nthreads, target_s=12.0):
    """Times the reference CPU algorithm (C++/OpenMP port, oracle/pftcc_cpu.cpp) on a bounded
    sample of the same workload and checks the GPU's assignments on that sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cpu_port
    import pftcc_oracle as O

    n, nk = w.pftsz, w.nk
    grid = O.make_grid(n, w.kfromto[0], w.kfromto[1], w.box)
    npt = 2 * nthreads                      # two particles per thread (static schedule)
    pt = np.stack([pftcc.get_ptcl_pft(i + 1) for i in range(npt)])
    ctf = np.stack([pftcc.get_ctfmat(i + 1) for i in range(npt)])
    sq = np.array([pftcc.get_sqsum_ptcl(i + 1) for i in range(npt)], np.float32)
    pa, pb = cpu_port.memoize(pt, nthreads)

    def run(nr, ns):
        sel = shifts[np.linspace(0, len(shifts) - 1, ns).astype(int)]
        shm = np.stack([O.shmat_8(grid, s) for s in sel]).astype(np.complex64)
        refs = np.stack([pftcc.get_ref_pft(r + 1, True) for r in range(nr)])
        t0 = time.perf_counter()
        res = cpu_port.search_fp32(refs, ctf, sq, pa, pb, shm, nk, nthreads)
        return time.perf_counter() - t0, sel, res

    nr, ns = 2, 2
    dt, _, _ = run(nr, ns)                   # calibration
    scale = target_s / max(dt, 1e-3)
    ns = int(min(len(shifts), max(2, ns * scale)))
    scale = target_s / max(dt * ns / 2, 1e-3)
    nr = int(min(w.nrefs, max(2, nr * scale)))
    dt, sel, (ccc, cir, cish, crot) = run(nr, ns)
    evals = npt * nr * ns * 2 * n
    gcc, gir, gish, grot = pftcc.search(np.arange(1, npt + 1), np.arange(1, nr + 1), sel)
    same = (gir == cir) & (gish == cish) & (grot == crot)
    return {
        "value": evals / dt, "unit": "evals/s", "cores": nthreads, "kind": "port",
        "sample": f"{npt} ptcls x {nr} refs x {ns} shifts x {2 * n} rots of the same workload "
                  f"({dt:.1f} s): the reference's FP32 FFT route (r2c_n + c2c_n + c2r_2n per shell), "
                  f"OpenMP over particles",
        "fft": "own mixed-radix Stockham FFT (FFTW is absent), eight shells per AVX2 vector, twiddles precomputed",
        "cpu": host_cpu(),
        "gflops_per_core": evals / dt / (2 * n) * plain_flops_per_pair(n, nk) / nthreads / 1e9,
        "gpu_assignment_agreement": float(np.mean(same)),
        "max_abs_best_cc_diff": float(np.max(np.abs(gcc - ccc))),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=1000)
    ap.add_argument("--nptcls", type=int, default=100000)
    ap.add_argument("--nrefs", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-precision", action="store_true", help="skip the MFMA precision-form line")
    ap.add_argument("--precision-steps", type=int, default=2)
    ap.add_argument("--mfma-ptcls", type=int, default=10000,
                    help="particles of the precision-form run on the main workload")
    ap.add_argument("--no-gradient", action="store_true", help="skip the adjoint-kernel line")
    ap.add_argument("--legs", default="plain,precision,coarse,gradient,shsrch,polarizer,prep,strategy",
                    help="comma list of plain, precision, coarse, gradient, shsrch, polarizer, prep, strategy "
                         "(profiling runs select one)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a 1 GiB device copy (k_copy), the reference for HBM traffic counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the search shards particles "
              "across GPUs by design but this build measures the single-GPU kernel", flush=True)
        return 2

    import pftcc
    from pftcc import Params, PolarFTCorrCalc
    from pftcc import synth

    legs = set(args.legs.split(","))
    if args.copy_calibration:
        S0 = synth.synth_lib()
        a, b = synth.DeviceBuffer(1 << 30), synth.DeviceBuffer(1 << 30)
        if S0.synth_copy(b.ptr, a.ptr, 1 << 30):
            raise RuntimeError("calibration copy failed")
        a.close()
        b.close()
        log("copy calibration: 1 GiB read + 1 GiB written by k_copy")
    if "plain" not in legs:
        out = {"metric": METRIC, "n_gpus": 1, "note": "plain leg skipped (--legs)"}
        if "precision" in legs:
            out["precision_form"] = precision_form(args.precision_steps)
        if "coarse" in legs:
            out["coarse_to_fine"] = coarse_fine_form()
        if "gradient" in legs:
            out["gradient_form"] = gradient_form(3)
        if "shsrch" in legs:
            out["shift_search"] = shift_search_form()
        if "polarizer" in legs:
            out["polarizer"] = polarizer_form()
        if "prep" in legs:
            out["prep"] = prep_form()
        print(json.dumps(out), flush=True)
        return 0

    box, smpd, kfromto = 200, 1.0, (2, 60)
    pftsz = pftcc.magic_pftsz(98)       # ring2 = round2even(box/2 - 2)
    w = synth.Workload(box=box, smpd=smpd, kfromto=kfromto, pftsz=pftsz, nrefs=args.nrefs,
                       nptcls=args.nptcls, snr=0.03, seed=20251202, dfrange=(0.8, 3.5), astig=0.1,
                       max_shift=3.0)
    synth.make_params(w)
    lib = PolarFTCorrCalc().new(w.nrefs, (1, w.nptcls), Params(boxmatch=box, kfromto=kfromto, pftsz=pftsz, ctf=True))
    t0 = time.perf_counter()
    synth.load_into(lib, w, chunk=10000)
    log(f"setup: {w.nptcls} particles x {w.nrefs} refs generated, CTF + memo in {time.perf_counter() - t0:.1f} s")

    g = np.arange(-3.0, 3.0 + 1e-6, 0.5)
    shifts = np.array([(x, y) for x in g for y in g], np.float32)
    ns = len(shifts)
    nrots = 2 * pftsz
    S = synth.synth_lib()
    B = min(args.batch, w.nptcls)
    d_ids = synth.DeviceBuffer.from_array(np.arange(1, w.nptcls + 1, dtype=np.int32))
    d_sh = synth.DeviceBuffer.from_array(shifts)
    d_cc = synth.DeviceBuffer(4 * B)
    d_ir = synth.DeviceBuffer(4 * B)
    d_is = synth.DeviceBuffer(4 * B)
    d_rt = synth.DeviceBuffer(4 * B)

    def step(i):
        first = (i * B) % max(1, w.nptcls - B + 1)
        lib.search_dev(B, d_ids.ptr + 4 * first, w.nrefs, None, ns, d_sh.ptr, d_cc.ptr, d_ir.ptr, d_is.ptr, d_rt.ptr)

    for i in range(args.warmup):
        step(i)
    S.synth_device_sync()
    lib.set_timing(True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(args.warmup + i)
    S.synth_device_sync()
    wall = time.perf_counter() - t0
    kern_ms, launches = lib.get_timing()
    lib.set_timing(False)

    evals_step = B * w.nrefs * ns * nrots
    value = evals_step * args.steps / wall
    pairs_launch = B * w.nrefs * ns
    flops_launch = plain_flops_per_pair(pftsz, w.nk) * pairs_launch
    avg_ms = kern_ms / max(launches, 1)
    achieved = flops_launch / (avg_ms * 1e-3) / 1e12
    # compulsory HBM bytes per launch: refs once (8 B per ring sample) + particle rings, CTF and
    # memoised spectra of the batch + the per-particle outputs
    bytes_launch = w.nrefs * w.nk * pftsz * 8 + B * w.nk * (pftsz * 8 + pftsz * 4 + (pftsz + 1) * 8) + B * 16
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain FP32 exhaustive search, 2D-classification shape",
                   "refs": w.nrefs, "ptcls_resident": w.nptcls, "ptcls_per_step": B, "box": box,
                   "smpd": smpd, "kfromto": list(kfromto), "pftsz": pftsz, "nrots": nrots,
                   "shifts": ns, "shift_grid": "+-3 px step 0.5", "ctf": "astigmatic, defocus U(0.8,3.5) um",
                   "snr": 0.03, "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "pipe": "VALU fp32 (no MFMA; packed-fp32 disabled, issue peak 78.6)",
                     "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": PLAIN_TRAFFIC_BYTES_B1000 if (B == 1000 and w.nrefs == 200) else None,
                     "traffic_note": "HBM bytes per launch from rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated; "
                                     "compulsory bytes per launch are in hbm_compulsory_bytes",
                     "hbm_compulsory_bytes": bytes_launch,
                     "kernel": "k_corr_fused<308,9,256,3,true>", "avg_launch_ms": avg_ms, "launches": launches,
                     "flops_per_pair": plain_flops_per_pair(pftsz, w.nk), "pairs_per_launch": pairs_launch,
                     "hbm_compulsory_GBs": bytes_launch / (avg_ms * 1e-3) / 1e9,
                     "hbm_frac": bytes_launch / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }
    if not args.no_cpu:
        nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        try:
            out["cpu_baseline"] = cpu_baseline(lib, w, shifts, nthreads)
        except Exception as e:  # the GPU figure stands on its own
            out["cpu_baseline"] = {"error": repr(e)}
    if not args.no_precision and "precision" in legs:
        out["precision_form_same_workload"] = precision_same_workload(lib, w, shifts, value, args.mfma_ptcls)
    if "strategy" in legs:
        out["search_strategies"] = strategy_form(lib, w, with_cpu=not args.no_cpu)
    lib.kill()
    if not args.no_precision and "precision" in legs:
        out["precision_form"] = precision_form(args.precision_steps)
    if not args.no_precision and "coarse" in legs:
        out["coarse_to_fine"] = coarse_fine_form()
    if not args.no_gradient and "gradient" in legs:
        out["gradient_form"] = gradient_form(3)
    if "shsrch" in legs:
        out["shift_search"] = shift_search_form()
    if "polarizer" in legs:
        out["polarizer"] = polarizer_form()
    if "prep" in legs:
        out["prep"] = prep_form()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
