"""Benchmark: four-phonon processes/s on one MI355X (dense form vs FP64 form).

Workload: the synthetic diamond crystal (a = 5.43 A, m = 28.085 amu, spring
FC2, seeded second-shell fourth-order IFCs) on a 16x16x16 Gamma-centred mesh
(4096 q, 6 branches, 145 irreducible q, 870 modes), RTA at 300 K,
scalebroad 1.  A step is one RTA_driver_4ph pass over all 870 modes
(all three channels) in the dense MFMA form (FP16 with hi/lo phase factors
against one FP16 A' operand by default; BF16 hi/lo, FP16 hi/lo or FP32 with
--precision); value = processes/s of that pass, with every input resident in
HBM before the timed region.

Also reported: the FP64 per-process form (the reference-faithful kernel) and
the dense form on the same stratified sample (every 29th mode plus the modes
with the fewest and most processes, second-phonon window q' = 1..64), the
per-mode rate and RTA kappa errors of the FP32 form against the dense FP64
form on the full workload, the dominant kernel's roofline numbers (HIP events
on the library's stream), and the CPU oracle timed on all host cores on a
bounded piece of the same stratified sample, with the GPU forms timed on that
same piece beside it.

Library calls are synchronous (they return after their stream has drained),
so a wall clock around the steps brackets exactly the device work.

--workload iterative20 instead runs the diamond 20x20x20 iterative workload
(the 3ph+4ph operator assembled on the device, iterative kappa at
100/300/700 K) and prints its own line;
--workload sample16 runs the sampling estimator of RTA_driver_4ph on the
16x16x16 crystal against the full rates; --workload wurtzite12 runs every
MFMA form on the 12x12x8 wurtzite against the dense FP64 form.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3     # MI355X dense FP32 matrix peak
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0    # dense BF16 matrix peak
FP16_PEAK_TFLOPS = 2500.0    # dense FP16 matrix peak
CPU_QP = 4                   # q' window of the CPU baseline piece (about 20 s on 16 cores)


# MFMA forms: (library precision name, matrix peak in TFLOP/s of the issued datatype, dtype label)
FORMS = {
    "fp16x2": ("PREC_FP16X2", FP16_PEAK_TFLOPS, "fp16x2 (fp16 MFMA, hi/lo phase factors against one fp16 A' operand, "
                                                "fp32 accumulate)"),
    "bf16x3": ("PREC_BF16X3", BF16_PEAK_TFLOPS, "bf16x3 (bf16 hi/lo MFMA, fp32 accumulate)"),
    "fp16x3": ("PREC_FP16X3", FP16_PEAK_TFLOPS, "fp16x3 (fp16 hi/lo MFMA, fp32 accumulate)"),
    "fp32": ("PREC_FP32", FP32_PEAK_TFLOPS, "fp32"),
}


def prec_of(rt, name):
    return getattr(rt, FORMS[name][0])


def kernel_name(ctx, nb):
    """Name of the last dense launch's kernel, as in the kernel trace."""
    shape, kg = ctx.dense_shape(), ctx.dense_kernel()
    form = ["valu", "fp32", "bf16x3", "fp16x3", "fp16x2"][shape["mfma"]]
    if kg > 0:
        return f"k_dense2<{nb},{shape['ib']},{form},kg{kg}>"
    return f"k_dense<{nb},{shape['ib']},{shape['tq']},{form}>"


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def kappa_rta(case, rate):
    """RTA kappa (sum over branches, W/m/K) from per-mode rates, FP64 on the host
    (TConduct, conductivity.f90:11-37, with F = tau v omega)."""
    m = case.modes
    r = np.asarray(rate).reshape(m.nbands, m.Nlist, order="F")
    tau = np.where(r != 0, 1.0 / np.where(r != 0, r, 1.0), 0.0)
    # expand irreducible tau to the full grid through the stars
    tq = np.zeros((m.nptk, m.nbands))
    for ll in range(m.Nlist):
        for kk in range(m.Nequi[ll]):
            tq[m.ALLEquiList[kk, ll] - 1, :] = tau[:, ll]
    hbar, kb = 1.05457172647e-22, 1.380648813e-23
    T, V = case.T, case.structure.volume
    w = m.energy[1:]
    v = m.velocity[1:]
    f = 1.0 / np.expm1(hbar * w / kb / T)
    k = np.einsum("qb,qba,qbc->ac", f * (f + 1) * w * w * tq[1:], v, v)
    return k * float(np.float32(1e21)) * hbar ** 2 / (kb * T * T * V * m.nptk)


def dense_algorithmic_flops(case, ntriples):
    """Algorithmic work of the MFMA form: per visited momentum triple
    (channel, q, q', q'') with at least one surviving process, covering all
    Nb^4 branch combinations of all Nb modes at q, the fold of every IFC
    cluster with its phase (324 * Ntri4: one real x complex MAC per IFC
    element) plus the four-stage complex contraction (4 * Nb^5 complex MACs,
    8 flops each)."""
    nb = case.nbands
    return (324.0 * case.ifc4.Ntri + 32.0 * nb ** 5) * ntriples


def dense_flops(case, ngroups, nfine, ntriples, nproc, nitems):
    """Flops the regrouped dense kernel actually executes (Nb = number of
    branches): per triple the skinny GEMM, 8 flops per complex MAC over
    ngroups * Nb^2 * 9 entries; per process the contraction, 8 * (Nb^2 + Nb);
    per work item (channel, q, q') the A' build,
    5 * 81 * Ntri + nfine * 8 * 27 * Nb * (3 + Nb)."""
    nb = case.nbands
    per_triple = 8.0 * ngroups * nb * nb * 9
    per_proc = 8.0 * (nb * nb + nb)
    per_item = 5.0 * 81 * case.ifc4.Ntri + nfine * 8.0 * 27 * nb * (3 + nb)
    return per_triple * ntriples + per_proc * nproc + per_item * nitems


def measured_traffic(kernel):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summaries (profiles/*traffic*.md, calibrated
    against a copy kernel), or None when no summary covers this kernel."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*traffic*.md"))):
        for line in open(path):
            mt = re.match(r"traffic_bytes_per_launch:\s*(\d+)\s+kernel\s+(\S+)", line.strip())
            if mt and mt.group(2) == kernel:
                return int(mt.group(1))
    return None


def copy_calibration(nbytes):
    """One hipMemcpy device-to-device of nbytes, so that a counter run can
    scale FETCH_SIZE / WRITE_SIZE by a kernel of known traffic."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3):   # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"copy calibration: {nbytes} bytes device-to-device")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-form and accuracy legs")
    ap.add_argument("--no-autotune", action="store_true", help="use the default kernel shape instead of the autotuner's")
    ap.add_argument("--precision", choices=list(FORMS), default="fp16x2",
                    help="MFMA form of the timed steps: FP16 with hi/lo phases (fp16x2), BF16 or FP16 hi/lo "
                         "(3 products), or FP32 MFMA; FP32 accumulation throughout")
    ap.add_argument("--workload", choices=["rta16", "iterative20", "sample16", "wurtzite12", "batch24"], default="rta16",
                    help="rta16: the 4ph RTA throughput on diamond 16^3 (the metric); iterative20: diamond 20^3, "
                         "iterative kappa with the assembled 3ph+4ph operator at 100/300/700 K; sample16: the sampling estimator on "
                         "diamond 16^3 against the full rates; wurtzite12: every MFMA form on the 12x12x8 "
                         "wurtzite against dense FP64 (per-mode and kappa errors); batch24: the 24-crystal batch, "
                         "all 24 crystals (Nb 6 to 24) in one graph of one persistent launch")
    ap.add_argument("--phase-profile", action="store_true",
                    help="load the profiling build (libfp4prof.so) and print the dense kernel's per-phase cycle "
                         "shares of the timed steps")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="first run one device-to-device copy of this many bytes (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    if args.phase_profile:
        os.environ["FP4_PROFILE_BUILD"] = "1"   # read by fourphonon_amd.runtime at import
    if args.copy_bytes > 0:
        copy_calibration(args.copy_bytes)
    if args.workload == "iterative20":
        run_iterative20(args)
        return
    if args.workload == "sample16":
        run_sample16(args)
        return
    if args.workload == "wurtzite12":
        run_wurtzite12(args)
        return
    if args.workload == "batch24":
        run_batch24(args)
        return

    from fourphonon_amd import crystal
    from fourphonon_amd import runtime as rt

    t0 = time.time()
    case = crystal.config_case(2)
    m = case.modes
    log(f"case built in {time.time() - t0:.1f}s: nptk={m.nptk} Nb={m.nbands} Nlist={m.Nlist} Ntri={case.ifc4.Ntri}")
    ctx = rt.Context(case)
    nm = ctx.nmodes
    all_modes = np.arange(1, nm + 1, dtype=np.int32)

    # ---- counts (prefilter), reported separately
    t1 = time.perf_counter()
    Npp, _, Npm, _, Nmm, _ = ctx.NP_driver_4ph()
    t_pref = time.perf_counter() - t1
    nproc_total = int(Npp.sum() + Npm.sum() + Nmm.sum())
    log(f"prefilter: {nproc_total} processes in {t_pref:.2f}s")

    # ---- timed dense (MFMA) steps over the whole workload
    prec = prec_of(rt, args.precision)
    tuned = None
    if not args.no_autotune:
        # kernel shape from the autotuner, on every 29th mode over q' = 1..64 (untimed)
        t1 = time.perf_counter()
        tuned = ctx.autotune_dense(prec, all_modes[::29], ii_lo=1, ii_hi=64)
        tuned["seconds"] = time.perf_counter() - t1
        log(f"autotune: IB {tuned['ib']}, TQ {tuned['tq']}, {tuned['slots']} blocks ({tuned['seconds']:.1f}s)")
    for w in range(args.warmup):
        ctx.RTA_driver_4ph_modes(all_modes, prec)
        log(f"warmup {w + 1}/{args.warmup} done")
    ctx.reset_stats()
    t1 = time.perf_counter()
    for s in range(args.steps):
        r32, nproc = ctx.RTA_driver_4ph_modes(all_modes, prec)
        log(f"step {s + 1}/{args.steps} done")
    wall = time.perf_counter() - t1
    assert nproc == nproc_total, (nproc, nproc_total)
    phases = None
    if args.phase_profile:
        ph = ctx.dense_phase_profile()
        tot_c = sum(ph.values())
        phases = {k: round(v / tot_c, 4) for k, v in ph.items()}
    k_ms, k_n = ctx.kernel_stats(rt.Context.KS_DENSE)
    ms_step = 1e3 * wall / args.steps
    value = nproc_total * args.steps / wall

    # ---- roofline of the dense kernel
    info = ctx.dense_info()
    ntriples = info["triples"]
    flops = dense_algorithmic_flops(case, ntriples)
    executed = dense_flops(case, info["ngroups"], info["nfine"], ntriples, nproc_total, info["items"])
    avg_ms = k_ms / max(k_n, 1)
    achieved = flops / (avg_ms * 1e-3) / 1e12
    kname = kernel_name(ctx, m.nbands)
    peak = FORMS[args.precision][1]
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "traffic": measured_traffic(kname),
                "kernel": kname, "kernel_ms": round(avg_ms, 3), "launches": k_n,
                "flops_per_launch": flops, "flops_per_triple": flops / max(ntriples, 1),
                "executed_flops_per_launch": executed,
                "executed_tflops": round(executed / (avg_ms * 1e-3) / 1e12, 3)}

    out = {
        "metric": "4ph processes/s on 1 MI355X, MFMA form vs FP64 form, kappa within 1e-4 of FP64",
        "value": value, "unit": "processes/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": FORMS[args.precision][2], "data": "synthetic",
        "config": {"workload": "diamond 16x16x16, 2nd-shell 4th-order IFCs (Ntri=1522), 4ph RTA at 300 K, "
                               "all 870 modes", "ngrid": [16, 16, 16], "nbands": 6, "nlist": int(m.Nlist),
                   "ntri4": int(case.ifc4.Ntri), "processes": nproc_total, "triples": ntriples},
        "prefilter_seconds": t_pref,
        "autotune": tuned,
        "roofline": roofline,
    }
    if phases is not None:
        out["phase_shares_of_wave_cycles"] = phases

    # ---- FP64 form and dense form on the same stratified sample
    if not args.no_fp64:
        tot = Npp + Npm + Nmm
        act = np.nonzero(tot > 0)[0] + 1
        sample = np.unique(np.concatenate([act[::29], [act[np.argmin(tot[act - 1])], act[np.argmax(tot[act - 1])]]]))
        lo, hi = 1, 64
        ctx.reset_stats()
        t1 = time.perf_counter()
        r64s, n64s = ctx.RTA_driver_4ph_modes(sample, rt.PREC_FP64, ii_lo=lo, ii_hi=hi)
        t64 = time.perf_counter() - t1
        ms_r64, _ = ctx.kernel_stats(rt.Context.KS_RATE64)
        ms_emit, _ = ctx.kernel_stats(rt.Context.KS_EMIT)
        ms_cnt, _ = ctx.kernel_stats(rt.Context.KS_PREFILTER)
        t1 = time.perf_counter()
        r32s, n32s = ctx.RTA_driver_4ph_modes(sample, prec, ii_lo=lo, ii_hi=hi)
        t32 = time.perf_counter() - t1
        assert n32s == n64s
        fp64_pp = n64s / t64
        fp64_flops = 1822.0 * case.ifc4.Ntri * n64s
        out["fp64_form"] = {"value": fp64_pp, "unit": "processes/s", "processes": n64s,
                            "sample": f"{len(sample)} modes (every 29th active mode + min/max count), q'={lo}..{hi}",
                            "rate_kernel_ms": ms_r64, "emit_ms": ms_emit, "count_ms": ms_cnt,
                            "rate_kernel_tflops": fp64_flops / (ms_r64 * 1e-3) / 1e12,
                            "rate_kernel_frac_fp64_peak": fp64_flops / (ms_r64 * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}
        out["dense_on_fp64_sample"] = {"value": n32s / t32, "unit": "processes/s"}
        out["speedup_vs_fp64_form"] = (n32s / t32) / fp64_pp
        e_s = np.abs(r32s["rate_scatt_4ph"].ravel(order="F")[sample - 1] - r64s["rate_scatt_4ph"].ravel(order="F")[sample - 1])
        out["sample_rate_max_rel_err"] = float(np.max(e_s / np.abs(r64s["rate_scatt_4ph"].ravel(order="F")[sample - 1])))
        log(f"FP64 form {fp64_pp:.3e} proc/s on sample; dense {n32s / t32:.3e}")

        # accuracy of the FP32 form on the whole workload, against dense FP64
        t1 = time.perf_counter()
        r64, _ = ctx.RTA_driver_4ph_modes(all_modes, rt.PREC_FP64_DENSE)
        out["dense_fp64_seconds"] = time.perf_counter() - t1
        a = r32["rate_scatt_4ph"].ravel(order="F")
        b = r64["rate_scatt_4ph"].ravel(order="F")
        nz = b != 0
        out["rate_max_rel_err"] = float(np.max(np.abs(a[nz] - b[nz]) / np.abs(b[nz])))
        out["rate_p99_rel_err"] = float(np.percentile(np.abs(a[nz] - b[nz]) / np.abs(b[nz]), 99))
        k32, k64 = kappa_rta(case, a), kappa_rta(case, b)
        out["kappa_rta_trace_fp64"] = float(np.trace(k64) / 3)
        out["kappa_rel_err"] = float(abs(np.trace(k32) - np.trace(k64)) / abs(np.trace(k64)))
        log(f"rate max rel err {out['rate_max_rel_err']:.2e}, kappa rel err {out['kappa_rel_err']:.2e}")

        # the other MFMA forms, one pass each over the whole workload
        out["other_mfma_forms"] = []
        for other in FORMS:
            if other == args.precision:
                continue
            oprec, opeak = prec_of(rt, other), FORMS[other][1]
            ctx.reset_stats()
            t1 = time.perf_counter()
            ro, no = ctx.RTA_driver_4ph_modes(all_modes, oprec)
            to = time.perf_counter() - t1
            o_ms, _ = ctx.kernel_stats(rt.Context.KS_DENSE)
            oach = dense_algorithmic_flops(case, ctx.dense_info()["triples"]) / (o_ms * 1e-3) / 1e12
            ao = ro["rate_scatt_4ph"].ravel(order="F")
            out["other_mfma_forms"].append({
                "precision": other, "kernel": kernel_name(ctx, m.nbands), "value": no / to, "unit": "processes/s",
                "kernel_ms": o_ms,
                "achieved_tflops": round(oach, 3), "peak": opeak, "frac": round(oach / opeak, 4),
                "rate_max_rel_err": float(np.max(np.abs(ao[nz] - b[nz]) / np.abs(b[nz]))),
                "kappa_rel_err": float(abs(np.trace(kappa_rta(case, ao)) - np.trace(k64)) / abs(np.trace(k64)))})
            log(f"{other} form {no / to:.3e} proc/s, rate max rel err {out['other_mfma_forms'][-1]['rate_max_rel_err']:.2e}")

    # ---- three-phonon companion on the same crystal (2nd-shell FC3): counts
    # and FP64 per-process rates, all modes; reported beside the headline
    if case.ifc3 is not None:
        ctx.reset_stats()
        t1 = time.perf_counter()
        Np3, _, Nm3, _ = ctx.NP_driver()
        t_np3 = time.perf_counter() - t1
        t1 = time.perf_counter()
        ctx.RTA_driver()
        t_r3 = time.perf_counter() - t1
        ms_r3, _ = ctx.kernel_stats(rt.Context.KS_RATE3)
        n3 = int(Np3.sum() + Nm3.sum())
        f3 = 442.0 * case.ifc3.Ntri * n3
        out["three_phonon"] = {"processes": n3, "ntri3": int(case.ifc3.Ntri), "np_seconds": t_np3,
                               "rta_seconds": t_r3, "value": n3 / t_r3, "unit": "processes/s",
                               "rate_kernel_ms": ms_r3,
                               "rate_kernel_frac_fp64_peak": f3 / (ms_r3 * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}
        log(f"3ph: {n3} processes, counts {t_np3:.2f}s, FP64 rates {t_r3:.2f}s")

    # ---- CPU baseline: the oracle on all host cores, on a bounded piece of
    # the stratified sample the FP64 form runs (its modes, q' = 1..CPU_QP),
    # split into (mode, channel, q') items over a thread pool (the oracle's C
    # calls release the GIL); the GPU forms are timed on the same piece
    if not args.no_cpu:
        try:
            from concurrent.futures import ThreadPoolExecutor
            from oracle.oracle import Oracle
            tot = Npp + Npm + Nmm
            act = np.nonzero(tot > 0)[0] + 1
            sample = np.unique(np.concatenate([act[::29], [act[np.argmin(tot[act - 1])],
                                                            act[np.argmax(tot[act - 1])]]]))
            cores = len(os.sched_getaffinity(0))
            qp = CPU_QP
            o = Oracle(case)
            items = [(ch, int(mm), q) for mm in sample for ch in range(3) for q in range(1, qp + 1)]
            t1 = time.perf_counter()
            with ThreadPoolExecutor(max_workers=cores) as pool:
                cpu_g = list(pool.map(lambda it: o.rta_mode(it[0], it[1], it[2], it[2])[0], items))
            tc = time.perf_counter() - t1
            ctx.reset_stats()
            t1 = time.perf_counter()
            g64, n_cs = ctx.RTA_driver_4ph_modes(sample, rt.PREC_FP64, ii_lo=1, ii_hi=qp)
            tg64 = time.perf_counter() - t1
            t1 = time.perf_counter()
            gd, _ = ctx.RTA_driver_4ph_modes(sample, prec, ii_lo=1, ii_hi=qp)
            tgd = time.perf_counter() - t1
            # the oracle's per-item sums, against the GPU FP64 form on the same piece
            cpu_rate = np.zeros(3 * len(sample))
            for (ch, mm, q), g in zip(items, cpu_g):
                cpu_rate[ch * len(sample) + int(np.searchsorted(sample, mm))] += g
            keys = ("rate_scatt_plusplus", "rate_scatt_plusminus", "rate_scatt_minusminus")
            gpu_rate = np.concatenate([g64[k].ravel(order="F")[sample - 1] for k in keys])
            nzc = gpu_rate != 0
            err = float(np.max(np.abs(cpu_rate[nzc] - gpu_rate[nzc]) / np.abs(gpu_rate[nzc])))
            out["cpu_baseline"] = {
                "value": n_cs / tc, "unit": "processes/s", "cores": cores, "kind": "port",
                "cpu_model": cpu_model(),
                "sample": f"FP64 C oracle, {len(sample)} modes (every 29th active mode + min/max count) x 3 "
                          f"channels, q'=1..{qp}: {len(items)} (mode, channel, q') items over {cores} threads, "
                          f"{n_cs} processes in {tc:.1f}s",
                "gpu_fp64_form_same_piece": {"value": n_cs / tg64, "seconds": tg64},
                "gpu_mfma_form_same_piece": {"value": n_cs / tgd, "seconds": tgd},
                "gpu_fp64_vs_cpu_rate_max_rel_err": err}
            log(f"CPU oracle {n_cs / tc:.3e} proc/s on {cores} cores; GPU FP64 form {n_cs / tg64:.3e}, "
                f"MFMA form {n_cs / tgd:.3e} on the same piece")
        except Exception as e:  # the baseline must not hide the GPU result
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    ctx.close()
    print(json.dumps(out), flush=True)


def run_sample16(args):
    """Diamond 16x16x16, 2nd-shell FC4 (the metric's crystal): the sampling
    estimator of RTA_driver_4ph (num_sample_process_4ph) at several sample
    counts against the full dense-form rates.  Prints one JSON line with the
    time, the sampled processes evaluated and the per-mode rate and RTA kappa
    errors per sample count."""
    from fourphonon_amd import crystal
    from fourphonon_amd import runtime as rt

    case = crystal.config_case(2)
    ctx = rt.Context(case)
    prec = prec_of(rt, args.precision)
    t1 = time.perf_counter()
    full = ctx.RTA_driver_4ph(prec)["rate_scatt_4ph"].ravel(order="F")
    t_full = time.perf_counter() - t1
    k_full = np.trace(kappa_rta(case, full)) / 3
    log(f"full dense rates in {t_full:.2f}s, RTA kappa {k_full:.4f}")
    ok = full > 0
    recs = []
    for ns in (10000, 100000, 1000000):
        t1 = time.perf_counter()
        r, _ = ctx.RTA_driver_4ph_sample(ns, 1)
        ts = time.perf_counter() - t1
        est = r["rate_scatt_4ph"].ravel(order="F")
        err = np.abs(est[ok] - full[ok]) / full[ok]
        k = np.trace(kappa_rta(case, est)) / 3
        ms, _ = ctx.kernel_stats(rt.Context.KS_SAMPLE)
        rec = {"nsample": ns, "seconds": ts, "rate_rel_err_median": float(np.median(err)),
               "rate_rel_err_p90": float(np.percentile(err, 90)), "kappa_rel_err": float(abs(k - k_full) / k_full)}
        recs.append(rec)
        log(f"nsample {ns}: {ts:.3f}s, per-mode rate error median {rec['rate_rel_err_median']:.2e} "
            f"p90 {rec['rate_rel_err_p90']:.2e}, kappa error {rec['kappa_rel_err']:.2e}")
    ctx.close()
    print(json.dumps({"workload": "diamond 16x16x16, 2nd-shell FC4, sampled 4ph RTA rates vs the full dense form",
                      "precision_full": args.precision, "full_seconds": t_full, "kappa_full": float(k_full),
                      "samples": recs}), flush=True)


def run_wurtzite12(args):
    """Precision form on the 4-atom wurtzite-like cell (a = 3.2 A, c = 5.2 A,
    masses 26.98/14.01 amu, 12 branches) at 12x12x8 with second-shell FC4,
    RTA at 300 K: every MFMA form over all modes against the dense FP64 form,
    which is itself checked against the oracle's FP64 rates (golden fixture
    cfg3.npz, its q' window) and against the per-process FP64 form on a
    stratified sample.  Prints one JSON line with per-form time, processes/s,
    per-mode max and p99 rate error and RTA kappa error."""
    from fourphonon_amd import crystal
    from fourphonon_amd import runtime as rt

    t0 = time.time()
    case = crystal.config_case(3)
    m = case.modes
    log(f"case built in {time.time() - t0:.1f}s: nptk={m.nptk} Nb={m.nbands} Nlist={m.Nlist} Ntri={case.ifc4.Ntri}")
    ctx = rt.Context(case)
    nm = ctx.nmodes
    all_modes = np.arange(1, nm + 1, dtype=np.int32)
    t1 = time.perf_counter()
    Npp, _, Npm, _, Nmm, _ = ctx.NP_driver_4ph()
    t_pref = time.perf_counter() - t1
    nproc_total = int(Npp.sum() + Npm.sum() + Nmm.sum())
    log(f"prefilter: {nproc_total} processes in {t_pref:.2f}s")
    keys = ("rate_scatt_plusplus", "rate_scatt_plusminus", "rate_scatt_minusminus")

    # dense FP64 over the whole mesh: the accuracy reference of the MFMA forms
    t1 = time.perf_counter()
    r64, n64 = ctx.RTA_driver_4ph_modes(all_modes, rt.PREC_FP64_DENSE)
    t64 = time.perf_counter() - t1
    b = r64["rate_scatt_4ph"].ravel(order="F")
    nz = b != 0
    k64 = np.trace(kappa_rta(case, b)) / 3
    log(f"dense FP64: {t64:.1f}s, RTA kappa {k64:.4f}")

    # the dense FP64 form against the oracle (golden window) and the
    # per-process FP64 form (stratified sample, q' = 1..16)
    checks = {}
    gpath = os.path.join(ROOT, "tests", "golden", "cfg3.npz")
    if os.path.exists(gpath):
        g = np.load(gpath)
        rm, lo, hi = g["rate_modes"], int(g["ii_lo"]), int(g["ii_hi"])
        rw, _ = ctx.RTA_driver_4ph_modes(rm, rt.PREC_FP64_DENSE, ii_lo=lo, ii_hi=hi)
        errs = []
        for key, gk in zip(keys, ("rate_pp", "rate_pm", "rate_mm")):
            got, want = rw[key].ravel(order="F")[rm - 1], g[gk]
            ok = want != 0
            errs.append(float(np.max(np.abs(got[ok] - want[ok]) / np.abs(want[ok]), initial=0.0)))
        checks["dense_fp64_vs_oracle_golden"] = {"modes": [int(x) for x in rm], "q_window": [lo, hi],
                                                 "max_rel_err": max(errs)}
    tot = Npp + Npm + Nmm
    act = np.nonzero(tot > 0)[0] + 1
    sample = np.unique(np.concatenate([act[::97], [act[np.argmin(tot[act - 1])], act[np.argmax(tot[act - 1])]]]))
    rp, np_s = ctx.RTA_driver_4ph_modes(sample, rt.PREC_FP64, ii_lo=1, ii_hi=16)
    rd, _ = ctx.RTA_driver_4ph_modes(sample, rt.PREC_FP64_DENSE, ii_lo=1, ii_hi=16)
    a_, b_ = (np.concatenate([r[k].ravel(order="F")[sample - 1] for k in keys]) for r in (rd, rp))
    ok = b_ != 0
    checks["dense_fp64_vs_per_process_fp64"] = {
        "sample": f"{len(sample)} modes (every 97th active mode + min/max count), q'=1..16, {np_s} processes",
        "max_rel_err": float(np.max(np.abs(a_[ok] - b_[ok]) / np.abs(b_[ok])))}
    log(f"checks: {checks}")

    forms = []
    for name in FORMS:
        prec = prec_of(rt, name)
        ctx.reset_stats()
        t1 = time.perf_counter()
        ro, no = ctx.RTA_driver_4ph_modes(all_modes, prec)
        to = time.perf_counter() - t1
        k_ms, _ = ctx.kernel_stats(rt.Context.KS_DENSE)
        a = ro["rate_scatt_4ph"].ravel(order="F")
        err = np.abs(a[nz] - b[nz]) / np.abs(b[nz])
        kk = np.trace(kappa_rta(case, a)) / 3
        forms.append({"precision": name, "kernel": kernel_name(ctx, m.nbands), "seconds": to, "kernel_ms": k_ms,
                      "value": no / to, "unit": "processes/s", "rate_max_rel_err": float(err.max()),
                      "rate_p99_rel_err": float(np.percentile(err, 99)),
                      "rate_median_rel_err": float(np.median(err)),
                      "kappa_rel_err": float(abs(kk - k64) / abs(k64))})
        log(f"{name}: {to:.1f}s, {no / to:.3e} proc/s, rate err max {err.max():.2e} p99 "
            f"{np.percentile(err, 99):.2e}, kappa err {forms[-1]['kappa_rel_err']:.2e}")
    ctx.close()
    print(json.dumps({"workload": "wurtzite-like 4-atom cell 12x12x8, 2nd-shell FC4, 4ph RTA at 300 K, all "
                                  f"{nm} modes: MFMA forms vs dense FP64",
                      "processes": nproc_total, "prefilter_seconds": t_pref, "dense_fp64_seconds": t64,
                      "kappa_rta_fp64": float(k64), "checks": checks, "forms": forms}), flush=True)


def run_batch24(args):
    """The 24-crystal batch (fcc / zincblende / rocksalt-like cells of 2, 4,
    6 and 8 atoms, Nb 6 to 24, meshes 6^3 to 14^3 as drawn): the RTA pass of
    every crystal through fp4_batch_rta_4ph -- one persistent FP32-MFMA
    launch (k_dense_multi) over one queue of all 24 crystals' work items,
    costliest first by the prefilter counts, in a captured HIP graph.  The
    whole pass is about 6.8e12 processes, so a step is a sample: 4 of its 64
    q' windows spread over the q' range, one graph each, timed over --steps
    steps (a warm-up step runs each window's count pass and captures its
    graph); the batch's rates of the first window against each crystal's
    FP64 per-process rates there (three modes per crystal) give the per-mode
    error.  Prints one JSON line."""
    from fourphonon_amd import crystal
    from fourphonon_amd import runtime as rt

    t0 = time.time()
    cases = [crystal.batch_case(c) for c in range(crystal.BATCH_SIZE)]
    log(f"24 cases built in {time.time() - t0:.1f}s")
    ctxs = [rt.Context(cs) for cs in cases]
    t1 = time.perf_counter()
    b = rt.Batch(ctxs)
    t_create = time.perf_counter() - t1
    prec = rt.PREC_FP32
    # a step is a sample of the pass: 4 of its 64 q' windows, spread over the
    # q' range, one graph each (the whole pass is about 6.8e12 processes)
    nparts = 64
    parts = [k * nparts // 4 for k in range(4)]

    def step(tag):
        np_tot, ms_tot = 0, 0.0
        for part in parts:
            res, nproc, ms, nl = b.RTA_driver_4ph(prec, part=part, nparts=nparts)
            np_tot = np_tot + nproc
            ms_tot += ms
            log(f"{tag} q' window {part + 1}/{nparts}: {int(nproc.sum())} processes, graph {ms:.1f} ms, "
                f"{nl} kernels")
        return np_tot, ms_tot, nl

    for w in range(args.warmup):
        step(f"warmup {w + 1}/{args.warmup}")
    t1 = time.perf_counter()
    kms = []
    for st in range(args.steps):
        nproc, ms, nl = step(f"step {st + 1}/{args.steps}")
        kms.append(ms)
    wall = (time.perf_counter() - t1) / args.steps
    # the batch's rates of window 1 against each crystal's FP64 per-process
    # rates over the same q' window, three modes per crystal
    res0, _, _, _ = b.RTA_driver_4ph(prec, part=0, nparts=nparts)
    errs = []
    for k, c in enumerate(ctxs):
        tot = res0[k]["rate_scatt_4ph"].ravel(order="F")
        act = np.nonzero(tot > 0)[0] + 1
        modes = act[np.linspace(0, len(act) - 1, 3).astype(int)]
        hi = c.nptk // nparts
        r64, _ = c.RTA_driver_4ph_modes(modes, rt.PREC_FP64, ii_lo=1, ii_hi=hi)
        a_, b_ = tot[modes - 1], r64["rate_scatt_4ph"].ravel(order="F")[modes - 1]
        ok = b_ != 0
        errs.append(float(np.max(np.abs(a_[ok] - b_[ok]) / np.abs(b_[ok]), initial=0.0)))
        log(f"crystal {k} (Nb {cases[k].nbands}): batch vs FP64 per-process over q' = 1..{hi}: "
            f"max rel err {errs[-1]:.2e}")
    b.close()
    for c in ctxs:
        c.close()
    n_all = int(nproc.sum())
    by_nb = {}
    for k, cs in enumerate(cases):
        by_nb[cs.nbands] = by_nb.get(cs.nbands, 0) + int(nproc[k])
    out = {"workload": "24-crystal batch (2/4/6/8-atom fcc, zincblende, rocksalt-like cells, Nb 6..24, meshes "
                       "6^3..14^3), 4ph RTA at 300 K: all crystals in one graph of one persistent FP32-MFMA launch",
           "sample": f"a step is {len(parts)} of the pass's {nparts} q' windows (windows {[p + 1 for p in parts]})",
           "precision": "fp32", "steps": args.steps, "warmup": args.warmup,
           "crystals": [{"crystal": k, "nbands": cases[k].nbands, "mesh": list(map(int, cases[k].modes.ngrid)),
                         "processes": int(nproc[k])} for k in range(len(cases))],
           "processes_per_step_by_nbands": {str(k): v for k, v in sorted(by_nb.items())},
           "processes_per_step": n_all, "seconds_per_step": wall, "value": n_all / wall,
           "unit": "processes/s",
           "graph_ms_per_step": float(np.mean(kms)), "graphs_per_step": len(parts), "kernels_per_graph": nl,
           "batch_create_seconds": t_create,
           "sample_rate_max_rel_err_vs_fp64": max(errs)}
    print(json.dumps(out), flush=True)


def run_iterative20(args):
    """Diamond 20x20x20 (8000 q, 1536 irreducible modes) with 1st-shell
    FC3/FC4: the iterative BTE with the full 3ph+4ph scattering operator
    (four_phonon_iteration = .true.) at 100, 300 and 700 K.  Per T the
    operator is assembled on the device, aggregated by partner mode (4ph
    Gamma in the MFMA form, 3ph Gamma in FP64; the 4.3e11 4ph processes are
    never stored), tau0 comes from the same processes' rates, and the
    reference's in-place iteration runs to a relative kappa change below
    1e-5 with the FP64 kappa residual.  At 300 K the operator is assembled a
    second time with FP64 dense 4ph Gamma, and the iterations and kappa of
    the MFMA form are compared with it.  Prints one JSON line."""
    from fourphonon_amd import crystal
    from fourphonon_amd import runtime as rt

    t0 = time.time()
    case = crystal.config_case(5)
    m = case.modes
    log(f"case built in {time.time() - t0:.1f}s: nptk={m.nptk} Nlist={m.Nlist} Ntri4={case.ifc4.Ntri} "
        f"Ntri3={case.ifc3.Ntri}")
    ctx = rt.Context(case)
    prec = prec_of(rt, args.precision)

    def solve(precision, check_sweep=False):
        ctx.reset_stats()
        t1 = time.perf_counter()
        # assembled in 8 pieces of whole irreducible points, a progress line each
        nm = ctx.nmodes
        r3 = np.zeros((m.nbands, m.Nlist), order="F")
        r4 = np.zeros((m.nbands, m.Nlist), order="F")
        n3 = n4 = 0
        ctx.bte_operator_begin()
        cuts = np.linspace(0, m.Nlist, 9).astype(int)
        for p in range(8):
            modes = np.arange(cuts[p] * m.nbands + 1, cuts[p + 1] * m.nbands + 1, dtype=np.int32)
            if len(modes) == 0:
                continue
            a3, a4 = ctx.bte_operator_add(modes, precision, r3=r3, r4=r4)
            n3, n4 = n3 + a3, n4 + a4
            log(f"  operator piece {p + 1}/8: modes {modes[0]}..{modes[-1]} of {nm}, {time.perf_counter() - t1:.1f}s")
        ctx.bte_operator_end()
        ta = time.perf_counter() - t1
        rate = r3 + r4
        tau = np.asfortranarray(np.where(rate > 0, 1.0 / np.where(rate > 0, rate, 1.0), 0.0))
        t1 = time.perf_counter()
        res = ctx.bte_solve(tau, maxiter=1000)
        ts = time.perf_counter() - t1
        sweep_ms, steps = ctx.kernel_stats(rt.Context.KS_ITER)
        k = res["kappa_iter"]
        extra = {}
        if check_sweep:
            # the same solve with one launch per (ll, i, kk) step: bit-identical F_n and kappa
            ctx.reset_stats()
            ctx.bte_set_sweep(False)
            t1 = time.perf_counter()
            res1 = ctx.bte_solve(tau, maxiter=1000)
            ts1 = time.perf_counter() - t1
            ctx.bte_set_sweep(True)
            ms1, steps1 = ctx.kernel_stats(rt.Context.KS_ITER)
            same = bool(np.array_equal(res1["F_n"], res["F_n"]) and np.array_equal(res1["kappa_iter"], k))
            extra = {"one_launch_per_step": {"solve_seconds": ts1, "seconds_per_sweep": ms1 * 1e-3 / max(int(res1["niter"]), 1),
                                             "launches_per_sweep": steps1 // max(int(res1["niter"]), 1),
                                             "niter": int(res1["niter"]), "identical_to_persistent": same}}
            log(f"  one launch per step: {ts1:.2f}s, {ms1 * 1e-3 / max(int(res1['niter']), 1):.3f} s/sweep, "
                f"bit-identical to the persistent sweep: {same}")
        return {**extra, "processes_3ph": int(n3), "processes_4ph": int(n4), "assemble_seconds": ta,
                "solve_seconds": ts, "niter": int(res["niter"]), "final_relchange": float(res["relchange"][-1]),
                "seconds_per_sweep": sweep_ms * 1e-3 / max(int(res["niter"]), 1), "launches_per_sweep": steps // max(int(res["niter"]), 1),
                "kappa_rta": float(np.trace(k[:, :, 0]) / 3), "kappa": float(np.trace(k[:, :, -1]) / 3),
                "kappa_tensor": k[:, :, -1].tolist()}

    per_T = []
    for T in (100.0, 300.0, 700.0):
        ctx.set_temperature(T)
        rec = {"T": T, **solve(prec, check_sweep=(T == 300.0))}
        log(f"T={T:.0f} K: operator {rec['assemble_seconds']:.1f}s ({rec['processes_4ph']} 4ph + "
            f"{rec['processes_3ph']} 3ph processes), solve {rec['solve_seconds']:.2f}s in {rec['niter']} iterations "
            f"({rec['seconds_per_sweep']:.3f} s/sweep), kappa {rec['kappa']:.4f} W/m/K (RTA {rec['kappa_rta']:.4f})")
        if T == 300.0 and not args.no_fp64:
            ref = solve(rt.PREC_FP64_DENSE)
            rec["fp64_dense_reference"] = {k: ref[k] for k in ("assemble_seconds", "niter", "kappa", "kappa_rta")}
            rec["kappa_rel_err_vs_fp64_dense"] = abs(rec["kappa"] - ref["kappa"]) / abs(ref["kappa"])
            rec["niter_fp64_dense"] = ref["niter"]
            log(f"  FP64 dense operator: {ref['assemble_seconds']:.1f}s, {ref['niter']} iterations, kappa "
                f"{ref['kappa']:.4f}; MFMA-form kappa rel err {rec['kappa_rel_err_vs_fp64_dense']:.2e}")
        per_T.append(rec)
    ctx.close()
    total = sum(r["assemble_seconds"] + r["solve_seconds"] for r in per_T)
    print(json.dumps({"workload": "diamond 20x20x20, 1st-shell FC3/FC4, iterative BTE with the 3ph+4ph operator "
                                  "assembled on the device (aggregated by partner mode), eps 1e-5",
                      "unit": "W/(m K)", "precision_4ph": args.precision, "total_seconds": total,
                      "temperatures": per_T}), flush=True)


if __name__ == "__main__":
    main()
