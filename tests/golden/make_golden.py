"""Writes the golden fixtures of the GPU parity tests with the CPU oracle.

Run in the development container (the oracle never runs on the GPU box):

    python tests/golden/make_golden.py cfg1 cfg2 cfg3 cfg4

Each part writes tests/golden/<part>.npz holding inputs-by-seed (the crystal
is regenerated from its seed by fourphonon_amd.crystal) and the oracle's
outputs: per-mode counts and phase space, order-sensitive per-mode list
hashes, and FP64 rates for a subset of modes (and a q' window for the
large meshes).
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fourphonon_amd import crystal  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402
from tests.golden.listhash import list_hash  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
THREADS = int(os.environ.get("GOLDEN_THREADS", "8"))


def log(*a):
    print(time.strftime("%H:%M:%S"), *a, flush=True)


def counts_and_hashes(o, modes):
    """Counts, phase space and list hashes of the given modes."""
    from concurrent.futures import ThreadPoolExecutor
    nm = len(modes)
    N = np.zeros((3, nm), np.int64)
    P = np.zeros((3, nm))
    H = np.zeros((3, nm), np.uint64)

    def one(arg):
        t, mm = arg
        res = []
        for ch in range(3):
            n, p = o.np_mode(ch, int(mm))
            lst = o.ind_only_mode(ch, int(mm), cap=n)
            res.append((n, p, list_hash(*lst)))
        return t, res

    with ThreadPoolExecutor(THREADS) as ex:
        for t, res in ex.map(one, list(enumerate(modes))):
            for ch, (n, p, h) in enumerate(res):
                N[ch, t], P[ch, t], H[ch, t] = n, p, h
    return N, P, H


def part_cfg1():
    case = crystal.config_case(1)
    o = Oracle(case)
    allm = np.arange(1, o.nmodes + 1, dtype=np.int32)
    log("cfg1: counts + hashes for", len(allm), "modes")
    N, P, H = counts_and_hashes(o, allm)
    rmodes = np.unique(np.concatenate([allm[3::7], [allm[-1]]])).astype(np.int32)
    log("cfg1: rates for", len(rmodes), "modes")
    t0 = time.time()
    r = o.rta_driver(modes=rmodes, nthreads=THREADS)
    log("cfg1: rates took", time.time() - t0)
    np.savez_compressed(os.path.join(OUT, "cfg1.npz"), ngrid=np.array([8, 8, 8]), shell=1, workload=1,
                        modes=allm, N=N, P=P, H=H, rate_modes=rmodes, rate_seconds=time.time() - t0,
                        rate_threads=THREADS, **{k: v[rmodes - 1] for k, v in r.items()})


def part_cfg1k():
    """cfg1 as the workload defines it: 4ph and 3ph RTA rates of every mode
    (the RTA kappa follows from them)."""
    case = crystal.config_case(1)
    o = Oracle(case)
    allm = np.arange(1, o.nmodes + 1, dtype=np.int32)
    log("cfg1k: 3ph rates for all", len(allm), "modes")
    r3 = o.rta3_driver(modes=allm, nthreads=THREADS)
    log("cfg1k: 4ph rates for all", len(allm), "modes")
    t0 = time.time()
    r4 = o.rta_driver(modes=allm, nthreads=THREADS)
    log("cfg1k: 4ph rates took", time.time() - t0)
    np.savez_compressed(os.path.join(OUT, "cfg1k.npz"), ngrid=np.array([8, 8, 8]), shell=1, workload=1,
                        rate_seconds=time.time() - t0, rate_threads=THREADS,
                        **{"r4_" + k: v for k, v in r4.items()}, **{"r3_" + k: v for k, v in r3.items()})


def part_cfg2():
    case = crystal.config_case(2)
    o = Oracle(case)
    nm = o.nmodes
    modes = np.unique(np.array([4, 5, 61, 118, 200, 333, 415, 500, 611, 700, 777, 812, 850, nm - 1, nm], np.int32))
    log("cfg2: counts + hashes for", len(modes), "modes")
    N, P, H = counts_and_hashes(o, modes)
    rmodes = np.array([5, 200, 611, nm], np.int32)
    lo, hi = 1, 3
    log("cfg2: windowed rates for", len(rmodes), "modes")
    t0 = time.time()
    r = o.rta_driver(modes=rmodes, ii_lo=lo, ii_hi=hi, nthreads=THREADS)
    log("cfg2: rates took", time.time() - t0)
    np.savez_compressed(os.path.join(OUT, "cfg2.npz"), ngrid=np.array([16, 16, 16]), shell=2, workload=2,
                        modes=modes, N=N, P=P, H=H, rate_modes=rmodes, ii_lo=lo, ii_hi=hi,
                        rate_seconds=time.time() - t0, rate_threads=THREADS,
                        **{k: v[rmodes - 1] for k, v in r.items()})


def streamed_counts_and_hashes(o, modes):
    """Counts, phase space and list hashes of the given modes from the
    oracle's streamed hash (no list is stored), one mode per task."""
    from concurrent.futures import ThreadPoolExecutor
    nm = len(modes)
    N = np.zeros((3, nm), np.int64)
    P = np.zeros((3, nm))
    H = np.zeros((3, nm), np.uint64)
    done = [0]

    def one(arg):
        t, mm = arg
        return t, [o.list_hash_mode(ch, int(mm)) for ch in range(3)]

    with ThreadPoolExecutor(THREADS) as ex:
        for t, res in ex.map(one, list(enumerate(modes))):
            for ch, (n, p, h) in enumerate(res):
                N[ch, t], P[ch, t], H[ch, t] = n, p, h
            done[0] += 1
            if done[0] % 29 == 0:
                log(f"  {done[0]}/{nm} modes")
    return N, P, H


def part_cfg2lists():
    """cfg2 at full size: counts, phase space and list hashes of all 870
    modes in all three channels (the whole 8.1e10-process list, hashed)."""
    case = crystal.config_case(2)
    o = Oracle(case)
    allm = np.arange(1, o.nmodes + 1, dtype=np.int32)
    log("cfg2lists: counts + streamed hashes for", len(allm), "modes on", THREADS, "threads")
    t0 = time.time()
    N, P, H = streamed_counts_and_hashes(o, allm)
    log("cfg2lists: took", time.time() - t0)
    np.savez_compressed(os.path.join(OUT, "cfg2_lists.npz"), ngrid=np.array([16, 16, 16]), shell=2, workload=2,
                        modes=allm, N=N, P=P, H=H, seconds=time.time() - t0, threads=THREADS)


def part_cfg3lists():
    """cfg3 at full size: counts, phase space and list hashes of all 1140
    modes of the 12x12x8 wurtzite in all three channels (the whole
    6.2e10-process list, hashed)."""
    case = crystal.config_case(3)
    o = Oracle(case)
    allm = np.arange(1, o.nmodes + 1, dtype=np.int32)
    log("cfg3lists: counts + streamed hashes for", len(allm), "modes on", THREADS, "threads")
    t0 = time.time()
    N, P, H = streamed_counts_and_hashes(o, allm)
    log("cfg3lists: took", time.time() - t0)
    np.savez_compressed(os.path.join(OUT, "cfg3_lists.npz"), ngrid=np.array([12, 12, 8]), shell=2, workload=3,
                        modes=allm, N=N, P=P, H=H, seconds=time.time() - t0, threads=THREADS)


def part_iter6():
    """The joint 3ph+4ph iterative solution on the 6x6x6 diamond (1st-shell
    FC3/FC4, 300 K): the oracle's counts, tau0, iteration count, kappa per
    iteration and converged F_n from its own lists and Gamma (5.9e7 4ph
    processes; about ten minutes on eight cores)."""
    from oracle.oracle import bte_solve
    case = crystal.config_case(1, ngrid=(6, 6, 6))
    o = Oracle(case)
    t0 = time.time()
    No3, _ = o.np3_driver(nthreads=THREADS)
    wl3, wg3, wr3 = o.ind3_driver(No3, nthreads=THREADS)
    No4, _ = o.np_driver(nthreads=THREADS)
    log("iter6: 4ph lists and Gamma for", sum(int(x.sum()) for x in No4), "processes")
    wl4, wg4, wr4 = o.ind_driver(No4, nthreads=THREADS)
    rate = (wr3["rate"] + wr4["rate_4ph"]).reshape((case.modes.nbands, case.modes.Nlist), order="F")
    tau = np.asfortranarray(np.where(rate > 0, 1.0 / np.where(rate > 0, rate, 1.0), 0.0))
    log("iter6: iterating")
    r = bte_solve(case, No4, wl4, wg4, tau, maxiter=60, N3=No3, lists3=wl3, gammas3=wg3)
    log("iter6: niter", r["niter"], "took", time.time() - t0)
    np.savez_compressed(os.path.join(OUT, "iter6.npz"), ngrid=np.array([6, 6, 6]), shell=1, workload=1, T=300.0,
                        N3=np.array([No3[0], No3[1]]), N4=np.array([No4[0], No4[1], No4[2]]),
                        rate_3ph=wr3["rate"], rate_4ph=wr4["rate_4ph"], tau=tau, niter=r["niter"],
                        kappa_iter=r["kappa_iter"], kappa=r["kappa"], relchange=r["relchange"], F_n=r["F_n"],
                        seconds=time.time() - t0, threads=THREADS)


def part_cfg3():
    case = crystal.config_case(3)
    o = Oracle(case)
    nm = o.nmodes
    modes = np.unique(np.array([13, 14, 40, 100, 250, nm // 2, nm - 5, nm], np.int32))
    log("cfg3: counts + hashes for", len(modes), "modes, Nlist", o.nlist)
    N, P, H = counts_and_hashes(o, modes)
    rmodes = np.array([14, nm // 2], np.int32)
    lo, hi = 2, 3
    t0 = time.time()
    r = o.rta_driver(modes=rmodes, ii_lo=lo, ii_hi=hi, nthreads=THREADS)
    log("cfg3: rates took", time.time() - t0)
    np.savez_compressed(os.path.join(OUT, "cfg3.npz"), ngrid=np.array([12, 12, 8]), shell=2, workload=3,
                        modes=modes, N=N, P=P, H=H, rate_modes=rmodes, ii_lo=lo, ii_hi=hi,
                        **{k: v[rmodes - 1] for k, v in r.items()})


CFG4_HASH_CAP = 30_000_000     # lists longer than this per mode are counted, not hashed


def part_cfg4():
    """The 24-crystal batch at its drawn meshes: three modes per crystal for
    counts, phase space and list hashes, one mode per crystal for rates over a
    one-point q' window (ii = 2)."""
    from concurrent.futures import ThreadPoolExecutor
    cases = []
    for c in range(crystal.BATCH_SIZE):
        cases.append(crystal.batch_case(c))
        m = cases[-1].modes
        log(f"cfg4: crystal {c} {cases[-1].structure.name} mesh {list(m.ngrid)} Nb {m.nbands} Nlist {m.Nlist}")
    oracles = [Oracle(cs) for cs in cases]
    modes = []
    for o, cs in zip(oracles, cases):
        nb, nm = cs.nbands, o.nmodes
        modes.append(np.array([nb + 4, nm // 2 + 1, nm], np.int32))
    rmode = np.array([m[1] for m in modes], np.int32)

    def counts_job(arg):
        c, t = arg
        o, mm = oracles[c], int(modes[c][t])
        res = []
        for ch in range(3):
            n, p = o.np_mode(ch, mm)
            h = list_hash(*o.ind_only_mode(ch, mm, cap=n)) if n <= CFG4_HASH_CAP else 0
            res.append((n, p, h, n <= CFG4_HASH_CAP))
        return c, t, res

    def rate_job(c):
        r = oracles[c].rta_driver(modes=rmode[c:c + 1], ii_lo=2, ii_hi=2, nthreads=1)
        return c, {k: v[rmode[c] - 1] for k, v in r.items()}

    nc = crystal.BATCH_SIZE
    N = np.zeros((nc, 3, 3), np.int64)
    P = np.zeros((nc, 3, 3))
    H = np.zeros((nc, 3, 3), np.uint64)
    hashed = np.zeros((nc, 3, 3), bool)
    t0 = time.time()
    jobs = [(c, t) for c in range(nc) for t in range(3)]
    jobs.sort(key=lambda ct: -cases[ct[0]].nptk ** 2 * cases[ct[0]].nbands ** 3)
    with ThreadPoolExecutor(THREADS) as ex:
        for c, t, res in ex.map(counts_job, jobs):
            for ch, (n, p, h, ok) in enumerate(res):
                N[c, ch, t], P[c, ch, t], H[c, ch, t], hashed[c, ch, t] = n, p, h, ok
            log(f"cfg4: counts crystal {c} mode {modes[c][t]} N {[r[0] for r in res]}")
    log("cfg4: counts took", time.time() - t0)
    rates = {}
    t0 = time.time()
    with ThreadPoolExecutor(THREADS) as ex:
        for c, r in ex.map(rate_job, range(nc)):
            for k, v in r.items():
                rates.setdefault(k, np.zeros(nc))[c] = v
            log(f"cfg4: rate crystal {c} mode {rmode[c]}: {r['rate_4ph']:.6e}")
    log("cfg4: rates took", time.time() - t0)
    np.savez_compressed(os.path.join(OUT, "cfg4.npz"), workload=4,
                        ngrid=np.array([list(cs.modes.ngrid) for cs in cases]),
                        nbands=np.array([cs.nbands for cs in cases]), shell=1,
                        modes=np.array(modes), N=N, P=P, H=H, hashed=hashed,
                        rate_modes=rmode, ii_lo=2, ii_hi=2, **rates)


def part_cfg4r():
    """The 24-crystal batch at its drawn meshes: FP64 rates (per channel, and
    WP4) of the three fixture modes of every crystal over the q' window
    ii = 1..3, from the oracle."""
    from concurrent.futures import ThreadPoolExecutor
    g = np.load(os.path.join(OUT, "cfg4.npz"))
    nc = crystal.BATCH_SIZE
    lo, hi = 1, 3

    def rate_job(c):
        case = crystal.batch_case(c)
        o = Oracle(case)
        mm = np.asarray(g["modes"][c], np.int32)
        t1 = time.time()
        r = o.rta_driver(modes=mm, ii_lo=lo, ii_hi=hi, nthreads=1)
        log(f"cfg4r: crystal {c} Nb {case.nbands} modes {list(mm)} in {time.time() - t1:.1f}s")
        return c, {k: v[mm - 1] for k, v in r.items()}

    rates = {}
    t0 = time.time()
    order = sorted(range(nc), key=lambda c: -int(g["nbands"][c]) ** 3 * int(np.prod(g["ngrid"][c])))
    with ThreadPoolExecutor(THREADS) as ex:
        for c, r in ex.map(rate_job, order):
            for k, v in r.items():
                rates.setdefault(k, np.zeros((nc, 3)))[c] = v
    log("cfg4r: rates took", time.time() - t0)
    np.savez_compressed(os.path.join(OUT, "cfg4r.npz"), workload=4, modes=g["modes"], ii_lo=lo, ii_hi=hi,
                        **rates)


if __name__ == "__main__":
    parts = sys.argv[1:] or ["cfg1", "cfg2", "cfg3"]
    for p in parts:
        globals()["part_" + p]()
        log("done", p)
