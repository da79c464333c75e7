"""Writes the golden fixtures of the GPU parity tests with the CPU oracle.

Run in the development container (the oracle never runs on the GPU box):

    python tests/golden/make_golden.py cfg1 cfg2 cfg3

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


if __name__ == "__main__":
    parts = sys.argv[1:] or ["cfg1", "cfg2", "cfg3"]
    for p in parts:
        globals()["part_" + p]()
        log("done", p)
