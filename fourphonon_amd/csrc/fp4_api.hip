// Host runtime and C ABI of the fp4 library (declared in include/fp4.h).
//
// The context owns one HIP stream (default priority), the device copies of
// the mode data and the IFC clusters, and a few grow-only scratch buffers.
// Every entry point runs the same pipeline:
//   count pass  -> per-item survivor counts (kernel a)
//   host scan   -> exclusive offsets, batch plan under a device-memory cap
//   per batch   -> emit pass (kernel a) -> rate kernel -> segment reduction
// and accumulates per-(channel, mode) results in batch order, so results do
// not depend on the batch size.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/fp4.h"
#include "fp4_dense.hpp"
#include "fp4_iterate.hpp"
#include "fp4_kernels.hpp"

using namespace fp4;

// Device scratch of the allocation-free entry point fp4_np_4ph_modes_noalloc
// (per-item counts and phase-space partials, per-segment totals, the mode
// subset).  Every GPU the library runs on holds its own copy for the life of
// the process.  In host code the name stands for no device address: the
// address on each device comes from hipGetSymbolAddress, taken once per
// device at context creation (scratch_setup) and kept in g_scratch_dev.
static __device__ __align__(256) unsigned char scratch[FP4_NOALLOC_SCRATCH_BYTES];

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHECK(x)                                                                        \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) return fail(FP4_ERR_DEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

// Device address of `scratch` per device ordinal, set by scratch_setup.
constexpr int kMaxScratchDevices = 64;
void *g_scratch_dev[kMaxScratchDevices] = {};
std::mutex g_scratch_mu;

// With `device` current: looks up the device's copy of `scratch` once.
int scratch_setup(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return 0;  // the no-alloc entry point then reports it
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[device]) return 0;
  void *p = nullptr;
  HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return fail(FP4_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned a null address");
  g_scratch_dev[device] = p;
  return 0;
}

void *scratch_ptr(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_dev[device];
}

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  int ensure(size_t need) {
    if (need <= bytes) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (hipMalloc(&p, need) != hipSuccess) {
      p = nullptr;
      return fail(FP4_ERR_NOMEM, "hipMalloc(%zu) failed", need);
    }
    bytes = need;
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T *as() const {
    return reinterpret_cast<T *>(p);
  }
};

enum {
  KS_PREFILTER = 0, KS_EMIT = 1, KS_RATE64 = 2, KS_DENSE = 3, KS_ITER = 4, KS_SCAN3 = 5, KS_RATE3 = 6, KS_SAMPLE = 7,
  KS_N = 8
};

}  // namespace

struct fp4_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int ngrid[3] = {0, 0, 0};
  int nptk = 0, nb = 0, natoms = 0, nelem = 0, nlist = 0;
  double rlattvec[9];
  double lattvec[9];
  double T = 300.0, scalebroad = 1.0, omega_max = 1e100;
  int onlyharmonic = 0;
  std::vector<double> masses;
  std::vector<int> types;
  bool have_modes = false, have_ifc = false, have_ifc3 = false;
  bool have_eig = false;         // eigenvectors given for the current mode data
  std::vector<double> h_energy, h_vel;   // host copies of the mode data on the device
  std::vector<int32_t> h_ijk, h_list;
  // the last 4ph count pass (per-item counts and phase-space terms in cnt /
  // psum), reused while its modes and q' window are asked for again and no
  // other pass or new mode data has touched those buffers
  bool cnt_valid = false;
  std::vector<int> cnt_modes;
  int cnt_lo0 = 0, cnt_nii = 0;
  // mode data
  DevBuf energy, vel, vbound, eig, ijk, list;
  // clusters
  int ntri = 0, L = 1;
  DevBuf clusters, phi, roots;
  std::vector<ClusterRec> h_clusters;
  std::vector<double> h_phi;
  // third-order clusters
  int ntri3 = 0;
  DevBuf clusters3, phi3;
  // scratch
  DevBuf modes_sub, cnt, psum, offsets, segN, segP, lists[3], blk, blk_g, blk_w, seg_first, seg_g, seg_w, proc_g,
      seg_hash;
  DenseState *dense = nullptr;
  BteState *bte = nullptr;
  long long list_cap = 1LL << 26;   // processes per batch
  int last_nblk = 0;                // rate blocks of the last batch (their descriptors are in blk)
  double *w_op = nullptr;           // aggregated operator being assembled (between begin and end)
  long long w_cols = 0;
  // timing
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double ks_ms[KS_N] = {};
  long long ks_n[KS_N] = {};
  DenseTuning dense_tuning;
};

namespace {

ModeView mode_view(const fp4_ctx *c) {
  ModeView m;
  m.energy = c->energy.as<double>();
  m.vel = c->vel.as<double>();
  m.vbound = c->vbound.as<float>();
  m.eig = c->eig.as<double2>();
  m.ijk = c->ijk.as<int>();
  m.list = c->list.as<int>();
  m.nptk = c->nptk;
  m.nb = c->nb;
  m.nlist = c->nlist;
  m.n1 = c->ngrid[0];
  m.n2 = c->ngrid[1];
  m.n3 = c->ngrid[2];
  for (int t = 0; t < 9; t++) m.b[t] = c->rlattvec[t];
  m.scalebroad = c->scalebroad;
  m.omega_max = c->omega_max;
  for (int nu = 0; nu < 3; nu++)
    for (int k = 0; k < 3; k++) m.bf[3 * nu + k] = (float)(c->rlattvec[3 * nu + k] / c->ngrid[nu]);
  m.sbf = (float)std::fabs(c->scalebroad);
  for (int nu = 0; nu < 3; nu++) {
    const int n = c->ngrid[nu];
    m.inv_n[nu] = (n > 0 && (n & (n - 1)) == 0) ? 1.0 / (double)n : 0.0;
  }
  return m;
}

Ifc3View ifc3_view(const fp4_ctx *c) {
  Ifc3View f;
  f.cl = c->clusters3.as<Cluster3Rec>();
  f.phi = c->phi3.as<double>();
  f.roots = c->roots.as<double2>();
  f.ntri = c->ntri3;
  f.L = c->L;
  return f;
}

IfcView ifc_view(const fp4_ctx *c) {
  IfcView f;
  f.cl = c->clusters.as<ClusterRec>();
  f.phi = c->phi.as<double>();
  f.roots = c->roots.as<double2>();
  f.ntri = c->ntri;
  f.L = c->L;
  return f;
}

bool nb_supported(int nb) { return nb == 3 || nb == 6 || nb == 9 || nb == 12 || nb == 18 || nb == 24; }

// Timed launch helper: HIP events on the context stream around one launch.
template <typename F>
int timed(fp4_ctx *c, int kind, F &&launch) {
  HIPCHECK(hipEventRecord(c->ev0, c->stream));
  int rc = launch();
  if (rc) return rc;
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipEventRecord(c->ev1, c->stream));
  HIPCHECK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HIPCHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->ks_ms[kind] += ms;
  c->ks_n[kind] += 1;
  return 0;
}

// ---------------------------------------------------------------- pipeline
struct Plan {
  std::vector<int> modes;          // mm (1-based) of the subset
  // items per (channel, mode) segment: the q' window (four phonons: ii_lo0,
  // nii) or the Nb second-phonon branches (three phonons: nii = Nb, window
  // win_lo0, win_n)
  int ii_lo0 = 0, nii = 0;
  int win_lo0 = 0, win_n = 0;
  long items = 0;
  std::vector<long long> off;      // exclusive offsets, size items+1
  std::vector<long long> segN;     // [3*nsub] processes per (ch, mode)
  std::vector<double> segP;        // [3*nsub] phase space partial sums
};

int finish_counts(fp4_ctx *c, Plan &pl, int nseg, bool want_offsets);

int count_pass(fp4_ctx *c, Plan &pl, bool want_offsets) {
  const int nsub = (int)pl.modes.size();
  pl.items = 3L * nsub * pl.nii;
  const int nseg = 3 * nsub;
  int rc;
  if (c->cnt_valid && c->cnt_lo0 == pl.ii_lo0 && c->cnt_nii == pl.nii && c->cnt_modes == pl.modes)
    return finish_counts(c, pl, nseg, want_offsets);
  c->cnt_valid = false;
  if ((rc = c->modes_sub.ensure(sizeof(int) * std::max(nsub, 1)))) return rc;
  if ((rc = c->cnt.ensure(sizeof(int) * std::max(pl.items, 1L)))) return rc;
  if ((rc = c->psum.ensure(sizeof(double) * std::max(pl.items, 1L)))) return rc;
  if ((rc = c->segN.ensure(sizeof(long long) * std::max(nseg, 1)))) return rc;
  if ((rc = c->segP.ensure(sizeof(double) * std::max(nseg, 1)))) return rc;
  HIPCHECK(hipMemcpyAsync(c->modes_sub.p, pl.modes.data(), sizeof(int) * nsub, hipMemcpyHostToDevice, c->stream));
  PrefilterArgs a{};
  a.m = mode_view(c);
  a.modes = c->modes_sub.as<int>();
  a.nsub = nsub;
  a.nii = pl.nii;
  a.ii_lo0 = pl.ii_lo0;
  a.item_lo = 0;
  a.item_hi = pl.items;
  a.cnt = c->cnt.as<int>();
  a.psum = c->psum.as<double>();
  if ((rc = timed(c, KS_PREFILTER, [&] {
         return launch_prefilter(a, 0, c->stream) ? fail(FP4_ERR_UNSUPPORTED, "Nbands=%d", c->nb) : 0;
       })))
    return rc;
  if ((rc = finish_counts(c, pl, nseg, want_offsets))) return rc;
  c->cnt_valid = true;
  c->cnt_modes = pl.modes;
  c->cnt_lo0 = pl.ii_lo0;
  c->cnt_nii = pl.nii;
  return 0;
}

// Per-segment totals of the count pass and, when asked, the exclusive item
// offsets (host and device copies).
int finish_counts(fp4_ctx *c, Plan &pl, int nseg, bool want_offsets) {
  int rc;
  launch_reduce_counts(c->cnt.as<int>(), c->psum.as<double>(), nseg, pl.nii, c->segN.as<long long>(),
                       c->segP.as<double>(), c->stream);
  HIPCHECK(hipGetLastError());
  pl.segN.assign(nseg, 0);
  pl.segP.assign(nseg, 0.0);
  HIPCHECK(hipMemcpyAsync(pl.segN.data(), c->segN.p, sizeof(long long) * nseg, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipMemcpyAsync(pl.segP.data(), c->segP.p, sizeof(double) * nseg, hipMemcpyDeviceToHost, c->stream));
  if (want_offsets) {
    std::vector<int> h(pl.items);
    HIPCHECK(hipMemcpyAsync(h.data(), c->cnt.p, sizeof(int) * pl.items, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    pl.off.resize(pl.items + 1);
    long long s = 0;
    for (long t = 0; t < pl.items; t++) {
      pl.off[t] = s;
      s += h[t];
    }
    pl.off[pl.items] = s;
    if ((rc = c->offsets.ensure(sizeof(long long) * (pl.items + 1)))) return rc;
    HIPCHECK(hipMemcpyAsync(c->offsets.p, pl.off.data(), sizeof(long long) * (pl.items + 1), hipMemcpyHostToDevice,
                            c->stream));
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

struct Batch {
  long item_lo, item_hi;
  long long p0, p1;
};

std::vector<Batch> plan_batches(const Plan &pl, long long cap) {
  std::vector<Batch> out;
  long lo = 0;
  while (lo < pl.items) {
    // largest hi with off[hi]-off[lo] <= cap
    const long long lim = pl.off[lo] + cap;
    long hi = (long)(std::upper_bound(pl.off.begin() + lo + 1, pl.off.end(), lim) - pl.off.begin()) - 1;
    if (hi <= lo) hi = lo + 1;
    if (pl.off[hi] > pl.off[lo]) out.push_back({lo, hi, pl.off[lo], pl.off[hi]});
    lo = hi;
  }
  return out;
}

// Block descriptors for one batch; returns the per-piece (ch, mode-in-subset) map.
void plan_blocks(const Plan &pl, const Batch &b, std::vector<BlockDesc> &blocks, std::vector<int> &seg_first,
                 std::vector<int> &piece_seg) {
  blocks.clear();
  seg_first.clear();
  piece_seg.clear();
  const int nsub = (int)pl.modes.size();
  const long s_lo = b.item_lo / pl.nii, s_hi = (b.item_hi - 1) / pl.nii;
  for (long s = s_lo; s <= s_hi; s++) {
    const long i0 = std::max(b.item_lo, s * pl.nii), i1 = std::min(b.item_hi, (s + 1) * pl.nii);
    const long long a = pl.off[i0], e = pl.off[i1];
    if (e <= a) continue;
    const int ch = (int)(s / nsub), mi = (int)(s % nsub);
    const int piece = (int)piece_seg.size();
    piece_seg.push_back((int)s);
    seg_first.push_back((int)blocks.size());
    for (long long p = a; p < e; p += 256) {
      BlockDesc d;
      d.start = p - b.p0;
      d.count = (int)std::min<long long>(256, e - p);
      d.ch = ch;
      d.mm = pl.modes[mi];
      d.seg = piece;
      blocks.push_back(d);
    }
  }
  seg_first.push_back((int)blocks.size());
}

int emit_batch(fp4_ctx *c, const Plan &pl, const Batch &b, bool wide) {
  int rc;
  const long long n = b.p1 - b.p0;
  for (int t = 0; t < 3; t++)
    if ((rc = c->lists[t].ensure((wide ? 8 : 4) * (size_t)std::max(n, 1LL)))) return rc;
  PrefilterArgs a{};
  a.m = mode_view(c);
  a.modes = c->modes_sub.as<int>();
  a.nsub = (int)pl.modes.size();
  a.nii = pl.nii;
  a.ii_lo0 = pl.ii_lo0;
  a.item_lo = b.item_lo;
  a.item_hi = b.item_hi;
  a.offsets = c->offsets.as<long long>();
  a.out_base = b.p0;
  for (int t = 0; t < 3; t++) {
    a.out32[t] = c->lists[t].as<int>();
    a.out64[t] = c->lists[t].as<long long>();
  }
  return timed(c, KS_EMIT, [&] { return launch_prefilter(a, wide ? 2 : 1, c->stream) ? fail(FP4_ERR_UNSUPPORTED, "nb") : 0; });
}

// Block plan upload, one rate launch, per-segment reduction and accumulation
// into acc_g / acc_w (in batch order) for the lists of one batch.
template <typename F>
int rate_common(fp4_ctx *c, const Plan &pl, const Batch &b, bool store_gamma, std::vector<double> &acc_g,
                std::vector<double> &acc_w, F &&launch) {
  std::vector<BlockDesc> blocks;
  std::vector<int> seg_first, piece_seg;
  plan_blocks(pl, b, blocks, seg_first, piece_seg);
  c->last_nblk = (int)blocks.size();
  if (blocks.empty()) return 0;
  int rc;
  const int nblk = (int)blocks.size(), npc = (int)piece_seg.size();
  if ((rc = c->blk.ensure(sizeof(BlockDesc) * nblk))) return rc;
  if ((rc = c->blk_g.ensure(sizeof(double) * nblk))) return rc;
  if ((rc = c->blk_w.ensure(sizeof(double) * nblk))) return rc;
  if ((rc = c->seg_first.ensure(sizeof(int) * (npc + 1)))) return rc;
  if ((rc = c->seg_g.ensure(sizeof(double) * npc))) return rc;
  if ((rc = c->seg_w.ensure(sizeof(double) * npc))) return rc;
  if (store_gamma && (rc = c->proc_g.ensure(sizeof(double) * (size_t)std::max(b.p1 - b.p0, 1LL)))) return rc;
  HIPCHECK(hipMemcpyAsync(c->blk.p, blocks.data(), sizeof(BlockDesc) * nblk, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(c->seg_first.p, seg_first.data(), sizeof(int) * (npc + 1), hipMemcpyHostToDevice, c->stream));
  if ((rc = launch(nblk))) return rc;
  launch_seg_reduce(c->blk_g.as<double>(), c->blk_w.as<double>(), c->seg_first.as<int>(), npc,
                    c->seg_g.as<double>(), c->seg_w.as<double>(), c->stream);
  HIPCHECK(hipGetLastError());
  std::vector<double> g(npc), w(npc);
  HIPCHECK(hipMemcpyAsync(g.data(), c->seg_g.p, sizeof(double) * npc, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipMemcpyAsync(w.data(), c->seg_w.p, sizeof(double) * npc, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  for (int t = 0; t < npc; t++) {
    acc_g[piece_seg[t]] = acc_g[piece_seg[t]] + g[t];
    acc_w[piece_seg[t]] = acc_w[piece_seg[t]] + w[t];
  }
  return 0;
}

// Four-phonon rate kernel over a batch whose lists are in c->lists (int32,
// or int64 when wide).
int rate_batch(fp4_ctx *c, const Plan &pl, const Batch &b, bool wide, bool store_gamma, std::vector<double> &acc_g,
               std::vector<double> &acc_w) {
  return rate_common(c, pl, b, store_gamma, acc_g, acc_w, [&](int nblk) {
    RateArgs a{};
    a.m = mode_view(c);
    a.f = ifc_view(c);
    for (int t = 0; t < 3; t++) {
      a.lst[t] = wide ? nullptr : c->lists[t].as<int>();
      a.lst64[t] = wide ? c->lists[t].as<long long>() : nullptr;
    }
    a.blocks = c->blk.as<BlockDesc>();
    a.nblocks = nblk;
    a.T = c->T;
    a.onlyharmonic = c->onlyharmonic;
    a.store_gamma = store_gamma;
    a.proc_gamma = c->proc_g.as<double>();
    a.blk_gamma = c->blk_g.as<double>();
    a.blk_wp4 = c->blk_w.as<double>();
    return timed(c, KS_RATE64, [&] { return launch_rate64(a, c->stream) ? fail(FP4_ERR_UNSUPPORTED, "nb") : 0; });
  });
}

// ------------------------------------------------------- three phonons
// Items are (channel, mode-in-subset, j); the q' window is win_lo0/win_n.
int count3_pass(fp4_ctx *c, Plan &pl, bool want_offsets) {
  c->cnt_valid = false;   // cnt / psum are this pass's now
  const int nsub = (int)pl.modes.size();
  pl.nii = c->nb;
  pl.items = 2L * nsub * c->nb;
  const int nseg = 2 * nsub;
  int rc;
  if ((rc = c->modes_sub.ensure(sizeof(int) * std::max(nsub, 1)))) return rc;
  if ((rc = c->cnt.ensure(sizeof(int) * std::max(pl.items, 1L)))) return rc;
  if ((rc = c->psum.ensure(sizeof(double) * std::max(pl.items, 1L)))) return rc;
  if ((rc = c->segN.ensure(sizeof(long long) * std::max(nseg, 1)))) return rc;
  if ((rc = c->segP.ensure(sizeof(double) * std::max(nseg, 1)))) return rc;
  HIPCHECK(hipMemcpyAsync(c->modes_sub.p, pl.modes.data(), sizeof(int) * nsub, hipMemcpyHostToDevice, c->stream));
  Scan3Args a{};
  a.m = mode_view(c);
  a.modes = c->modes_sub.as<int>();
  a.nsub = nsub;
  a.nii = pl.win_n;
  a.ii_lo0 = pl.win_lo0;
  a.item_lo = 0;
  a.item_hi = pl.items;
  a.cnt = c->cnt.as<int>();
  a.psum = c->psum.as<double>();
  if ((rc = timed(c, KS_SCAN3, [&] {
         return launch_scan3(a, 0, c->stream) ? fail(FP4_ERR_UNSUPPORTED, "Nbands=%d", c->nb) : 0;
       })))
    return rc;
  return finish_counts(c, pl, nseg, want_offsets);
}

int emit3_batch(fp4_ctx *c, const Plan &pl, const Batch &b) {
  int rc;
  const long long n = b.p1 - b.p0;
  for (int t = 0; t < 2; t++)
    if ((rc = c->lists[t].ensure(4 * (size_t)std::max(n, 1LL)))) return rc;
  Scan3Args a{};
  a.m = mode_view(c);
  a.modes = c->modes_sub.as<int>();
  a.nsub = (int)pl.modes.size();
  a.nii = pl.win_n;
  a.ii_lo0 = pl.win_lo0;
  a.item_lo = b.item_lo;
  a.item_hi = b.item_hi;
  a.offsets = c->offsets.as<long long>();
  a.out_base = b.p0;
  for (int t = 0; t < 2; t++) a.out[t] = c->lists[t].as<int>();
  return timed(c, KS_SCAN3, [&] { return launch_scan3(a, 1, c->stream) ? fail(FP4_ERR_UNSUPPORTED, "nb") : 0; });
}

// Three-phonon rate kernel over a batch whose int32 lists are in c->lists[0..1].
int rate3_batch(fp4_ctx *c, const Plan &pl, const Batch &b, bool store_gamma, std::vector<double> &acc_g,
                std::vector<double> &acc_w) {
  return rate_common(c, pl, b, store_gamma, acc_g, acc_w, [&](int nblk) {
    Rate3Args a{};
    a.m = mode_view(c);
    a.f = ifc3_view(c);
    for (int t = 0; t < 2; t++) a.lst[t] = c->lists[t].as<int>();
    a.blocks = c->blk.as<BlockDesc>();
    a.nblocks = nblk;
    a.T = c->T;
    a.onlyharmonic = c->onlyharmonic;
    a.store_gamma = store_gamma;
    a.proc_gamma = c->proc_g.as<double>();
    a.blk_gamma = c->blk_g.as<double>();
    a.blk_wp = c->blk_w.as<double>();
    return timed(c, KS_RATE3, [&] { return launch_rate3(a, c->stream) ? fail(FP4_ERR_UNSUPPORTED, "nb") : 0; });
  });
}

// Per-mode outputs of the three-phonon drivers: "+" Gamma * 5.60626442 * 1e8
// / nptk (processes.f90:1389), "-" the same times 1/2 (:1742); WP3 / nptk and
// WP3 * 0.5 / nptk (:1387, :1467).  scale_gamma = false when the segment sums
// are already sums of per-process Gamma in THz (Ind_driver, :623, :637).
void write3_rates(const fp4_ctx *c, const Plan &pl, const std::vector<double> &g, const std::vector<double> &w,
                  double *rate, double *r_plus, double *r_minus, double *w_plus, double *w_minus, bool scale_gamma) {
  const int nsub = (int)pl.modes.size();
  for (int mi = 0; mi < nsub; mi++) {
    const int m = pl.modes[mi] - 1;
    const double gp = g[mi], gm = g[nsub + mi];
    r_plus[m] = scale_gamma ? gp * (double)kGamma3Units * 1.e8 / c->nptk : gp;
    r_minus[m] = (scale_gamma ? gm * (double)kGamma3Units * 1.e8 / c->nptk : gm) * 5.e-1;
    w_plus[m] = w[mi] / c->nptk;
    w_minus[m] = w[nsub + mi] * 5.e-1 / c->nptk;
    rate[m] = r_plus[m] + r_minus[m];
  }
}

int check_ready3(const fp4_ctx *c, bool need_ifc3) {
  if (!c) return fail(FP4_ERR_ARG, "null context");
  if (!c->have_modes) return fail(FP4_ERR_STATE, "fp4_ctx_set_modes has not been called");
  if (need_ifc3 && !c->have_ifc3) return fail(FP4_ERR_STATE, "fp4_ctx_set_ifc3 has not been called");
  if (need_ifc3 && !c->have_eig) return fail(FP4_ERR_STATE, "fp4_ctx_set_modes was called without eigenvectors");
  HIPCHECK(hipSetDevice(c->device));
  return 0;
}

// e^{2 pi i k / L} for k < L = lcm(N1, N2, N3), shared by both cluster sets.
int ensure_roots(fp4_ctx *c) {
  const int L = std::lcm(std::lcm(c->ngrid[0], c->ngrid[1]), c->ngrid[2]);
  if (c->L == L && c->roots.p) return 0;
  std::vector<double> roots(2 * (size_t)L);
  for (int k = 0; k < L; k++) {
    const double x = 2.0 * M_PI * (double)k / (double)L;
    roots[2 * k] = std::cos(x);
    roots[2 * k + 1] = std::sin(x);
  }
  int rc;
  if ((rc = c->roots.ensure(sizeof(double) * 2 * L))) return rc;
  HIPCHECK(hipMemcpy(c->roots.p, roots.data(), sizeof(double) * 2 * L, hipMemcpyHostToDevice));
  c->L = L;
  return 0;
}

// Lattice coordinates of R (nm) times L / N_nu: the phase multipliers of a cluster site.
int phase_multipliers(const fp4_ctx *c, const double *R, int L, int *out, int cluster) {
  const int ng[3] = {c->ngrid[0], c->ngrid[1], c->ngrid[2]};
  for (int nu = 0; nu < 3; nu++) {
    // b_nu . R = 2 pi m_nu for a lattice vector R
    const double *bv = c->rlattvec + 3 * nu;
    const double x = (bv[0] * R[0] + bv[1] * R[1] + bv[2] * R[2]) / (2.0 * M_PI);
    const double xr = std::nearbyint(x);
    if (std::fabs(x - xr) > 1e-6) return fail(FP4_ERR_ARG, "R of cluster %d is not a lattice vector", cluster);
    out[nu] = (int)xr * (L / ng[nu]);
  }
  return 0;
}

int check_ready(const fp4_ctx *c, bool need_ifc) {
  if (!c) return fail(FP4_ERR_ARG, "null context");
  if (!c->have_modes) return fail(FP4_ERR_STATE, "fp4_ctx_set_modes has not been called");
  if (need_ifc && !c->have_ifc) return fail(FP4_ERR_STATE, "fp4_ctx_set_ifc4 has not been called");
  if (need_ifc && !c->have_eig) return fail(FP4_ERR_STATE, "fp4_ctx_set_modes was called without eigenvectors");
  HIPCHECK(hipSetDevice(c->device));
  return 0;
}

void write_rates(const fp4_ctx *c, const Plan &pl, const std::vector<double> &g, const std::vector<double> &w,
                 double *rate_4ph, double *r_pp, double *r_pm, double *r_mm, double *w_pp, double *w_pm, double *w_mm,
                 bool scale_gamma) {
  const int nsub = (int)pl.modes.size();
  double *R[3] = {r_pp, r_pm, r_mm};
  double *W[3] = {w_pp, w_pm, w_mm};
  std::vector<double> host_e(1);
  for (int mi = 0; mi < nsub; mi++) {
    const int m = pl.modes[mi] - 1;
    for (int ch = 0; ch < 3; ch++) {
      const double gs = g[ch * nsub + mi], ws = w[ch * nsub + mi];
      R[ch][m] = scale_gamma ? gs * (double)kGammaUnits * 1.e9 / c->nptk / c->nptk : gs;
      W[ch][m] = ws / c->nptk / c->nptk;
    }
    rate_4ph[m] = r_pp[m] + r_pm[m] + r_mm[m];
  }
}

}  // namespace

// ======================================================================= ABI
extern "C" {

const char *fp4_last_error(void) { return g_err.c_str(); }

int fp4_ctx_create(const fp4_config *cfg, fp4_ctx **out) {
  if (!cfg || !out) return fail(FP4_ERR_ARG, "null argument");
  *out = nullptr;
  if (cfg->natoms < 1 || cfg->nbands != 3 * cfg->natoms) return fail(FP4_ERR_ARG, "nbands must be 3*natoms");
  if (!nb_supported(cfg->nbands)) return fail(FP4_ERR_UNSUPPORTED, "Nbands=%d not built", cfg->nbands);
  for (int t = 0; t < 3; t++)
    if (cfg->ngrid[t] < 1) return fail(FP4_ERR_ARG, "ngrid must be positive");
  if (cfg->nelements < 1 || !cfg->masses || !cfg->types) return fail(FP4_ERR_ARG, "masses/types missing");
  for (int a = 0; a < cfg->natoms; a++)
    if (cfg->types[a] < 1 || cfg->types[a] > cfg->nelements) return fail(FP4_ERR_ARG, "types out of range");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FP4_ERR_DEVICE, "no HIP device");
  if (cfg->device < 0 || cfg->device >= ndev) return fail(FP4_ERR_DEVICE, "device %d out of range", cfg->device);
  fp4_ctx *c = new fp4_ctx();
  c->device = cfg->device;
  if (hipSetDevice(c->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return fail(FP4_ERR_DEVICE, "stream/event creation failed");
  }
  if (int rc = scratch_setup(c->device)) {
    fp4_ctx_destroy(c);
    return rc;
  }
  for (int t = 0; t < 3; t++) c->ngrid[t] = cfg->ngrid[t];
  c->nptk = cfg->ngrid[0] * cfg->ngrid[1] * cfg->ngrid[2];
  c->nb = cfg->nbands;
  c->natoms = cfg->natoms;
  c->nelem = cfg->nelements;
  std::memcpy(c->rlattvec, cfg->rlattvec, sizeof(c->rlattvec));
  // lattvec = 2 pi inverse(rlattvec)^T
  const double *b = cfg->rlattvec;
  double det = b[0] * (b[4] * b[8] - b[7] * b[5]) - b[3] * (b[1] * b[8] - b[7] * b[2]) + b[6] * (b[1] * b[5] - b[4] * b[2]);
  if (!(std::fabs(det) > 0)) {
    fp4_ctx_destroy(c);
    return fail(FP4_ERR_ARG, "singular rlattvec");
  }
  c->T = cfg->T;
  c->scalebroad = cfg->scalebroad;
  c->omega_max = cfg->omega_max;
  c->onlyharmonic = cfg->onlyharmonic;
  c->masses.assign(cfg->masses, cfg->masses + cfg->nelements);
  c->types.assign(cfg->types, cfg->types + cfg->natoms);
  *out = c;
  return FP4_OK;
}

int fp4_ctx_destroy(fp4_ctx *c) {
  if (!c) return FP4_OK;
  (void)hipSetDevice(c->device);
  DevBuf *bufs[] = {&c->energy, &c->vel, &c->vbound, &c->eig, &c->ijk, &c->list, &c->clusters, &c->phi, &c->roots,
                    &c->clusters3, &c->phi3,
                    &c->modes_sub, &c->cnt, &c->psum, &c->offsets, &c->segN, &c->segP, &c->lists[0], &c->lists[1],
                    &c->lists[2], &c->blk, &c->blk_g, &c->blk_w, &c->seg_first, &c->seg_g, &c->seg_w, &c->proc_g,
                    &c->seg_hash};
  for (DevBuf *b : bufs) b->release();
  dense_destroy(c->dense);
  c->dense = nullptr;
  bte_destroy(c->bte);
  c->bte = nullptr;
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return FP4_OK;
}

int fp4_ctx_set_temperature(fp4_ctx *c, double T) {
  if (!c) return fail(FP4_ERR_ARG, "null context");
  if (!(T > 0)) return fail(FP4_ERR_ARG, "T must be positive");
  c->T = T;
  return FP4_OK;
}

int fp4_ctx_set_modes(fp4_ctx *c, const double *energy, const double *velocity, const double *eigenvect,
                      const int32_t *IJK, int32_t nlist, const int32_t *List) {
  if (!c || !energy || !velocity || !IJK || !List) return fail(FP4_ERR_ARG, "null argument");
  if (nlist < 1 || nlist > c->nptk) return fail(FP4_ERR_ARG, "Nlist out of range");
  for (int t = 0; t < nlist; t++)
    if (List[t] < 1 || List[t] > c->nptk) return fail(FP4_ERR_ARG, "List(%d)=%d out of range", t + 1, List[t]);
  for (int q = 0; q < c->nptk; q++)
    for (int t = 0; t < 3; t++)
      if (IJK[3 * q + t] < 0 || IJK[3 * q + t] >= c->ngrid[t]) return fail(FP4_ERR_ARG, "IJK out of range");
  HIPCHECK(hipSetDevice(c->device));
  const size_t np = c->nptk, nb = c->nb;
  // the reference's drivers take the mode data as arguments, so a caller
  // passes the same arrays again and again: unchanged frequencies,
  // velocities, grid and irreducible list stay on the device (and keep the
  // cached count pass valid); only eigenvectors, when given, are uploaded
  const bool same = c->have_modes && c->nlist == nlist && c->h_energy.size() == np * nb &&
                    std::memcmp(c->h_energy.data(), energy, sizeof(double) * np * nb) == 0 &&
                    std::memcmp(c->h_vel.data(), velocity, sizeof(double) * np * nb * 3) == 0 &&
                    std::memcmp(c->h_ijk.data(), IJK, sizeof(int32_t) * 3 * np) == 0 &&
                    std::memcmp(c->h_list.data(), List, sizeof(int32_t) * nlist) == 0;
  if (same) {
    if (eigenvect) {
      std::vector<double> eig(2 * np * nb * nb);
      for (size_t comp = 0; comp < nb; comp++)
        for (size_t b = 0; b < nb; b++)
          for (size_t q = 0; q < np; q++) {
            const size_t src = q + np * (b + nb * comp), dst = (q * nb + b) * nb + comp;
            eig[2 * dst] = eigenvect[2 * src];
            eig[2 * dst + 1] = eigenvect[2 * src + 1];
          }
      HIPCHECK(hipMemcpy(c->eig.p, eig.data(), sizeof(double) * 2 * np * nb * nb, hipMemcpyHostToDevice));
      c->have_eig = true;
    }
    return FP4_OK;
  }
  int rc;
  if ((rc = c->energy.ensure(sizeof(double) * np * nb)) || (rc = c->vel.ensure(sizeof(double) * np * nb * 3)) ||
      (rc = c->vbound.ensure(sizeof(float) * np * nb)) || (rc = c->eig.ensure(sizeof(double) * 2 * np * nb * nb)) ||
      (rc = c->ijk.ensure(sizeof(int) * 3 * np)) || (rc = c->list.ensure(sizeof(int) * nlist)))
    return rc;
  // eigenvectors: reference (q, b, comp) with q fastest -> [q][b][comp]
  std::vector<double> eig(eigenvect ? 2 * np * nb * nb : 0);
  if (eigenvect)
    for (size_t comp = 0; comp < nb; comp++)
      for (size_t b = 0; b < nb; b++)
        for (size_t q = 0; q < np; q++) {
          const size_t src = q + np * (b + nb * comp), dst = (q * nb + b) * nb + comp;
          eig[2 * dst] = eigenvect[2 * src];
          eig[2 * dst + 1] = eigenvect[2 * src + 1];
        }
  // conservative broadening bound per mode (see fp4_prefilter.hip)
  double bm = 0.0;
  for (int nu = 0; nu < 3; nu++) {
    const double *bv = c->rlattvec + 3 * nu;
    bm = std::max(bm, std::sqrt(bv[0] * bv[0] + bv[1] * bv[1] + bv[2] * bv[2]) / c->ngrid[nu]);
  }
  std::vector<float> vb(np * nb);
  for (size_t b = 0; b < nb; b++)
    for (size_t q = 0; q < np; q++) {
      const double *v = velocity;
      const double vx = v[q + np * (b + nb * 0)], vy = v[q + np * (b + nb * 1)], vz = v[q + np * (b + nb * 2)];
      const double val = std::fabs(c->scalebroad) * bm * std::sqrt(vx * vx + vy * vy + vz * vz) / std::sqrt(2.0);
      float f = (float)(val * (1.0 + 1e-6));
      vb[b * np + q] = std::nextafter(f, INFINITY);
    }
  HIPCHECK(hipMemcpy(c->energy.p, energy, sizeof(double) * np * nb, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->vel.p, velocity, sizeof(double) * np * nb * 3, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->vbound.p, vb.data(), sizeof(float) * np * nb, hipMemcpyHostToDevice));
  if (eigenvect) HIPCHECK(hipMemcpy(c->eig.p, eig.data(), sizeof(double) * 2 * np * nb * nb, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->ijk.p, IJK, sizeof(int) * 3 * np, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->list.p, List, sizeof(int) * nlist, hipMemcpyHostToDevice));
  c->nlist = nlist;
  c->have_modes = true;
  c->have_eig = eigenvect != nullptr;
  c->cnt_valid = false;
  c->h_energy.assign(energy, energy + np * nb);
  c->h_vel.assign(velocity, velocity + np * nb * 3);
  c->h_ijk.assign(IJK, IJK + 3 * np);
  c->h_list.assign(List, List + nlist);
  return FP4_OK;
}

int fp4_ctx_set_ifc4(fp4_ctx *c, int32_t ntri, const double *Phi, const double *R_j, const double *R_k,
                     const double *R_l, const int32_t *Ii, const int32_t *Ij, const int32_t *Ik, const int32_t *Il) {
  if (!c || !Phi || !R_j || !R_k || !R_l || !Ii || !Ij || !Ik || !Il) return fail(FP4_ERR_ARG, "null argument");
  if (ntri < 1) return fail(FP4_ERR_ARG, "Ntri must be positive");
  HIPCHECK(hipSetDevice(c->device));
  int rc;
  if ((rc = ensure_roots(c))) return rc;
  const int L = c->L;
  c->h_clusters.assign(ntri, ClusterRec{});
  for (int t = 0; t < ntri; t++) {
    ClusterRec &r = c->h_clusters[t];
    const int32_t *ix[4] = {Ii, Ij, Ik, Il};
    for (int a = 0; a < 4; a++) {
      if (ix[a][t] < 1 || ix[a][t] > c->natoms) return fail(FP4_ERR_ARG, "Index out of range at cluster %d", t + 1);
      r.atom[a] = ix[a][t] - 1;
    }
    const double *Rs[3] = {R_j + 3 * (size_t)t, R_k + 3 * (size_t)t, R_l + 3 * (size_t)t};
    int *outs[3] = {r.pj, r.pk, r.pl};
    for (int w = 0; w < 3; w++)
      if ((rc = phase_multipliers(c, Rs[w], L, outs[w], t + 1))) return rc;
    double mp = c->masses[c->types[r.atom[0]] - 1] * c->masses[c->types[r.atom[1]] - 1];
    mp = mp * c->masses[c->types[r.atom[2]] - 1];
    mp = mp * c->masses[c->types[r.atom[3]] - 1];
    r.minv = 1.0 / std::sqrt(mp);
  }
  if ((rc = c->clusters.ensure(sizeof(ClusterRec) * ntri)) || (rc = c->phi.ensure(sizeof(double) * 81 * ntri)))
    return rc;
  c->h_phi.assign(Phi, Phi + 81 * (size_t)ntri);
  HIPCHECK(hipMemcpy(c->clusters.p, c->h_clusters.data(), sizeof(ClusterRec) * ntri, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->phi.p, Phi, sizeof(double) * 81 * ntri, hipMemcpyHostToDevice));
  c->ntri = ntri;
  dense_destroy(c->dense);
  int drc = 0;
  c->dense = dense_create(c->h_clusters, c->h_phi, c->natoms, c->nb, L, &drc);
  if (!c->dense) return fail(drc, "%s", dense_error());
  c->have_ifc = true;
  return FP4_OK;
}

static void all_modes(const fp4_ctx *c, Plan &pl) {
  pl.modes.resize((size_t)c->nb * c->nlist);
  std::iota(pl.modes.begin(), pl.modes.end(), 1);
  pl.ii_lo0 = 0;
  pl.nii = c->nptk;
}

int fp4_np_4ph(fp4_ctx *c, int64_t *N_pp, double *P_pp, int64_t *N_pm, double *P_pm, int64_t *N_mm, double *P_mm) {
  int rc;
  if ((rc = check_ready(c, false))) return rc;
  if (!N_pp || !P_pp || !N_pm || !P_pm || !N_mm || !P_mm) return fail(FP4_ERR_ARG, "null output");
  Plan pl;
  all_modes(c, pl);
  if ((rc = count_pass(c, pl, false))) return rc;
  const int nm = (int)pl.modes.size();
  int64_t *N[3] = {N_pp, N_pm, N_mm};
  double *P[3] = {P_pp, P_pm, P_mm};
  for (int ch = 0; ch < 3; ch++)
    for (int m = 0; m < nm; m++) {
      N[ch][m] = pl.segN[ch * nm + m];
      P[ch][m] = pl.segP[ch * nm + m];
    }
  return FP4_OK;
}

static int verify_counts(const Plan &pl, const int64_t *N_pp, const int64_t *N_pm, const int64_t *N_mm) {
  const int nm = (int)pl.modes.size();
  const int64_t *N[3] = {N_pp, N_pm, N_mm};
  for (int ch = 0; ch < 3; ch++)
    for (int m = 0; m < nm; m++)
      if (N[ch][m] != pl.segN[ch * nm + m])
        return fail(FP4_ERR_OVERFLOW, "count mismatch: channel %d mode %d has %lld processes, caller passed %lld", ch,
                    m + 1, (long long)pl.segN[ch * nm + m], (long long)N[ch][m]);
  return 0;
}

int fp4_ind_only_4ph(fp4_ctx *c, const int64_t *N_pp, const int64_t *N_pm, const int64_t *N_mm, int64_t *a2,
                     int64_t *a3, int64_t *a4, int64_t *b2, int64_t *b3, int64_t *b4, int64_t *c2, int64_t *c3,
                     int64_t *c4) {
  int rc;
  if ((rc = check_ready(c, false))) return rc;
  if (!N_pp || !N_pm || !N_mm) return fail(FP4_ERR_ARG, "null counts");
  Plan pl;
  all_modes(c, pl);
  if ((rc = count_pass(c, pl, true))) return rc;
  if ((rc = verify_counts(pl, N_pp, N_pm, N_mm))) return rc;
  int64_t *outs[3][3] = {{a2, a3, a4}, {b2, b3, b4}, {c2, c3, c4}};
  const long per_ch = pl.items / 3;
  for (const Batch &b : plan_batches(pl, c->list_cap)) {
    if ((rc = emit_batch(c, pl, b, true))) return rc;
    // copy each channel's part of the batch to its host lists
    for (int ch = 0; ch < 3; ch++) {
      const long lo = std::max(b.item_lo, ch * per_ch), hi = std::min(b.item_hi, (ch + 1) * per_ch);
      if (hi <= lo) continue;
      const long long p0 = pl.off[lo], p1 = pl.off[hi], base = pl.off[ch * per_ch];
      if (p1 <= p0) continue;
      for (int t = 0; t < 3; t++) {
        if (!outs[ch][t]) return fail(FP4_ERR_ARG, "null list output");
        HIPCHECK(hipMemcpyAsync(outs[ch][t] + (p0 - base), c->lists[t].as<long long>() + (p0 - b.p0),
                                sizeof(long long) * (p1 - p0), hipMemcpyDeviceToHost, c->stream));
      }
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
  }
  return FP4_OK;
}

int fp4_rta_4ph_modes(fp4_ctx *c, int32_t precision, int32_t nmodes, const int32_t *modes, int32_t ii_lo,
                      int32_t ii_hi, double *rate_4ph, double *r_pp, double *r_pm, double *r_mm, double *w_pp,
                      double *w_pm, double *w_mm, int64_t *nproc) {
  int rc;
  if ((rc = check_ready(c, true))) return rc;
  if (!rate_4ph || !r_pp || !r_pm || !r_mm || !w_pp || !w_pm || !w_mm) return fail(FP4_ERR_ARG, "null output");
  if (nmodes < 1 || !modes) return fail(FP4_ERR_ARG, "empty mode subset");
  if (ii_lo < 1 || ii_hi > c->nptk || ii_hi < ii_lo) return fail(FP4_ERR_ARG, "bad ii window %d..%d", ii_lo, ii_hi);
  Plan pl;
  pl.modes.assign(modes, modes + nmodes);
  for (size_t t = 0; t < pl.modes.size(); t++) {
    if (pl.modes[t] < 1 || pl.modes[t] > c->nb * c->nlist) return fail(FP4_ERR_ARG, "mode %d out of range", pl.modes[t]);
    if (t && pl.modes[t] <= pl.modes[t - 1]) return fail(FP4_ERR_ARG, "modes must be strictly increasing");
  }
  pl.ii_lo0 = ii_lo - 1;
  pl.nii = ii_hi - ii_lo + 1;
  const int nseg = 3 * nmodes;
  std::vector<double> g(nseg, 0.0), w(nseg, 0.0);
  if (precision == FP4_PREC_FP64) {
    if ((rc = count_pass(c, pl, true))) return rc;
    for (const Batch &b : plan_batches(pl, c->list_cap)) {
      if ((rc = emit_batch(c, pl, b, false))) return rc;
      if ((rc = rate_batch(c, pl, b, false, false, g, w))) return rc;
    }
  } else if (precision == FP4_PREC_FP32 || precision == FP4_PREC_FP64_DENSE || precision == FP4_PREC_BF16X3 ||
             precision == FP4_PREC_FP16X3 || precision == FP4_PREC_FP16X2) {
    DenseRun dr;
    dr.m = mode_view(c);
    dr.roots = c->roots.as<double2>();
    dr.T = c->T;
    dr.onlyharmonic = c->onlyharmonic;
    dr.natoms = c->natoms;
    dr.precision = precision;
    dr.stream = c->stream;
    dr.tuning = &c->dense_tuning;
    double ms = 0.0;
    std::vector<long long> n;
    rc = dense_rates(c->dense, dr, pl.modes, pl.ii_lo0, pl.nii, g, w, n, &ms);
    if (rc) return fail(rc, "dense path: %s", dense_error());
    c->ks_ms[KS_DENSE] += ms;
    c->ks_n[KS_DENSE] += 1;
    pl.segN = n;
  } else {
    return fail(FP4_ERR_ARG, "unknown precision %d", precision);
  }
  write_rates(c, pl, g, w, rate_4ph, r_pp, r_pm, r_mm, w_pp, w_pm, w_mm, true);
  if (nproc) *nproc = std::accumulate(pl.segN.begin(), pl.segN.end(), 0LL);
  return FP4_OK;
}

int fp4_rta_4ph(fp4_ctx *c, int32_t precision, double *rate_4ph, double *r_pp, double *r_pm, double *r_mm,
                double *w_pp, double *w_pm, double *w_mm) {
  int rc;
  if ((rc = check_ready(c, true))) return rc;
  std::vector<int> modes((size_t)c->nb * c->nlist);
  std::iota(modes.begin(), modes.end(), 1);
  const size_t nm = modes.size();
  double *outs[7] = {rate_4ph, r_pp, r_pm, r_mm, w_pp, w_pm, w_mm};
  for (double *o : outs) {
    if (!o) return fail(FP4_ERR_ARG, "null output");
    std::fill(o, o + nm, 0.0);
  }
  return fp4_rta_4ph_modes(c, precision, (int)nm, modes.data(), 1, c->nptk, rate_4ph, r_pp, r_pm, r_mm, w_pp, w_pm,
                           w_mm, nullptr);
}

int fp4_rta_4ph_using_ind(fp4_ctx *c, const int64_t *N_pp, const int64_t *N_pm, const int64_t *N_mm,
                          const int64_t *a2, const int64_t *a3, const int64_t *a4, const int64_t *b2,
                          const int64_t *b3, const int64_t *b4, const int64_t *c2, const int64_t *c3,
                          const int64_t *c4, double *rate_4ph, double *r_pp, double *r_pm, double *r_mm,
                          double *w_pp, double *w_pm, double *w_mm) {
  int rc;
  if ((rc = check_ready(c, true))) return rc;
  const int nm = c->nb * c->nlist;
  const int64_t *N[3] = {N_pp, N_pm, N_mm};
  const int64_t *L3[3][3] = {{a2, a3, a4}, {b2, b3, b4}, {c2, c3, c4}};
  double *outs[7] = {rate_4ph, r_pp, r_pm, r_mm, w_pp, w_pm, w_mm};
  for (double *o : outs) {
    if (!o) return fail(FP4_ERR_ARG, "null output");
    std::fill(o, o + nm, 0.0);
  }
  // Items are (channel, mode) with one "ii" each; offsets from the given counts.
  Plan pl;
  pl.modes.resize(nm);
  std::iota(pl.modes.begin(), pl.modes.end(), 1);
  pl.nii = 1;
  pl.items = 3L * nm;
  pl.off.assign(pl.items + 1, 0);
  for (int ch = 0; ch < 3; ch++)
    for (int m = 0; m < nm; m++) {
      if (!N[ch] || N[ch][m] < 0) return fail(FP4_ERR_ARG, "bad counts");
      pl.off[(long)ch * nm + m + 1] = pl.off[(long)ch * nm + m] + N[ch][m];
    }
  const long long tot[3] = {pl.off[nm] - pl.off[0], pl.off[2L * nm] - pl.off[nm], pl.off[3L * nm] - pl.off[2L * nm]};
  for (int ch = 0; ch < 3; ch++)
    for (int t = 0; t < 3; t++)
      if (tot[ch] > 0 && !L3[ch][t]) return fail(FP4_ERR_ARG, "null list");
  // range check of the caller's entries happens on the host before upload
  const long long maxv = (long long)c->nptk * c->nb;
  std::vector<double> g(3 * nm, 0.0), w(3 * nm, 0.0);
  for (const Batch &b : plan_batches(pl, c->list_cap)) {
    const long long n = b.p1 - b.p0;
    for (int t = 0; t < 3; t++)
      if ((rc = c->lists[t].ensure(8 * (size_t)n))) return rc;
    for (long it = b.item_lo; it < b.item_hi; it++) {
      const int ch = (int)(it / nm);
      const long long p0 = pl.off[it], p1 = pl.off[it + 1];
      if (p1 <= p0) continue;
      const long long base = pl.off[(long)ch * nm];
      for (int t = 0; t < 3; t++) {
        const int64_t *src = L3[ch][t] + (p0 - base);
        for (long long x = 0; x < p1 - p0; x++)
          if (src[x] < 1 || src[x] > maxv) return fail(FP4_ERR_ARG, "list entry out of range");
        HIPCHECK(hipMemcpyAsync(c->lists[t].as<long long>() + (p0 - b.p0), src, sizeof(long long) * (p1 - p0),
                                hipMemcpyHostToDevice, c->stream));
      }
    }
    if ((rc = rate_batch(c, pl, b, true, false, g, w))) return rc;
  }
  write_rates(c, pl, g, w, rate_4ph, r_pp, r_pm, r_mm, w_pp, w_pm, w_mm, true);
  return FP4_OK;
}

int fp4_ind_4ph(fp4_ctx *c, const int64_t *N_pp, const int64_t *N_pm, const int64_t *N_mm, int64_t *a2, int64_t *a3,
                int64_t *a4, int64_t *b2, int64_t *b3, int64_t *b4, int64_t *c2, int64_t *c3, int64_t *c4,
                double *G_pp, double *G_pm, double *G_mm, double *rate_4ph, double *r_pp, double *r_pm,
                double *r_mm, double *w_pp, double *w_pm, double *w_mm) {
  int rc;
  if ((rc = check_ready(c, true))) return rc;
  Plan pl;
  all_modes(c, pl);
  const int nm = (int)pl.modes.size();
  double *outs[7] = {rate_4ph, r_pp, r_pm, r_mm, w_pp, w_pm, w_mm};
  for (double *o : outs) {
    if (!o) return fail(FP4_ERR_ARG, "null output");
    std::fill(o, o + nm, 0.0);
  }
  if ((rc = count_pass(c, pl, true))) return rc;
  if ((rc = verify_counts(pl, N_pp, N_pm, N_mm))) return rc;
  int64_t *outl[3][3] = {{a2, a3, a4}, {b2, b3, b4}, {c2, c3, c4}};
  double *outg[3] = {G_pp, G_pm, G_mm};
  const long per_ch = pl.items / 3;
  std::vector<double> g(3 * nm, 0.0), w(3 * nm, 0.0);
  for (const Batch &b : plan_batches(pl, c->list_cap)) {
    if ((rc = emit_batch(c, pl, b, true))) return rc;
    if ((rc = rate_batch(c, pl, b, true, true, g, w))) return rc;
    for (int ch = 0; ch < 3; ch++) {
      const long lo = std::max(b.item_lo, ch * per_ch), hi = std::min(b.item_hi, (ch + 1) * per_ch);
      if (hi <= lo) continue;
      const long long p0 = pl.off[lo], p1 = pl.off[hi], base = pl.off[ch * per_ch];
      if (p1 <= p0) continue;
      for (int t = 0; t < 3; t++) {
        if (!outl[ch][t]) return fail(FP4_ERR_ARG, "null list output");
        HIPCHECK(hipMemcpyAsync(outl[ch][t] + (p0 - base), c->lists[t].as<long long>() + (p0 - b.p0),
                                sizeof(long long) * (p1 - p0), hipMemcpyDeviceToHost, c->stream));
      }
      if (!outg[ch]) return fail(FP4_ERR_ARG, "null Gamma output");
      HIPCHECK(hipMemcpyAsync(outg[ch] + (p0 - base), c->proc_g.as<double>() + (p0 - b.p0),
                              sizeof(double) * (p1 - p0), hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
  }
  // per-process Gamma are already in THz; Ind_driver_4ph sums them per mode
  write_rates(c, pl, g, w, rate_4ph, r_pp, r_pm, r_mm, w_pp, w_pm, w_mm, false);
  return FP4_OK;
}

int fp4_vp_4ph(fp4_ctx *c, int32_t channel, int64_t n, const int32_t *idx, double *out) {
  int rc;
  if ((rc = check_ready(c, true))) return rc;
  if (channel < 0 || channel > 2) return fail(FP4_ERR_ARG, "channel must be 0, 1 or 2");
  if (n < 0 || (n > 0 && (!idx || !out))) return fail(FP4_ERR_ARG, "bad arguments");
  if (n == 0) return FP4_OK;
  for (int64_t t = 0; t < n; t++) {
    const int32_t *r = idx + 8 * t;
    for (int a = 0; a < 4; a++)
      if (r[a] < 1 || r[a] > c->nb) return fail(FP4_ERR_ARG, "branch index out of range in row %lld", (long long)t);
    for (int a = 4; a < 8; a++)
      if (r[a] < 1 || r[a] > c->nptk) return fail(FP4_ERR_ARG, "q index out of range in row %lld", (long long)t);
  }
  DevBuf di, dout;
  if ((rc = di.ensure(sizeof(int) * 8 * n)) || (rc = dout.ensure(sizeof(double) * 2 * n))) {
    di.release();
    return rc;
  }
  VpArgs a{};
  a.m = mode_view(c);
  a.f = ifc_view(c);
  a.ch = channel;
  a.n = n;
  a.idx = di.as<int>();
  a.out = dout.as<double2>();
  hipError_t e = hipMemcpyAsync(di.p, idx, sizeof(int) * 8 * n, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    launch_vp64(a, c->stream);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout.p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  di.release();
  dout.release();
  if (e != hipSuccess) return fail(FP4_ERR_DEVICE, "fp4_vp_4ph: %s", hipGetErrorString(e));
  return FP4_OK;
}

// A strictly increasing subset of modes over the whole q' range.
static int subset_plan(const fp4_ctx *c, int32_t nmodes, const int32_t *modes, Plan &pl) {
  if (nmodes < 1 || !modes) return fail(FP4_ERR_ARG, "bad mode subset");
  pl.modes.assign(modes, modes + nmodes);
  for (size_t t = 0; t < pl.modes.size(); t++) {
    if (pl.modes[t] < 1 || pl.modes[t] > c->nb * c->nlist) return fail(FP4_ERR_ARG, "mode %d out of range", pl.modes[t]);
    if (t && pl.modes[t] <= pl.modes[t - 1]) return fail(FP4_ERR_ARG, "modes must be strictly increasing");
  }
  pl.ii_lo0 = 0;
  pl.nii = c->nptk;
  return 0;
}

int fp4_np_4ph_modes(fp4_ctx *c, int32_t nmodes, const int32_t *modes, int64_t *N_pp, double *P_pp, int64_t *N_pm,
                     double *P_pm, int64_t *N_mm, double *P_mm) {
  int rc;
  if ((rc = check_ready(c, false))) return rc;
  if (!N_pp || !P_pp || !N_pm || !P_pm || !N_mm || !P_mm) return fail(FP4_ERR_ARG, "null output");
  Plan pl;
  if ((rc = subset_plan(c, nmodes, modes, pl))) return rc;
  if ((rc = count_pass(c, pl, false))) return rc;
  int64_t *N[3] = {N_pp, N_pm, N_mm};
  double *P[3] = {P_pp, P_pm, P_mm};
  for (int ch = 0; ch < 3; ch++)
    for (int m = 0; m < nmodes; m++) {
      N[ch][m] = pl.segN[ch * nmodes + m];
      P[ch][m] = pl.segP[ch * nmodes + m];
    }
  return FP4_OK;
}

// fp4_np_4ph_modes on the device's `scratch` array instead of context
// buffers: the same count-pass and reduction kernels, no device allocation.
// Scratch layout (256-byte aligned parts): modes[nmodes] int32,
// segN[3*nmodes] int64, segP[3*nmodes] double, cnt[items] int32,
// psum[items] double, items = 3*nmodes*nptk.  The count pass writes every
// cnt/psum entry and the reduction every segN/segP entry, so nothing left
// by an earlier call is read.
int fp4_np_4ph_modes_noalloc(fp4_ctx *c, int32_t nmodes, const int32_t *modes, int64_t *N_pp, double *P_pp,
                             int64_t *N_pm, double *P_pm, int64_t *N_mm, double *P_mm) {
  int rc;
  if ((rc = check_ready(c, false))) return rc;
  if (!N_pp || !P_pp || !N_pm || !P_pm || !N_mm || !P_mm) return fail(FP4_ERR_ARG, "null output");
  if (nmodes < 1 || !modes) return fail(FP4_ERR_ARG, "bad mode subset");
  for (int t = 0; t < nmodes; t++) {
    if (modes[t] < 1 || modes[t] > c->nb * c->nlist) return fail(FP4_ERR_ARG, "mode %d out of range", modes[t]);
    if (t && modes[t] <= modes[t - 1]) return fail(FP4_ERR_ARG, "modes must be strictly increasing");
  }
  unsigned char *base = static_cast<unsigned char *>(scratch_ptr(c->device));
  if (!base) return fail(FP4_ERR_STATE, "no scratch address for device %d", c->device);
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t nseg = 3 * (size_t)nmodes, items = nseg * (size_t)c->nptk;
  const size_t o_segN = up(sizeof(int) * (size_t)nmodes);
  const size_t o_segP = o_segN + up(sizeof(long long) * nseg);
  const size_t o_cnt = o_segP + up(sizeof(double) * nseg);
  const size_t o_psum = o_cnt + up(sizeof(int) * items);
  const size_t need = o_psum + sizeof(double) * items;
  if (need > (size_t)FP4_NOALLOC_SCRATCH_BYTES) {
    const size_t per_mode = 3 * (size_t)c->nptk * (sizeof(int) + sizeof(double)) + 3 * 16 + 4;
    return fail(FP4_ERR_ARG,
                "%d modes need %zu bytes of scratch, the limit is %zu (FP4_NOALLOC_SCRATCH_BYTES): pass at most "
                "about %zu modes per call",
                nmodes, need, (size_t)FP4_NOALLOC_SCRATCH_BYTES,
                (size_t)(FP4_NOALLOC_SCRATCH_BYTES - 4 * 256) / per_mode);
  }
  int *d_modes = reinterpret_cast<int *>(base);
  long long *d_segN = reinterpret_cast<long long *>(base + o_segN);
  double *d_segP = reinterpret_cast<double *>(base + o_segP);
  int *d_cnt = reinterpret_cast<int *>(base + o_cnt);
  double *d_psum = reinterpret_cast<double *>(base + o_psum);
  HIPCHECK(hipMemcpyAsync(d_modes, modes, sizeof(int) * nmodes, hipMemcpyHostToDevice, c->stream));
  PrefilterArgs a{};
  a.m = mode_view(c);
  a.modes = d_modes;
  a.nsub = nmodes;
  a.nii = c->nptk;
  a.ii_lo0 = 0;
  a.item_lo = 0;
  a.item_hi = (long)items;
  a.cnt = d_cnt;
  a.psum = d_psum;
  if ((rc = timed(c, KS_PREFILTER, [&] {
         return launch_prefilter(a, 0, c->stream) ? fail(FP4_ERR_UNSUPPORTED, "Nbands=%d", c->nb) : 0;
       })))
    return rc;
  launch_reduce_counts(d_cnt, d_psum, (int)nseg, c->nptk, d_segN, d_segP, c->stream);
  HIPCHECK(hipGetLastError());
  int64_t *N[3] = {N_pp, N_pm, N_mm};
  double *P[3] = {P_pp, P_pm, P_mm};
  for (int ch = 0; ch < 3; ch++) {
    HIPCHECK(hipMemcpyAsync(N[ch], d_segN + (size_t)ch * nmodes, sizeof(long long) * nmodes, hipMemcpyDeviceToHost,
                            c->stream));
    HIPCHECK(hipMemcpyAsync(P[ch], d_segP + (size_t)ch * nmodes, sizeof(double) * nmodes, hipMemcpyDeviceToHost,
                            c->stream));
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
  return FP4_OK;
}

int fp4_list_hash_4ph_modes(fp4_ctx *c, int32_t nmodes, const int32_t *modes, int64_t *N_pp, int64_t *N_pm,
                            int64_t *N_mm, uint64_t *H_pp, uint64_t *H_pm, uint64_t *H_mm) {
  int rc;
  if ((rc = check_ready(c, false))) return rc;
  if (!N_pp || !N_pm || !N_mm || !H_pp || !H_pm || !H_mm) return fail(FP4_ERR_ARG, "null output");
  Plan pl;
  if ((rc = subset_plan(c, nmodes, modes, pl))) return rc;
  if ((rc = count_pass(c, pl, true))) return rc;
  const int nseg = 3 * nmodes;
  if ((rc = c->seg_hash.ensure(sizeof(unsigned long long) * nseg))) return rc;
  HIPCHECK(hipMemsetAsync(c->seg_hash.p, 0, sizeof(unsigned long long) * nseg, c->stream));
  PrefilterArgs a{};
  a.m = mode_view(c);
  a.modes = c->modes_sub.as<int>();
  a.nsub = nmodes;
  a.nii = pl.nii;
  a.ii_lo0 = pl.ii_lo0;
  a.item_lo = 0;
  a.item_hi = pl.items;
  a.offsets = c->offsets.as<long long>();
  a.out_base = 0;
  a.seg_hash = c->seg_hash.as<unsigned long long>();
  if ((rc = timed(c, KS_EMIT, [&] {
         return launch_prefilter(a, 3, c->stream) ? fail(FP4_ERR_UNSUPPORTED, "Nbands=%d", c->nb) : 0;
       })))
    return rc;
  int64_t *N[3] = {N_pp, N_pm, N_mm};
  uint64_t *H[3] = {H_pp, H_pm, H_mm};
  for (int ch = 0; ch < 3; ch++) {
    for (int m = 0; m < nmodes; m++) N[ch][m] = pl.segN[ch * nmodes + m];
    HIPCHECK(hipMemcpyAsync(H[ch], c->seg_hash.as<unsigned long long>() + (size_t)ch * nmodes,
                            sizeof(unsigned long long) * nmodes, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
  return FP4_OK;
}

int fp4_ind_only_4ph_modes(fp4_ctx *c, int32_t nmodes, const int32_t *modes, int64_t cap, int64_t *N_pp,
                           int64_t *N_pm, int64_t *N_mm, int64_t *a2, int64_t *a3, int64_t *a4, int64_t *b2,
                           int64_t *b3, int64_t *b4, int64_t *c2, int64_t *c3, int64_t *c4) {
  int rc;
  if ((rc = check_ready(c, false))) return rc;
  if (nmodes < 1 || !modes || !N_pp || !N_pm || !N_mm) return fail(FP4_ERR_ARG, "bad arguments");
  Plan pl;
  if ((rc = subset_plan(c, nmodes, modes, pl))) return rc;
  if ((rc = count_pass(c, pl, true))) return rc;
  int64_t *N[3] = {N_pp, N_pm, N_mm};
  int64_t *outs[3][3] = {{a2, a3, a4}, {b2, b3, b4}, {c2, c3, c4}};
  const long per_ch = pl.items / 3;
  for (int ch = 0; ch < 3; ch++) {
    for (int m = 0; m < nmodes; m++) N[ch][m] = pl.segN[ch * nmodes + m];
    if (pl.off[(ch + 1) * per_ch] - pl.off[ch * per_ch] > cap)
      return fail(FP4_ERR_OVERFLOW, "channel %d needs %lld entries, cap is %lld", ch,
                  (long long)(pl.off[(ch + 1) * per_ch] - pl.off[ch * per_ch]), (long long)cap);
  }
  for (const Batch &b : plan_batches(pl, c->list_cap)) {
    if ((rc = emit_batch(c, pl, b, true))) return rc;
    for (int ch = 0; ch < 3; ch++) {
      const long lo = std::max(b.item_lo, ch * per_ch), hi = std::min(b.item_hi, (ch + 1) * per_ch);
      if (hi <= lo) continue;
      const long long p0 = pl.off[lo], p1 = pl.off[hi], base = pl.off[ch * per_ch];
      if (p1 <= p0) continue;
      for (int t = 0; t < 3; t++) {
        if (!outs[ch][t]) return fail(FP4_ERR_ARG, "null list output");
        HIPCHECK(hipMemcpyAsync(outs[ch][t] + (p0 - base), c->lists[t].as<long long>() + (p0 - b.p0),
                                sizeof(long long) * (p1 - p0), hipMemcpyDeviceToHost, c->stream));
      }
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
  }
  return FP4_OK;
}

// ---- sampling estimators (four phonons) -----------------------------------
// Segments are (channel, mode): one item each.
static int sample_check(const fp4_ctx *c, int32_t nsample) {
  (void)c;
  if (nsample < 1) return fail(FP4_ERR_ARG, "nsample must be positive (got %d)", nsample);
  return 0;
}

static long long sample_total(const fp4_ctx *c) {
  // Nbands*nptk*Nbands*nptk*Nbands in int64, processes.f90:2497, 4836
  return (long long)c->nb * c->nptk * c->nb * c->nptk * c->nb;
}

static SampleArgs sample_args(const fp4_ctx *c, const Plan &pl, int32_t nsample, uint64_t seed) {
  SampleArgs a{};
  a.m = mode_view(c);
  a.modes = c->modes_sub.as<int>();
  a.nsub = (int)pl.modes.size();
  a.nsample = nsample;
  a.seed = seed;
  return a;
}

// nch = 3: four-phonon channels (++, +-, --); nch = 2: three-phonon (+, -).
static int sample_launch(const SampleArgs &a, int nch, int mode, hipStream_t st) {
  return nch == 3 ? launch_sample4(a, mode, st) : launch_sample3(a, mode, st);
}

static int sample_count_pass(fp4_ctx *c, Plan &pl, int32_t nsample, uint64_t seed, int nch) {
  c->cnt_valid = false;   // cnt / psum are this pass's now
  const int nsub = (int)pl.modes.size();
  pl.ii_lo0 = 0;
  pl.nii = 1;
  pl.items = (long)nch * nsub;
  const int nseg = nch * nsub;
  int rc;
  if ((rc = c->modes_sub.ensure(sizeof(int) * nsub))) return rc;
  if ((rc = c->cnt.ensure(sizeof(int) * nseg))) return rc;
  if ((rc = c->psum.ensure(sizeof(double) * nseg))) return rc;
  if ((rc = c->segN.ensure(sizeof(long long) * nseg))) return rc;
  if ((rc = c->segP.ensure(sizeof(double) * nseg))) return rc;
  HIPCHECK(hipMemcpyAsync(c->modes_sub.p, pl.modes.data(), sizeof(int) * nsub, hipMemcpyHostToDevice, c->stream));
  SampleArgs a = sample_args(c, pl, nsample, seed);
  a.seg_lo = 0;
  a.seg_hi = nseg;
  a.cnt = c->cnt.as<int>();
  a.psum = c->psum.as<double>();
  if ((rc = timed(c, KS_SAMPLE, [&] {
         return sample_launch(a, nch, 0, c->stream) ? fail(FP4_ERR_UNSUPPORTED, "Nbands=%d", c->nb) : 0;
       })))
    return rc;
  return finish_counts(c, pl, nseg, true);
}

static int sample_emit_batch(fp4_ctx *c, const Plan &pl, const Batch &b, int32_t nsample, uint64_t seed, int nch) {
  int rc;
  const long long n = b.p1 - b.p0;
  for (int t = 0; t < nch; t++)
    if ((rc = c->lists[t].ensure(sizeof(int) * (size_t)std::max(n, 1LL)))) return rc;
  SampleArgs a = sample_args(c, pl, nsample, seed);
  a.seg_lo = (int)b.item_lo;
  a.seg_hi = (int)b.item_hi;
  a.offsets = c->offsets.as<long long>();
  a.out_base = b.p0;
  for (int t = 0; t < nch; t++) a.out32[t] = c->lists[t].as<int>();
  return timed(c, KS_SAMPLE,
               [&] { return sample_launch(a, nch, 1, c->stream) ? fail(FP4_ERR_UNSUPPORTED, "nb") : 0; });
}

int fp4_np_4ph_sample(fp4_ctx *c, int32_t nsample, uint64_t seed, int64_t *N_pp, double *P_pp, int64_t *N_pm,
                      double *P_pm, int64_t *N_mm, double *P_mm) {
  int rc;
  if ((rc = check_ready(c, false)) || (rc = sample_check(c, nsample))) return rc;
  if (!N_pp || !P_pp || !N_pm || !P_pm || !N_mm || !P_mm) return fail(FP4_ERR_ARG, "null output");
  Plan pl;
  all_modes(c, pl);
  if ((rc = sample_count_pass(c, pl, nsample, seed, 3))) return rc;
  const int nm = (int)pl.modes.size();
  const double total = (double)sample_total(c), num = (double)nsample;
  int64_t *N[3] = {N_pp, N_pm, N_mm};
  double *P[3] = {P_pp, P_pm, P_mm};
  for (int ch = 0; ch < 3; ch++)
    for (int m = 0; m < nm; m++) {
      // INT(REAL(N,8)/num*total, 8) and P*total/num, processes.f90:2562-2564
      N[ch][m] = (int64_t)((double)pl.segN[ch * nm + m] / num * total);
      P[ch][m] = pl.segP[ch * nm + m] * total / num;
    }
  return FP4_OK;
}

int fp4_rta_4ph_sample_modes(fp4_ctx *c, int32_t nsample, uint64_t seed, int32_t nmodes, const int32_t *modes,
                             double *rate_4ph, double *r_pp, double *r_pm, double *r_mm, double *w_pp, double *w_pm,
                             double *w_mm, int64_t *nhits) {
  int rc;
  if ((rc = check_ready(c, true)) || (rc = sample_check(c, nsample))) return rc;
  if (!rate_4ph || !r_pp || !r_pm || !r_mm || !w_pp || !w_pm || !w_mm) return fail(FP4_ERR_ARG, "null output");
  Plan pl;
  if ((rc = subset_plan(c, nmodes, modes, pl))) return rc;
  if ((rc = sample_count_pass(c, pl, nsample, seed, 3))) return rc;
  const int nseg = 3 * nmodes;
  std::vector<double> g(nseg, 0.0), w(nseg, 0.0);
  for (const Batch &b : plan_batches(pl, c->list_cap)) {
    if ((rc = sample_emit_batch(c, pl, b, nsample, seed, 3))) return rc;
    if ((rc = rate_batch(c, pl, b, false, false, g, w))) return rc;
  }
  write_rates(c, pl, g, w, rate_4ph, r_pp, r_pm, r_mm, w_pp, w_pm, w_mm, true);
  // Gamma/num*total and WP4*total/num, processes.f90:4922-4923
  const double total = (double)sample_total(c), num = (double)nsample;
  double *R[3] = {r_pp, r_pm, r_mm};
  double *W[3] = {w_pp, w_pm, w_mm};
  for (int mi = 0; mi < nmodes; mi++) {
    const int m = pl.modes[mi] - 1;
    for (int ch = 0; ch < 3; ch++) {
      R[ch][m] = R[ch][m] / num * total;
      W[ch][m] = W[ch][m] * total / num;
    }
    rate_4ph[m] = r_pp[m] + r_pm[m] + r_mm[m];
  }
  if (nhits) *nhits = std::accumulate(pl.segN.begin(), pl.segN.end(), 0LL);
  return FP4_OK;
}

int fp4_rta_4ph_sample(fp4_ctx *c, int32_t nsample, uint64_t seed, double *rate_4ph, double *r_pp, double *r_pm,
                       double *r_mm, double *w_pp, double *w_pm, double *w_mm) {
  int rc;
  if ((rc = check_ready(c, true))) return rc;
  std::vector<int> modes((size_t)c->nb * c->nlist);
  std::iota(modes.begin(), modes.end(), 1);
  const size_t nm = modes.size();
  double *outs[7] = {rate_4ph, r_pp, r_pm, r_mm, w_pp, w_pm, w_mm};
  for (double *o : outs) {
    if (!o) return fail(FP4_ERR_ARG, "null output");
    std::fill(o, o + nm, 0.0);
  }
  return fp4_rta_4ph_sample_modes(c, nsample, seed, (int)nm, modes.data(), rate_4ph, r_pp, r_pm, r_mm, w_pp, w_pm,
                                  w_mm, nullptr);
}

// ============================================================ three phonons
int fp4_ctx_set_ifc3(fp4_ctx *c, int32_t ntri, const double *Phi, const double *R_j, const double *R_k,
                     const int32_t *Ii, const int32_t *Ij, const int32_t *Ik) {
  if (!c || !Phi || !R_j || !R_k || !Ii || !Ij || !Ik) return fail(FP4_ERR_ARG, "null argument");
  if (ntri < 1) return fail(FP4_ERR_ARG, "Ntri must be positive");
  HIPCHECK(hipSetDevice(c->device));
  int rc;
  if ((rc = ensure_roots(c))) return rc;
  std::vector<Cluster3Rec> cl(ntri, Cluster3Rec{});
  for (int t = 0; t < ntri; t++) {
    Cluster3Rec &r = cl[t];
    const int32_t *ix[3] = {Ii, Ij, Ik};
    for (int a = 0; a < 3; a++) {
      if (ix[a][t] < 1 || ix[a][t] > c->natoms) return fail(FP4_ERR_ARG, "Index out of range at cluster %d", t + 1);
      r.atom[a] = ix[a][t] - 1;
    }
    if ((rc = phase_multipliers(c, R_j + 3 * (size_t)t, c->L, r.pj, t + 1))) return rc;
    if ((rc = phase_multipliers(c, R_k + 3 * (size_t)t, c->L, r.pk, t + 1))) return rc;
    // 1/sqrt(m_i*m_j*m_k), processes.f90:59-60
    double mp = c->masses[c->types[r.atom[0]] - 1] * c->masses[c->types[r.atom[1]] - 1];
    mp = mp * c->masses[c->types[r.atom[2]] - 1];
    r.minv = 1.0 / std::sqrt(mp);
  }
  if ((rc = c->clusters3.ensure(sizeof(Cluster3Rec) * ntri)) || (rc = c->phi3.ensure(sizeof(double) * 27 * ntri)))
    return rc;
  HIPCHECK(hipMemcpy(c->clusters3.p, cl.data(), sizeof(Cluster3Rec) * ntri, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->phi3.p, Phi, sizeof(double) * 27 * ntri, hipMemcpyHostToDevice));
  c->ntri3 = ntri;
  c->have_ifc3 = true;
  return FP4_OK;
}

static void all_modes3(const fp4_ctx *c, Plan &pl) {
  pl.modes.resize((size_t)c->nb * c->nlist);
  std::iota(pl.modes.begin(), pl.modes.end(), 1);
  pl.win_lo0 = 0;
  pl.win_n = c->nptk;
}

// counts of the reference's int32 kind; totals must fit the int32 list sizes
static int counts3_fit(const Plan &pl) {
  const int nsub = (int)pl.modes.size();
  for (int ch = 0; ch < 2; ch++) {
    long long tot = 0;
    for (int m = 0; m < nsub; m++) tot += pl.segN[ch * nsub + m];
    if (tot > INT32_MAX)
      return fail(FP4_ERR_OVERFLOW, "channel %d has %lld three-phonon processes, beyond the int32 lists", ch, tot);
  }
  return 0;
}

static int verify_counts3(const Plan &pl, const int32_t *N_plus, const int32_t *N_minus) {
  const int nm = (int)pl.modes.size();
  const int32_t *N[2] = {N_plus, N_minus};
  for (int ch = 0; ch < 2; ch++)
    for (int m = 0; m < nm; m++)
      if ((long long)N[ch][m] != pl.segN[ch * nm + m])
        return fail(FP4_ERR_OVERFLOW, "count mismatch: channel %d mode %d has %lld processes, caller passed %d", ch,
                    m + 1, (long long)pl.segN[ch * nm + m], (int)N[ch][m]);
  return 0;
}

int fp4_np_3ph(fp4_ctx *c, int32_t *N_plus, double *P_plus, int32_t *N_minus, double *P_minus) {
  int rc;
  if ((rc = check_ready3(c, false))) return rc;
  if (!N_plus || !P_plus || !N_minus || !P_minus) return fail(FP4_ERR_ARG, "null output");
  Plan pl;
  all_modes3(c, pl);
  if ((rc = count3_pass(c, pl, false))) return rc;
  if ((rc = counts3_fit(pl))) return rc;
  const int nm = (int)pl.modes.size();
  int32_t *N[2] = {N_plus, N_minus};
  double *P[2] = {P_plus, P_minus};
  for (int ch = 0; ch < 2; ch++)
    for (int m = 0; m < nm; m++) {
      N[ch][m] = (int32_t)pl.segN[ch * nm + m];
      P[ch][m] = pl.segP[ch * nm + m];
    }
  return FP4_OK;
}

// Copies each channel's part of a batch from c->lists (and c->proc_g) to the host.
static int copy3_batch(fp4_ctx *c, const Plan &pl, const Batch &b, int32_t *const outl[2][2], double *const outg[2]) {
  const long per_ch = pl.items / 2;
  for (int ch = 0; ch < 2; ch++) {
    const long lo = std::max(b.item_lo, ch * per_ch), hi = std::min(b.item_hi, (ch + 1) * per_ch);
    if (hi <= lo) continue;
    const long long p0 = pl.off[lo], p1 = pl.off[hi], base = pl.off[ch * per_ch];
    if (p1 <= p0) continue;
    for (int t = 0; t < 2; t++) {
      if (!outl[ch][t]) return fail(FP4_ERR_ARG, "null list output");
      HIPCHECK(hipMemcpyAsync(outl[ch][t] + (p0 - base), c->lists[t].as<int>() + (p0 - b.p0),
                              sizeof(int) * (p1 - p0), hipMemcpyDeviceToHost, c->stream));
    }
    if (outg) {
      if (!outg[ch]) return fail(FP4_ERR_ARG, "null Gamma output");
      HIPCHECK(hipMemcpyAsync(outg[ch] + (p0 - base), c->proc_g.as<double>() + (p0 - b.p0),
                              sizeof(double) * (p1 - p0), hipMemcpyDeviceToHost, c->stream));
    }
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int fp4_ind_only_3ph(fp4_ctx *c, const int32_t *N_plus, const int32_t *N_minus, int32_t *a2, int32_t *a3,
                     int32_t *b2, int32_t *b3) {
  int rc;
  if ((rc = check_ready3(c, false))) return rc;
  if (!N_plus || !N_minus) return fail(FP4_ERR_ARG, "null counts");
  Plan pl;
  all_modes3(c, pl);
  if ((rc = count3_pass(c, pl, true))) return rc;
  if ((rc = counts3_fit(pl)) || (rc = verify_counts3(pl, N_plus, N_minus))) return rc;
  int32_t *const outl[2][2] = {{a2, a3}, {b2, b3}};
  for (const Batch &b : plan_batches(pl, c->list_cap)) {
    if ((rc = emit3_batch(c, pl, b))) return rc;
    if ((rc = copy3_batch(c, pl, b, outl, nullptr))) return rc;
  }
  return FP4_OK;
}

int fp4_rta_3ph_modes(fp4_ctx *c, int32_t nmodes, const int32_t *modes, int32_t ii_lo, int32_t ii_hi, double *rate,
                      double *r_plus, double *r_minus, double *w_plus, double *w_minus, int64_t *nproc) {
  int rc;
  if ((rc = check_ready3(c, true))) return rc;
  if (!rate || !r_plus || !r_minus || !w_plus || !w_minus) return fail(FP4_ERR_ARG, "null output");
  if (ii_lo < 1 || ii_hi > c->nptk || ii_hi < ii_lo) return fail(FP4_ERR_ARG, "bad ii window %d..%d", ii_lo, ii_hi);
  Plan pl;
  if ((rc = subset_plan(c, nmodes, modes, pl))) return rc;
  pl.win_lo0 = ii_lo - 1;
  pl.win_n = ii_hi - ii_lo + 1;
  if ((rc = count3_pass(c, pl, true))) return rc;
  std::vector<double> g(2 * (size_t)nmodes, 0.0), w(2 * (size_t)nmodes, 0.0);
  for (const Batch &b : plan_batches(pl, c->list_cap)) {
    if ((rc = emit3_batch(c, pl, b))) return rc;
    if ((rc = rate3_batch(c, pl, b, false, g, w))) return rc;
  }
  write3_rates(c, pl, g, w, rate, r_plus, r_minus, w_plus, w_minus, true);
  if (nproc) *nproc = std::accumulate(pl.segN.begin(), pl.segN.end(), 0LL);
  return FP4_OK;
}

int fp4_rta_3ph(fp4_ctx *c, double *rate, double *r_plus, double *r_minus, double *w_plus, double *w_minus) {
  int rc;
  if ((rc = check_ready3(c, true))) return rc;
  const size_t nm = (size_t)c->nb * c->nlist;
  double *outs[5] = {rate, r_plus, r_minus, w_plus, w_minus};
  for (double *o : outs) {
    if (!o) return fail(FP4_ERR_ARG, "null output");
    std::fill(o, o + nm, 0.0);
  }
  std::vector<int> modes(nm);
  std::iota(modes.begin(), modes.end(), 1);
  return fp4_rta_3ph_modes(c, (int)nm, modes.data(), 1, c->nptk, rate, r_plus, r_minus, w_plus, w_minus, nullptr);
}

int fp4_rta_3ph_using_ind(fp4_ctx *c, const int32_t *N_plus, const int32_t *N_minus, const int32_t *a2,
                          const int32_t *b2, const int32_t *a3, const int32_t *b3, double *rate, double *r_plus,
                          double *r_minus, double *w_plus, double *w_minus) {
  int rc;
  if ((rc = check_ready3(c, true))) return rc;
  const int nm = c->nb * c->nlist;
  const int32_t *N[2] = {N_plus, N_minus};
  const int32_t *L2[2][2] = {{a2, a3}, {b2, b3}};
  double *outs[5] = {rate, r_plus, r_minus, w_plus, w_minus};
  for (double *o : outs) {
    if (!o) return fail(FP4_ERR_ARG, "null output");
    std::fill(o, o + nm, 0.0);
  }
  // one item per (channel, mode); offsets from the given counts
  Plan pl;
  pl.modes.resize(nm);
  std::iota(pl.modes.begin(), pl.modes.end(), 1);
  pl.nii = 1;
  pl.items = 2L * nm;
  pl.off.assign(pl.items + 1, 0);
  for (int ch = 0; ch < 2; ch++)
    for (int m = 0; m < nm; m++) {
      if (!N[ch] || N[ch][m] < 0) return fail(FP4_ERR_ARG, "bad counts");
      pl.off[(long)ch * nm + m + 1] = pl.off[(long)ch * nm + m] + N[ch][m];
    }
  const long long maxv = (long long)c->nptk * c->nb;
  std::vector<double> g(2 * (size_t)nm, 0.0), w(2 * (size_t)nm, 0.0);
  for (const Batch &b : plan_batches(pl, c->list_cap)) {
    const long long n = b.p1 - b.p0;
    for (int t = 0; t < 2; t++)
      if ((rc = c->lists[t].ensure(4 * (size_t)n))) return rc;
    for (long it = b.item_lo; it < b.item_hi; it++) {
      const int ch = (int)(it / nm);
      const long long p0 = pl.off[it], p1 = pl.off[it + 1];
      if (p1 <= p0) continue;
      const long long base = pl.off[(long)ch * nm];
      for (int t = 0; t < 2; t++) {
        if (!L2[ch][t]) return fail(FP4_ERR_ARG, "null list");
        const int32_t *src = L2[ch][t] + (p0 - base);
        for (long long x = 0; x < p1 - p0; x++)
          if (src[x] < 1 || src[x] > maxv) return fail(FP4_ERR_ARG, "list entry out of range");
        HIPCHECK(hipMemcpyAsync(c->lists[t].as<int>() + (p0 - b.p0), src, sizeof(int) * (p1 - p0),
                                hipMemcpyHostToDevice, c->stream));
      }
    }
    if ((rc = rate3_batch(c, pl, b, false, g, w))) return rc;
  }
  write3_rates(c, pl, g, w, rate, r_plus, r_minus, w_plus, w_minus, true);
  return FP4_OK;
}

int fp4_ind_3ph(fp4_ctx *c, const int32_t *N_plus, const int32_t *N_minus, int32_t *a2, int32_t *a3, double *G_plus,
                int32_t *b2, int32_t *b3, double *G_minus, double *rate, double *r_plus, double *r_minus,
                double *w_plus, double *w_minus) {
  int rc;
  if ((rc = check_ready3(c, true))) return rc;
  if (!N_plus || !N_minus) return fail(FP4_ERR_ARG, "null counts");
  const size_t nm = (size_t)c->nb * c->nlist;
  double *outs[5] = {rate, r_plus, r_minus, w_plus, w_minus};
  for (double *o : outs) {
    if (!o) return fail(FP4_ERR_ARG, "null output");
    std::fill(o, o + nm, 0.0);
  }
  Plan pl;
  all_modes3(c, pl);
  if ((rc = count3_pass(c, pl, true))) return rc;
  if ((rc = counts3_fit(pl)) || (rc = verify_counts3(pl, N_plus, N_minus))) return rc;
  int32_t *const outl[2][2] = {{a2, a3}, {b2, b3}};
  double *const outg[2] = {G_plus, G_minus};
  std::vector<double> g(2 * nm, 0.0), w(2 * nm, 0.0);
  for (const Batch &b : plan_batches(pl, c->list_cap)) {
    if ((rc = emit3_batch(c, pl, b))) return rc;
    if ((rc = rate3_batch(c, pl, b, true, g, w))) return rc;
    if ((rc = copy3_batch(c, pl, b, outl, outg))) return rc;
  }
  // per-process Gamma are already in THz; Ind_driver sums them per mode
  write3_rates(c, pl, g, w, rate, r_plus, r_minus, w_plus, w_minus, false);
  return FP4_OK;
}

int fp4_vp_3ph(fp4_ctx *c, int32_t channel, int64_t n, const int32_t *idx, double *out) {
  int rc;
  if ((rc = check_ready3(c, true))) return rc;
  if (channel < 0 || channel > 1) return fail(FP4_ERR_ARG, "channel must be 0 or 1");
  if (n < 0 || (n > 0 && (!idx || !out))) return fail(FP4_ERR_ARG, "bad arguments");
  if (n == 0) return FP4_OK;
  for (int64_t t = 0; t < n; t++) {
    const int32_t *r = idx + 6 * t;
    for (int a = 0; a < 3; a++)
      if (r[a] < 1 || r[a] > c->nb) return fail(FP4_ERR_ARG, "branch index out of range in row %lld", (long long)t);
    for (int a = 3; a < 6; a++)
      if (r[a] < 1 || r[a] > c->nptk) return fail(FP4_ERR_ARG, "q index out of range in row %lld", (long long)t);
  }
  DevBuf di, dout;
  if ((rc = di.ensure(sizeof(int) * 6 * n)) || (rc = dout.ensure(sizeof(double) * 2 * n))) {
    di.release();
    return rc;
  }
  Vp3Args a{};
  a.m = mode_view(c);
  a.f = ifc3_view(c);
  a.ch = channel;
  a.n = n;
  a.idx = di.as<int>();
  a.out = dout.as<double2>();
  hipError_t e = hipMemcpyAsync(di.p, idx, sizeof(int) * 6 * n, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    launch_vp3(a, c->stream);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout.p, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  di.release();
  dout.release();
  if (e != hipSuccess) return fail(FP4_ERR_DEVICE, "fp4_vp_3ph: %s", hipGetErrorString(e));
  return FP4_OK;
}

// ---- three-phonon sampling estimators -------------------------------------
// Scaling in the reference's left-to-right
// order: N = INT(REAL(hits)/num*Nb*nptk*Nb, 4), P*Nb*nptk*Nb/num
// (processes.f90:898-899); Gamma*Nb*nptk*Nb/num then *5.60626442*1e8/nptk,
// WP3/nptk (WP3*0.5/nptk for "-") then *Nb*nptk*Nb/num (:1559-1570,
// :1660-1671); the "-" rate halved in the driver (:1759).
int fp4_np_3ph_sample(fp4_ctx *c, int32_t nsample, uint64_t seed, int32_t *N_plus, double *P_plus, int32_t *N_minus,
                      double *P_minus) {
  int rc;
  if ((rc = check_ready3(c, false)) || (rc = sample_check(c, nsample))) return rc;
  if (!N_plus || !P_plus || !N_minus || !P_minus) return fail(FP4_ERR_ARG, "null output");
  Plan pl;
  all_modes3(c, pl);
  if ((rc = sample_count_pass(c, pl, nsample, seed, 2))) return rc;
  const int nm = (int)pl.modes.size(), nb = c->nb, nptk = c->nptk;
  const double num = (double)nsample;
  int32_t *N[2] = {N_plus, N_minus};
  double *P[2] = {P_plus, P_minus};
  for (int ch = 0; ch < 2; ch++)
    for (int m = 0; m < nm; m++) {
      const double n = (double)pl.segN[ch * nm + m] / num * nb * nptk * nb;
      if (n > 2147483647.0)
        return fail(FP4_ERR_OVERFLOW, "channel %d mode %d: count estimate %.3e exceeds int32", ch, m + 1, n);
      N[ch][m] = (int32_t)n;
      P[ch][m] = pl.segP[ch * nm + m] * nb * nptk * nb / num;
    }
  return FP4_OK;
}

int fp4_rta_3ph_sample_modes(fp4_ctx *c, int32_t nsample, uint64_t seed, int32_t nmodes, const int32_t *modes,
                             double *rate, double *r_plus, double *r_minus, double *w_plus, double *w_minus,
                             int64_t *nhits) {
  int rc;
  if ((rc = check_ready3(c, true)) || (rc = sample_check(c, nsample))) return rc;
  if (!rate || !r_plus || !r_minus || !w_plus || !w_minus) return fail(FP4_ERR_ARG, "null output");
  Plan pl;
  if ((rc = subset_plan(c, nmodes, modes, pl))) return rc;
  if ((rc = sample_count_pass(c, pl, nsample, seed, 2))) return rc;
  std::vector<double> g(2 * (size_t)nmodes, 0.0), w(2 * (size_t)nmodes, 0.0);
  for (const Batch &b : plan_batches(pl, c->list_cap)) {
    if ((rc = sample_emit_batch(c, pl, b, nsample, seed, 2))) return rc;
    if ((rc = rate3_batch(c, pl, b, false, g, w))) return rc;
  }
  const int nb = c->nb, nptk = c->nptk;
  const double num = (double)nsample;
  double *R[2] = {r_plus, r_minus};
  double *W[2] = {w_plus, w_minus};
  for (int mi = 0; mi < nmodes; mi++) {
    const int m = pl.modes[mi] - 1;
    for (int ch = 0; ch < 2; ch++) {
      double gs = g[ch * nmodes + mi], ws = w[ch * nmodes + mi];
      ws = (ch == 0) ? ws / nptk : ws * 5.e-1 / nptk;
      gs = gs * nb * nptk * nb / num;
      ws = ws * nb * nptk * nb / num;
      gs = gs * (double)kGamma3Units * 1.e8 / nptk;
      R[ch][m] = (ch == 0) ? gs : gs * 5.e-1;
      W[ch][m] = ws;
    }
    rate[m] = r_plus[m] + r_minus[m];
  }
  if (nhits) *nhits = std::accumulate(pl.segN.begin(), pl.segN.end(), 0LL);
  return FP4_OK;
}

int fp4_rta_3ph_sample(fp4_ctx *c, int32_t nsample, uint64_t seed, double *rate, double *r_plus, double *r_minus,
                       double *w_plus, double *w_minus) {
  int rc;
  if ((rc = check_ready3(c, true))) return rc;
  std::vector<int> modes((size_t)c->nb * c->nlist);
  std::iota(modes.begin(), modes.end(), 1);
  const size_t nm = modes.size();
  double *outs[5] = {rate, r_plus, r_minus, w_plus, w_minus};
  for (double *o : outs) {
    if (!o) return fail(FP4_ERR_ARG, "null output");
    std::fill(o, o + nm, 0.0);
  }
  return fp4_rta_3ph_sample_modes(c, nsample, seed, (int)nm, modes.data(), rate, r_plus, r_minus, w_plus, w_minus,
                                  nullptr);
}

// ============================================================ harmonic front end
// Image list of phonon_espresso (Src/phonon_routines.f90:440-537): for each
// atom pair, the supercell images t (bohr) inside the Wigner-Seitz cell of
// the supercell with their weights 1/nreq, and the constants of each image.
namespace {
struct HarmonicImages {
  std::vector<int> start;
  std::vector<double> t, fc;
};

int espresso_images(const fp4_harmonic *h, const double *fc_in, HarmonicImages &out, std::vector<double> &mm) {
  const int nat = h->natoms, s1 = h->scell[0], s2 = h->scell[1], s3 = h->scell[2];
  const size_t ncell = (size_t)s1 * s2 * s3;
  const float f_bohr2nm = 0.052917721092f, f_mf1 = 1.8218779f, f_mf2 = 6.022e-4f;
  const double bohr2nm = (double)f_bohr2nm, massfactor = (double)(f_mf1 * f_mf2), tol = (double)1e-6f;
  // fc_s(ipol,jpol,iat,jat,t1,t2,t3), Fortran order
  auto at = [&](int a, int b, int i, int j, size_t cell) {
    return (size_t)a + 3 * ((size_t)b + 3 * ((size_t)i + nat * ((size_t)j + nat * cell)));
  };
  std::vector<double> fc(fc_in, fc_in + 9 * (size_t)nat * nat * ncell);
  // momentum conservation on the on-site block, Src/phonon_routines.f90:416-423
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      for (int i = 0; i < nat; i++) {
        double s = 0.0;
        for (size_t cell = 0; cell < ncell; cell++)
          for (int j = 0; j < nat; j++) s = s + fc[at(a, b, i, j, cell)];
        fc[at(a, b, i, i, 0)] = fc[at(a, b, i, i, 0)] - s;
      }
  double cell_r[3][3], r[64][3];   // rows a_i in bohr; atom positions in bohr
  for (int i = 0; i < 3; i++)
    for (int c = 0; c < 3; c++) cell_r[i][c] = h->lattvec[c + 3 * i] / bohr2nm;
  for (int i = 0; i < nat; i++)
    for (int c = 0; c < 3; c++) {
      double x = 0.0;
      for (int k = 0; k < 3; k++) x = x + h->lattvec[c + 3 * k] * h->positions[k + 3 * i];
      r[i][c] = x / bohr2nm;
    }
  double rws[124][4];
  int nr = 0;
  for (int m1 = -2; m1 <= 2; m1++)
    for (int m2 = -2; m2 <= 2; m2++)
      for (int m3 = -2; m3 <= 2; m3++) {
        if (!m1 && !m2 && !m3) continue;
        double d = 0.0;
        for (int c = 0; c < 3; c++) {
          rws[nr][c + 1] = cell_r[0][c] * s1 * m1 + cell_r[1][c] * s2 * m2 + cell_r[2][c] * s3 * m3;
          d = d + rws[nr][c + 1] * rws[nr][c + 1];
        }
        rws[nr][0] = 0.5 * d;
        nr++;
      }
  out.start.assign(1, 0);
  out.t.clear();
  out.fc.clear();
  for (int i = 0; i < nat; i++)
    for (int j = 0; j < nat; j++) {
      for (int m1 = -2 * s1; m1 <= 2 * s1; m1++)
        for (int m2 = -2 * s2; m2 <= 2 * s2; m2++)
          for (int m3 = -2 * s3; m3 <= 2 * s3; m3++) {
            double t[3], rw[3];
            for (int c = 0; c < 3; c++) {
              t[c] = m1 * cell_r[0][c] + m2 * cell_r[1][c] + m3 * cell_r[2][c];
              rw[c] = t[c] + (r[i][c] - r[j][c]);
            }
            int nreq = 1;
            bool outside = false;
            for (int k = 0; k < nr; k++) {
              const double ck = rw[0] * rws[k][1] + rw[1] * rws[k][2] + rw[2] * rws[k][3] - rws[k][0];
              if (ck > tol) outside = true;
              else if (std::fabs(ck) < tol) nreq++;
            }
            if (outside) continue;
            const double w = 1.0 / (double)nreq;
            const size_t cell = (size_t)(((m1 % s1) + s1) % s1) +
                                (size_t)s1 * ((size_t)(((m2 % s2) + s2) % s2) + (size_t)s2 * (((m3 % s3) + s3) % s3));
            for (int c = 0; c < 3; c++) out.t.push_back(t[c]);
            for (int a = 0; a < 3; a++)
              for (int b = 0; b < 3; b++) out.fc.push_back(fc[at(a, b, i, j, cell)] * w);
          }
      out.start.push_back((int)(out.t.size() / 3));
    }
  mm.assign((size_t)nat * nat, 0.0);
  for (int i = 0; i < nat; i++)
    for (int j = 0; j < nat; j++) {
      const double mi = h->masses[h->types[i] - 1] / massfactor, mj = h->masses[h->types[j] - 1] / massfactor;
      mm[(size_t)i * nat + j] = (i == j) ? mi : std::sqrt(mi * mj);
    }
  return 0;
}

// rlattvec = 2 pi inv(lattvec)^T, columns b_i; returns det(lattvec)
double reciprocal(const double *A, double *rl) {
  const double det = A[0] * (A[4] * A[8] - A[7] * A[5]) - A[3] * (A[1] * A[8] - A[7] * A[2]) +
                     A[6] * (A[1] * A[5] - A[4] * A[2]);
  if (!(std::fabs(det) > 0.0)) return 0.0;
  const double cof[9] = {A[4] * A[8] - A[5] * A[7], A[5] * A[6] - A[3] * A[8], A[3] * A[7] - A[4] * A[6],
                         A[2] * A[7] - A[1] * A[8], A[0] * A[8] - A[2] * A[6], A[1] * A[6] - A[0] * A[7],
                         A[1] * A[5] - A[2] * A[4], A[2] * A[3] - A[0] * A[5], A[0] * A[4] - A[1] * A[3]};
  // b_i . a_j = 2 pi delta_ij: b_i = 2 pi (a_j x a_k) / det
  for (int i = 0; i < 3; i++)
    for (int c = 0; c < 3; c++) rl[c + 3 * i] = 2.0 * M_PI * cof[c + 3 * i] / det;
  return det;
}

int check_harmonic(const fp4_harmonic *h) {
  if (!h || !h->positions || !h->masses || !h->types) return fail(FP4_ERR_ARG, "null argument");
  const int nat = h->natoms, nb = 3 * nat;
  if (nat < 1 || nat > 64) return fail(FP4_ERR_ARG, "natoms out of range");
  if (!nb_supported(nb)) return fail(FP4_ERR_UNSUPPORTED, "Nbands=%d not built", nb);
  for (int c = 0; c < 3; c++)
    if (h->ngrid[c] < 1 || h->scell[c] < 1) return fail(FP4_ERR_ARG, "bad ngrid or scell");
  for (int i = 0; i < nat; i++)
    if (h->types[i] < 1 || h->types[i] > h->nelements) return fail(FP4_ERR_ARG, "types out of range");
  return 0;
}

// phonon_phonopy's images (Src/phonon_routines.f90:210-240): per atom pair
// (row atom a1, column atom a2) and supercell cell, the supercell images of
// r = r_a1 - r_a2 + R_cell nearest to the origin, ties within 1e-5 nm kept in
// the loop order, each with weight 1/neq and the block
// Phi_ij = fc(a2, j, cell, a1, i) of read2fc's reduced constants.
void phonopy_images(const fp4_harmonic *h, const double *fc, HarmonicImages &out, std::vector<double> &w) {
  const int nat = h->natoms, s1 = h->scell[0], s2 = h->scell[1], s3 = h->scell[2];
  const double *L = h->lattvec;
  auto at = [&](int a, int i, int x, int y, int z, int b, int j) -> size_t {
    return (size_t)a + (size_t)nat * (i + 3 * (x + (size_t)s1 * (y + (size_t)s2 * (z + (size_t)s3 * (b + (size_t)nat * j)))));
  };
  std::vector<double> cart((size_t)3 * nat);
  for (int i = 0; i < nat; i++)
    for (int c = 0; c < 3; c++) {
      double x = 0.0;
      for (int k = 0; k < 3; k++) x = x + L[c + 3 * k] * h->positions[k + 3 * i];
      cart[3 * i + c] = x;
    }
  out.start.assign(1, 0);
  out.t.clear();
  out.fc.clear();
  w.clear();
  for (int a1 = 0; a1 < nat; a1++)
    for (int a2 = 0; a2 < nat; a2++) {
      for (int ix = 0; ix < s1; ix++)
        for (int iy = 0; iy < s2; iy++)
          for (int iz = 0; iz < s3; iz++) {
            double rcell[3], r[3];
            for (int c = 0; c < 3; c++) {
              rcell[c] = L[c] * ix + L[c + 3] * iy + L[c + 6] * iz;
              r[c] = cart[3 * a1 + c] - cart[3 * a2 + c] + rcell[c];
            }
            double dmin = std::numeric_limits<double>::max();
            std::vector<std::array<double, 3>> rr;
            for (int x = -2; x <= 2; x++)
              for (int y = -2; y <= 2; y++)
                for (int z = -2; z <= 2; z++) {
                  double rl[3], n2 = 0.0;
                  for (int c = 0; c < 3; c++) {
                    rl[c] = x * s1 * L[c] + y * s2 * L[c + 3] + z * s3 * L[c + 6];
                    n2 = n2 + (rl[c] + r[c]) * (rl[c] + r[c]);
                  }
                  const double rn = std::sqrt(n2);
                  const std::array<double, 3> v{rl[0] + rcell[0], rl[1] + rcell[1], rl[2] + rcell[2]};
                  if (std::fabs(rn - dmin) > 1e-5) {
                    if (rn < dmin) {
                      dmin = rn;
                      rr.assign(1, v);
                    }
                  } else {
                    rr.push_back(v);
                  }
                }
            const double wt = 1.0 / (double)rr.size();
            for (const auto &v : rr) {
              for (int c = 0; c < 3; c++) out.t.push_back(v[c]);
              for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) out.fc.push_back(fc[at(a2, j, ix, iy, iz, a1, i)] * wt);
              w.push_back(wt);
            }
          }
      out.start.push_back((int)(out.t.size() / 3));
    }
}
}  // namespace

int fp4_eigen_dm_espresso(const fp4_harmonic *h, const double *fc, const double *symmetrizers, double *omega,
                          double *eigenvect, double *velocity) {
  if (!h || !fc || !omega || !eigenvect || !velocity || !h->positions || !h->masses || !h->types)
    return fail(FP4_ERR_ARG, "null argument");
  const int nat = h->natoms, nb = 3 * nat;
  if (nat < 1 || nat > 64) return fail(FP4_ERR_ARG, "natoms out of range");
  if (!nb_supported(nb)) return fail(FP4_ERR_UNSUPPORTED, "Nbands=%d not built", nb);
  for (int c = 0; c < 3; c++)
    if (h->ngrid[c] < 1 || h->scell[c] < 1) return fail(FP4_ERR_ARG, "bad ngrid or scell");
  for (int i = 0; i < nat; i++)
    if (h->types[i] < 1 || h->types[i] > h->nelements) return fail(FP4_ERR_ARG, "types out of range");
  const int nptk = h->ngrid[0] * h->ngrid[1] * h->ngrid[2];
  HarmonicImages im;
  std::vector<double> mm;
  espresso_images(h, fc, im, mm);
  double rl[9];
  if (!(std::fabs(reciprocal(h->lattvec, rl)) > 0.0)) return fail(FP4_ERR_ARG, "singular lattice");
  HIPCHECK(hipSetDevice(h->device));
  DevBuf d_start, d_t, d_fc, d_mm, d_symm, d_om, d_eig, d_vel;
  int rc;
  if ((rc = d_start.ensure(sizeof(int) * im.start.size())) || (rc = d_t.ensure(sizeof(double) * im.t.size())) ||
      (rc = d_fc.ensure(sizeof(double) * im.fc.size())) || (rc = d_mm.ensure(sizeof(double) * mm.size())) ||
      (rc = d_om.ensure(sizeof(double) * (size_t)nptk * nb)) ||
      (rc = d_eig.ensure(sizeof(double2) * (size_t)nptk * nb * nb)) ||
      (rc = d_vel.ensure(sizeof(double) * (size_t)nptk * nb * 3)) ||
      (symmetrizers && (rc = d_symm.ensure(sizeof(double) * 9 * (size_t)nptk))))
    return rc;
  HIPCHECK(hipMemcpy(d_start.p, im.start.data(), sizeof(int) * im.start.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(d_t.p, im.t.data(), sizeof(double) * im.t.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(d_fc.p, im.fc.data(), sizeof(double) * im.fc.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(d_mm.p, mm.data(), sizeof(double) * mm.size(), hipMemcpyHostToDevice));
  if (symmetrizers)
    HIPCHECK(hipMemcpy(d_symm.p, symmetrizers, sizeof(double) * 9 * (size_t)nptk, hipMemcpyHostToDevice));
  EigenDMArgs a{};
  a.nptk = nptk;
  a.nb = nb;
  a.nat = nat;
  a.n1 = h->ngrid[0];
  a.n2 = h->ngrid[1];
  a.n3 = h->ngrid[2];
  for (int t = 0; t < 9; t++) a.b[t] = rl[t];
  a.bohr2nm = (double)0.052917721092f;
  a.tothz = (double)20670.687f;
  a.img_start = d_start.as<int>();
  a.img_t = d_t.as<double>();
  a.img_fc = d_fc.as<double>();
  a.mm = d_mm.as<double>();
  a.symm = symmetrizers ? d_symm.as<double>() : nullptr;
  a.omega = d_om.as<double>();
  a.eig = d_eig.as<double2>();
  a.vel = d_vel.as<double>();
  if (launch_eigen_dm(a, nullptr)) return fail(FP4_ERR_UNSUPPORTED, "Nbands=%d", nb);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpy(omega, d_om.p, sizeof(double) * (size_t)nptk * nb, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(eigenvect, d_eig.p, sizeof(double2) * (size_t)nptk * nb * nb, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(velocity, d_vel.p, sizeof(double) * (size_t)nptk * nb * 3, hipMemcpyDeviceToHost));
  return FP4_OK;
}

int fp4_eigen_dm_phonopy(const fp4_harmonic *h, const double *fc, const double *born, const double *epsilon,
                         const double *nac_images, const double *symmetrizers, double *omega, double *eigenvect,
                         double *velocity) {
  int rc;
  if ((rc = check_harmonic(h))) return rc;
  if (!fc || !omega || !eigenvect || !velocity) return fail(FP4_ERR_ARG, "null argument");
  if (!born != !epsilon) return fail(FP4_ERR_ARG, "born and epsilon go together");
  const int nat = h->natoms, nb = 3 * nat;
  const int nptk = h->ngrid[0] * h->ngrid[1] * h->ngrid[2];
  double rl[9];
  const double det = reciprocal(h->lattvec, rl);
  if (!(std::fabs(det) > 0.0)) return fail(FP4_ERR_ARG, "singular lattice");
  HarmonicImages im;
  std::vector<double> wimg;
  phonopy_images(h, fc, im, wimg);
  std::vector<double> ones((size_t)nat * nat, 1.0), nmm((size_t)nat * nat);
  for (int i = 0; i < nat; i++)
    for (int j = 0; j < nat; j++) {
      const double mi = h->masses[h->types[i] - 1], mj = h->masses[h->types[j] - 1];
      nmm[(size_t)i * nat + j] = (i == j) ? mi : std::sqrt(mi * mj);
    }
  // the 1st-BZ image of each q (Src/phonon_routines.f90:143-159): q itself
  // unless an image in x, y, z order is strictly shorter
  std::vector<double> qimg;
  if (born) {
    if (nac_images) {
      qimg.assign(nac_images, nac_images + 3 * (size_t)nptk);
    } else {
      qimg.resize(3 * (size_t)nptk);
      for (int q0 = 0; q0 < nptk; q0++) {
        const int i1 = q0 % h->ngrid[0], i2 = (q0 / h->ngrid[0]) % h->ngrid[1], i3 = q0 / (h->ngrid[0] * h->ngrid[1]);
        double k[3];
        for (int c = 0; c < 3; c++)
          k[c] = ((rl[c] * (double)i1) / (double)h->ngrid[0] + (rl[3 + c] * (double)i2) / (double)h->ngrid[1]) +
                 (rl[6 + c] * (double)i3) / (double)h->ngrid[2];
        double best = std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]), sh[3] = {k[0], k[1], k[2]};
        for (int x = -2; x <= 2; x++)
          for (int y = -2; y <= 2; y++)
            for (int z = -2; z <= 2; z++) {
              double r[3];
              for (int c = 0; c < 3; c++) r[c] = k[c] + x * rl[c] + y * rl[3 + c] + z * rl[6 + c];
              const double d = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
              if (d < best) {
                best = d;
                for (int c = 0; c < 3; c++) sh[c] = r[c];
              }
            }
        for (int c = 0; c < 3; c++) qimg[3 * (size_t)q0 + c] = sh[c];
      }
    }
  }
  HIPCHECK(hipSetDevice(h->device));
  DevBuf d_start, d_t, d_fc, d_w, d_mm, d_nmm, d_q, d_born, d_symm, d_om, d_eig, d_vel;
  if ((rc = d_start.ensure(sizeof(int) * im.start.size())) || (rc = d_t.ensure(sizeof(double) * im.t.size())) ||
      (rc = d_fc.ensure(sizeof(double) * im.fc.size())) || (rc = d_w.ensure(sizeof(double) * wimg.size())) ||
      (rc = d_mm.ensure(sizeof(double) * ones.size())) || (rc = d_nmm.ensure(sizeof(double) * nmm.size())) ||
      (rc = d_om.ensure(sizeof(double) * (size_t)nptk * nb)) ||
      (rc = d_eig.ensure(sizeof(double2) * (size_t)nptk * nb * nb)) ||
      (rc = d_vel.ensure(sizeof(double) * (size_t)nptk * nb * 3)) ||
      (born && ((rc = d_q.ensure(sizeof(double) * qimg.size())) || (rc = d_born.ensure(sizeof(double) * 9 * nat)))) ||
      (symmetrizers && (rc = d_symm.ensure(sizeof(double) * 9 * (size_t)nptk))))
    return rc;
  HIPCHECK(hipMemcpy(d_start.p, im.start.data(), sizeof(int) * im.start.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(d_t.p, im.t.data(), sizeof(double) * im.t.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(d_fc.p, im.fc.data(), sizeof(double) * im.fc.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(d_w.p, wimg.data(), sizeof(double) * wimg.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(d_mm.p, ones.data(), sizeof(double) * ones.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(d_nmm.p, nmm.data(), sizeof(double) * nmm.size(), hipMemcpyHostToDevice));
  if (born) {
    HIPCHECK(hipMemcpy(d_q.p, qimg.data(), sizeof(double) * qimg.size(), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(d_born.p, born, sizeof(double) * 9 * nat, hipMemcpyHostToDevice));
  }
  if (symmetrizers)
    HIPCHECK(hipMemcpy(d_symm.p, symmetrizers, sizeof(double) * 9 * (size_t)nptk, hipMemcpyHostToDevice));
  EigenDMArgs a{};
  a.nptk = nptk;
  a.nb = nb;
  a.nat = nat;
  a.n1 = h->ngrid[0];
  a.n2 = h->ngrid[1];
  a.n3 = h->ngrid[2];
  for (int t = 0; t < 9; t++) a.b[t] = rl[t];
  // constants already in (rad/ps)^2 per amu-reduced unit, lengths in nm
  a.bohr2nm = 1.0;
  a.tothz = 1.0;
  a.img_start = d_start.as<int>();
  a.img_t = d_t.as<double>();
  a.img_fc = d_fc.as<double>();
  a.mm = d_mm.as<double>();
  a.symm = symmetrizers ? d_symm.as<double>() : nullptr;
  a.omega = d_om.as<double>();
  a.eig = d_eig.as<double2>();
  a.vel = d_vel.as<double>();
  if (born) {
    a.img_w = d_w.as<double>();
    a.nac_q = d_q.as<double>();
    a.nac_mm = d_nmm.as<double>();
    a.born = d_born.as<double>();
    for (int t = 0; t < 9; t++) a.eps[t] = epsilon[t];
    // prefactor 1745.91429109 THz^2 amu nm^3 is a default-real literal (Src/phonon_routines.f90:91)
    a.nac_scale = (double)1745.91429109f / std::fabs(det) / ((double)h->scell[0] * h->scell[1] * h->scell[2]);
  }
  if (launch_eigen_dm(a, nullptr)) return fail(FP4_ERR_UNSUPPORTED, "Nbands=%d", nb);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpy(omega, d_om.p, sizeof(double) * (size_t)nptk * nb, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(eigenvect, d_eig.p, sizeof(double2) * (size_t)nptk * nb * nb, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(velocity, d_vel.p, sizeof(double) * (size_t)nptk * nb * 3, hipMemcpyDeviceToHost));
  return FP4_OK;
}

// ---- iterative solution ---------------------------------------------------
static BteModes bte_modes(const fp4_ctx *c) {
  BteModes m;
  m.omega = c->energy.as<double>();
  m.vel = c->vel.as<double>();
  m.nptk = c->nptk;
  m.nb = c->nb;
  m.nlist = c->nlist;
  m.stream = c->stream;
  return m;
}

static int bte_ready(fp4_ctx *c) {
  int rc;
  if ((rc = check_ready(c, false))) return rc;
  if (!c->bte) {
    c->bte = bte_create();
    if (!c->bte) return fail(FP4_ERR_DEVICE, "event creation failed");
  }
  return 0;
}

#define BTE_CALL(x)                                      \
  do {                                                   \
    const int r_ = (x);                                  \
    if (r_) return fail(r_, "%s", bte_error());          \
  } while (0)

int fp4_ctx_set_symmetry(fp4_ctx *c, int32_t nsymm_rot, const int32_t *ID_equi, const double *symmetrizers,
                         const double *crotations) {
  int rc;
  if ((rc = bte_ready(c))) return rc;
  BTE_CALL(bte_set_symmetry(c->bte, bte_modes(c), nsymm_rot, ID_equi, symmetrizers, crotations));
  return FP4_OK;
}

int fp4_iteration0(fp4_ctx *c, int32_t nlist, const int32_t *Nequi, const int32_t *ALLEquiList,
                   const double *tau_zero, double *F_n) {
  int rc;
  if ((rc = bte_ready(c))) return rc;
  if (nlist != c->nlist) return fail(FP4_ERR_ARG, "Nlist=%d differs from the modes' %d", nlist, c->nlist);
  const BteModes m = bte_modes(c);
  BTE_CALL(bte_set_orbits(c->bte, m, Nequi, ALLEquiList, nullptr));
  BTE_CALL(bte_iteration0(c->bte, m, tau_zero));
  BTE_CALL(bte_get_F(c->bte, m, F_n));
  return FP4_OK;
}

int fp4_bte_set_operator(fp4_ctx *c, int32_t nlist, const int32_t *Nequi, const int32_t *ALLEquiList,
                         const int32_t *TypeofSymmetry, const int32_t *N_plus, const int32_t *N_minus,
                         int32_t Ntotal_plus, int32_t Ntotal_minus, const int64_t *N_pp, const int64_t *N_pm,
                         const int64_t *N_mm, int64_t Ntotal_pp, int64_t Ntotal_pm, int64_t Ntotal_mm,
                         const int32_t *p2, const int32_t *p3, const int32_t *m2, const int32_t *m3,
                         const int64_t *a2, const int64_t *a3, const int64_t *a4, const int64_t *b2,
                         const int64_t *b3, const int64_t *b4, const int64_t *c2, const int64_t *c3,
                         const int64_t *c4, const double *G_plus, const double *G_minus, const double *G_pp,
                         const double *G_pm, const double *G_mm) {
  int rc;
  if ((rc = bte_ready(c))) return rc;
  if (nlist != c->nlist) return fail(FP4_ERR_ARG, "Nlist=%d differs from the modes' %d", nlist, c->nlist);
  if (!TypeofSymmetry) return fail(FP4_ERR_ARG, "null TypeofSymmetry");
  if (!N_pp || !N_pm || !N_mm) return fail(FP4_ERR_ARG, "null 4ph counts");
  const int nm = c->nb * c->nlist;
  // the stated list lengths must match the counts
  const int32_t *N3[2] = {N_plus, N_minus};
  const long long tot3[2] = {Ntotal_plus, Ntotal_minus};
  for (int t = 0; t < 2; t++) {
    long long s = 0;
    if (N3[t])
      for (int mm = 0; mm < nm; mm++) s += N3[t][mm];
    if (s != tot3[t]) return fail(FP4_ERR_ARG, "3ph counts sum to %lld, Ntotal is %lld", s, tot3[t]);
  }
  const int64_t *N4[3] = {N_pp, N_pm, N_mm};
  const long long tot4[3] = {Ntotal_pp, Ntotal_pm, Ntotal_mm};
  for (int t = 0; t < 3; t++) {
    long long s = 0;
    for (int mm = 0; mm < nm; mm++) s += N4[t][mm];
    if (s != tot4[t]) return fail(FP4_ERR_ARG, "4ph counts (channel %d) sum to %lld, Ntotal is %lld", t, s, tot4[t]);
  }
  const BteModes m = bte_modes(c);
  const int32_t *L3[2][2] = {{p2, p3}, {m2, m3}};
  const double *G3[2] = {G_plus, G_minus};
  const int64_t *L4[3][3] = {{a2, a3, a4}, {b2, b3, b4}, {c2, c3, c4}};
  const double *G4[3] = {G_pp, G_pm, G_mm};
  BTE_CALL(bte_set_orbits(c->bte, m, Nequi, ALLEquiList, TypeofSymmetry));
  BTE_CALL(bte_set_operator(c->bte, m, N3, L3, G3, N4, L4, G4));
  return FP4_OK;
}

// Adds the per-process Gamma of the last rate batch (its block descriptors
// in c->blk, lists in c->lists, Gamma in c->proc_g) to the operator W.
static int scatter_batch(fp4_ctx *c, double *W, long long cols, int order) {
  WScatterArgs a{};
  a.blocks = c->blk.as<BlockDesc>();
  a.nblocks = c->last_nblk;
  for (int t = 0; t < 3; t++) a.lst[t] = c->lists[t].as<int>();
  a.gamma = c->proc_g.as<double>();
  a.W = W;
  a.wcols = cols;
  a.order = order;
  launch_w_scatter(a, c->stream);
  HIPCHECK(hipGetLastError());
  return 0;
}

int fp4_bte_operator_begin(fp4_ctx *c, int32_t nlist, const int32_t *Nequi, const int32_t *ALLEquiList,
                           const int32_t *TypeofSymmetry) {
  int rc;
  if ((rc = bte_ready(c))) return rc;
  if (nlist != c->nlist) return fail(FP4_ERR_ARG, "Nlist=%d differs from the modes' %d", nlist, c->nlist);
  if (!TypeofSymmetry) return fail(FP4_ERR_ARG, "null TypeofSymmetry");
  if ((long long)c->nptk * c->nb >= (1LL << 31)) return fail(FP4_ERR_UNSUPPORTED, "nptk*Nbands must be < 2^31");
  const BteModes m = bte_modes(c);
  BTE_CALL(bte_set_orbits(c->bte, m, Nequi, ALLEquiList, TypeofSymmetry));
  double *W = nullptr;
  long long cols = 0;
  BTE_CALL(bte_w_begin(c->bte, m, &W, &cols));
  c->w_op = W;
  c->w_cols = cols;
  return FP4_OK;
}

int fp4_bte_operator_add(fp4_ctx *c, int32_t precision, int32_t with_3ph, int32_t with_4ph, int32_t nmodes,
                         const int32_t *modes, int64_t *nproc3, int64_t *nproc4, double *rate_3ph,
                         double *rate_4ph) {
  int rc;
  if ((rc = bte_ready(c))) return rc;
  if (!c->w_op) return fail(FP4_ERR_STATE, "fp4_bte_operator_begin has not been called");
  if (with_3ph && (rc = check_ready3(c, true))) return rc;
  if (with_4ph && (rc = check_ready(c, true))) return rc;
  Plan sub;
  if ((rc = subset_plan(c, nmodes, modes, sub))) return rc;
  double *W = c->w_op;
  const long long cols = c->w_cols;
  const int nm = c->nb * c->nlist;
  long long n3 = 0, n4 = 0;
  if (with_3ph) {
    // Ind_driver's lists and per-process Gamma, batch by batch, straight into W
    Plan pl;
    pl.modes = sub.modes;
    pl.win_lo0 = 0;
    pl.win_n = c->nptk;
    if ((rc = count3_pass(c, pl, true))) return rc;
    if ((rc = counts3_fit(pl))) return rc;
    std::vector<double> g(2 * (size_t)nmodes, 0.0), w(2 * (size_t)nmodes, 0.0);
    for (const Batch &b : plan_batches(pl, c->list_cap)) {
      if ((rc = emit3_batch(c, pl, b))) return rc;
      if ((rc = rate3_batch(c, pl, b, true, g, w))) return rc;
      if ((rc = scatter_batch(c, W, cols, 3))) return rc;
    }
    n3 = pl.off[pl.items];
    if (rate_3ph) {
      // Ind_driver's per-mode rates (sums of the per-process Gamma in THz)
      std::vector<double> rp(nm), rm(nm), wp(nm), wm(nm);
      write3_rates(c, pl, g, w, rate_3ph, rp.data(), rm.data(), wp.data(), wm.data(), false);
    }
  }
  if (with_4ph) {
    std::vector<double> t(7 * (size_t)nm);
    if (precision == FP4_PREC_FP64) {
      // Ind_driver_4ph's per-process FP64 Gamma, batch by batch
      Plan pl = sub;
      if ((rc = count_pass(c, pl, true))) return rc;
      std::vector<double> g(3 * (size_t)nmodes, 0.0), w(3 * (size_t)nmodes, 0.0);
      for (const Batch &b : plan_batches(pl, c->list_cap)) {
        if ((rc = emit_batch(c, pl, b, false))) return rc;
        if ((rc = rate_batch(c, pl, b, false, true, g, w))) return rc;
        if ((rc = scatter_batch(c, W, cols, 4))) return rc;
      }
      n4 = pl.off[pl.items];
      if (rate_4ph)
        write_rates(c, pl, g, w, rate_4ph, &t[0], &t[nm], &t[2 * nm], &t[3 * nm], &t[4 * nm], &t[5 * nm], false);
    } else if (precision == FP4_PREC_FP32 || precision == FP4_PREC_FP64_DENSE || precision == FP4_PREC_BF16X3 ||
               precision == FP4_PREC_FP16X3 || precision == FP4_PREC_FP16X2) {
      // the dense form computes every process's Gamma and adds it to W
      DenseRun dr;
      dr.m = mode_view(c);
      dr.roots = c->roots.as<double2>();
      dr.T = c->T;
      dr.onlyharmonic = c->onlyharmonic;
      dr.natoms = c->natoms;
      dr.precision = precision;
      dr.stream = c->stream;
      dr.tuning = &c->dense_tuning;
      dr.wop = W;
      dr.wcols = cols;
      double ms = 0.0;
      std::vector<double> g, w;
      std::vector<long long> n;
      rc = dense_rates(c->dense, dr, sub.modes, 0, c->nptk, g, w, n, &ms);
      if (rc) return fail(rc, "dense path: %s", dense_error());
      c->ks_ms[KS_DENSE] += ms;
      c->ks_n[KS_DENSE] += 1;
      n4 = std::accumulate(n.begin(), n.end(), 0LL);
      if (rate_4ph)
        write_rates(c, sub, g, w, rate_4ph, &t[0], &t[nm], &t[2 * nm], &t[3 * nm], &t[4 * nm], &t[5 * nm], true);
    } else {
      return fail(FP4_ERR_ARG, "unknown precision %d", precision);
    }
  }
  if (nproc3) *nproc3 = n3;
  if (nproc4) *nproc4 = n4;
  return FP4_OK;
}

int fp4_bte_operator_end(fp4_ctx *c) {
  int rc;
  if ((rc = bte_ready(c))) return rc;
  if (!c->w_op) return fail(FP4_ERR_STATE, "fp4_bte_operator_begin has not been called");
  BTE_CALL(bte_w_finish(c->bte, bte_modes(c)));
  c->w_op = nullptr;
  return FP4_OK;
}

int fp4_bte_assemble_operator(fp4_ctx *c, int32_t nlist, const int32_t *Nequi, const int32_t *ALLEquiList,
                              const int32_t *TypeofSymmetry, int32_t precision, int32_t with_3ph, int32_t with_4ph,
                              int64_t *nproc3, int64_t *nproc4, double *rate_3ph, double *rate_4ph) {
  int rc;
  if ((rc = fp4_bte_operator_begin(c, nlist, Nequi, ALLEquiList, TypeofSymmetry))) return rc;
  std::vector<int> modes((size_t)c->nb * c->nlist);
  std::iota(modes.begin(), modes.end(), 1);
  if ((rc = fp4_bte_operator_add(c, precision, with_3ph, with_4ph, (int)modes.size(), modes.data(), nproc3, nproc4,
                                 rate_3ph, rate_4ph)))
    return rc;
  return fp4_bte_operator_end(c);
}

int fp4_bte_iterate(fp4_ctx *c, const double *tau_zero, double *F_n) {
  int rc;
  if ((rc = bte_ready(c))) return rc;
  const BteModes m = bte_modes(c);
  BTE_CALL(bte_put_F(c->bte, m, F_n));
  BTE_CALL(bte_sweep(c->bte, m, tau_zero));
  c->ks_ms[KS_ITER] += bte_last_ms(c->bte);
  c->ks_n[KS_ITER] += bte_last_launches(c->bte);
  BTE_CALL(bte_get_F(c->bte, m, F_n));
  return FP4_OK;
}

int fp4_bte_set_sweep(fp4_ctx *c, int32_t persistent) {
  int rc;
  if ((rc = bte_ready(c))) return rc;
  BTE_CALL(bte_set_persistent(c->bte, persistent != 0));
  return FP4_OK;
}

int fp4_iteration_4ph(fp4_ctx *c, int32_t nlist, const int32_t *Nequi, const int32_t *ALLEquiList,
                      const int32_t *TypeofSymmetry, const int32_t *N_plus, const int32_t *N_minus,
                      int32_t Ntotal_plus, int32_t Ntotal_minus, const int64_t *N_pp, const int64_t *N_pm,
                      const int64_t *N_mm, int64_t Ntotal_pp, int64_t Ntotal_pm, int64_t Ntotal_mm,
                      const int32_t *p2, const int32_t *p3, const int32_t *m2, const int32_t *m3, const int64_t *a2,
                      const int64_t *a3, const int64_t *a4, const int64_t *b2, const int64_t *b3, const int64_t *b4,
                      const int64_t *c2, const int64_t *c3, const int64_t *c4, const double *G_plus,
                      const double *G_minus, const double *G_pp, const double *G_pm, const double *G_mm,
                      const double *tau_zero, double *F_n) {
  int rc;
  if ((rc = fp4_bte_set_operator(c, nlist, Nequi, ALLEquiList, TypeofSymmetry, N_plus, N_minus, Ntotal_plus,
                                 Ntotal_minus, N_pp, N_pm, N_mm, Ntotal_pp, Ntotal_pm, Ntotal_mm, p2, p3, m2, m3, a2,
                                 a3, a4, b2, b3, b4, c2, c3, c4, G_plus, G_minus, G_pp, G_pm, G_mm)))
    return rc;
  return fp4_bte_iterate(c, tau_zero, F_n);
}

int fp4_tconduct(fp4_ctx *c, double V, const double *F_n, double *kappa, double *mode) {
  int rc;
  if ((rc = bte_ready(c))) return rc;
  const BteModes m = bte_modes(c);
  BTE_CALL(bte_put_F(c->bte, m, F_n));
  BTE_CALL(bte_tconduct(c->bte, m, c->T, V, kappa, mode));
  return FP4_OK;
}

int fp4_bte_solve(fp4_ctx *c, double V, const double *tau_zero, int32_t maxiter, double eps, double *F_n,
                  double *kappa, double *kappa_iter, double *relchange, int32_t *niter) {
  int rc;
  if ((rc = bte_ready(c))) return rc;
  if (!F_n || !niter) return fail(FP4_ERR_ARG, "null output");
  if (c->nb > 64) return fail(FP4_ERR_UNSUPPORTED, "Nbands > 64");
  const BteModes m = bte_modes(c);
  int n = 0;
  BTE_CALL(bte_solve(c->bte, m, c->T, V, tau_zero, maxiter, eps, kappa, kappa_iter, relchange, &n));
  c->ks_ms[KS_ITER] += bte_last_ms(c->bte);
  c->ks_n[KS_ITER] += bte_last_launches(c->bte);
  *niter = n;
  BTE_CALL(bte_get_F(c->bte, m, F_n));
  return FP4_OK;
}

// ---- batches of crystals ---------------------------------------------------
}  // extern "C"

struct fp4_batch {
  std::vector<fp4_ctx *> ctx;
  int device = 0;
  hipStream_t stream = nullptr;
  DenseBatchState *dense = nullptr;
  // per (part, nparts) and context: the count pass over the part's q'
  // window, per (channel, mode, q'); it orders the queue and does not depend on T
  std::map<std::pair<int, int>, std::vector<std::vector<int>>> counts;
};

extern "C" {

int fp4_batch_destroy(fp4_batch *b) {
  if (!b) return FP4_OK;
  (void)hipSetDevice(b->device);
  dense_batch_destroy(b->dense);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return FP4_OK;
}

int fp4_batch_create(int32_t n, fp4_ctx *const *ctxs, fp4_batch **out) {
  if (!out || n < 1 || !ctxs) return fail(FP4_ERR_ARG, "bad arguments");
  *out = nullptr;
  int rc;
  for (int i = 0; i < n; i++) {
    if ((rc = check_ready(ctxs[i], true))) return rc;
    if (ctxs[i]->device != ctxs[0]->device) return fail(FP4_ERR_ARG, "contexts on different devices");
  }
  fp4_batch *b = new fp4_batch();
  b->ctx.assign(ctxs, ctxs + n);
  b->device = ctxs[0]->device;
  if (hipSetDevice(b->device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      !(b->dense = dense_batch_create())) {
    fp4_batch_destroy(b);
    return fail(FP4_ERR_DEVICE, "batch: stream or event creation failed");
  }
  *out = b;
  return FP4_OK;
}

int fp4_batch_rta_4ph(fp4_batch *b, int32_t precision, int32_t use_graph, int32_t part, int32_t nparts,
                      double *const *rates, int64_t *nproc, double *kernel_ms, int32_t *launches) {
  if (!b || !rates) return fail(FP4_ERR_ARG, "bad arguments");
  if (nparts < 1 || part < 0 || part >= nparts) return fail(FP4_ERR_ARG, "part %d of %d", part, nparts);
  if (precision != FP4_PREC_FP32 && precision != FP4_PREC_FP16X2)
    return fail(FP4_ERR_UNSUPPORTED, "batch: FP4_PREC_FP32 or FP4_PREC_FP16X2 only");
  HIPCHECK(hipSetDevice(b->device));
  const int n = (int)b->ctx.size();
  for (int i = 0; i < n; i++) {
    if (!rates[i]) return fail(FP4_ERR_ARG, "null output of crystal %d", i);
    std::fill(rates[i], rates[i] + 7 * (size_t)b->ctx[i]->nb * b->ctx[i]->nlist, 0.0);
    if (nproc) nproc[i] = 0;
  }
  // the prefilter counts of this part's windows, once per (part, nparts)
  auto key = std::make_pair((int)part, (int)nparts);
  if (!b->counts.count(key)) {
    std::vector<std::vector<int>> cs(n);
    for (int i = 0; i < n; i++) {
      fp4_ctx *c = b->ctx[i];
      if (c->natoms > kDenseMaxAtoms) continue;
      const long lo = (long)part * c->nptk / nparts, hi = (long)(part + 1) * c->nptk / nparts;
      if (hi <= lo) continue;
      Plan pl;
      all_modes(c, pl);
      pl.ii_lo0 = (int)lo;
      pl.nii = (int)(hi - lo);
      int rc;
      if ((rc = count_pass(c, pl, false))) return rc;
      cs[i].resize(pl.items);
      HIPCHECK(hipMemcpy(cs[i].data(), c->cnt.p, sizeof(int) * pl.items, hipMemcpyDeviceToHost));
    }
    b->counts[key] = std::move(cs);
  }
  const std::vector<std::vector<int>> &counts = b->counts[key];
  std::vector<DenseBatchMember> mem;
  std::vector<int> who;
  for (int i = 0; i < n; i++) {
    fp4_ctx *c = b->ctx[i];
    if (c->natoms > kDenseMaxAtoms) continue;
    DenseBatchMember x;
    x.s = c->dense;
    x.r.m = mode_view(c);
    x.r.roots = c->roots.as<double2>();
    x.r.T = c->T;
    x.r.onlyharmonic = c->onlyharmonic;
    x.r.natoms = c->natoms;
    x.r.precision = precision;
    x.r.stream = b->stream;
    x.r.tuning = nullptr;
    x.modes.resize((size_t)c->nb * c->nlist);
    std::iota(x.modes.begin(), x.modes.end(), 1);
    // q' window of this part: [part*nptk/nparts, (part+1)*nptk/nparts)
    const long lo = (long)part * c->nptk / nparts, hi = (long)(part + 1) * c->nptk / nparts;
    if (hi <= lo) continue;
    x.ii_lo0 = (int)lo;
    x.nii = (int)(hi - lo);
    x.counts = counts[i];     // (channel, mode, q') over the window, the count pass's layout
    mem.push_back(std::move(x));
    who.push_back(i);
  }
  double ms = 0.0;
  int nl = 0;
  if (!mem.empty()) {
    const int rc = dense_rates_batch(b->dense, mem, precision, b->stream, use_graph != 0, part, &ms, &nl);
    if (rc) return fail(rc, "batch: %s", dense_error());
  }
  for (size_t k = 0; k < mem.size(); k++) {
    const int i = who[k];
    fp4_ctx *c = b->ctx[i];
    const size_t nm = (size_t)c->nb * c->nlist;
    Plan pl;
    all_modes(c, pl);
    double *o = rates[i];
    std::fill(o, o + 7 * nm, 0.0);
    write_rates(c, pl, mem[k].g, mem[k].w, o, o + nm, o + 2 * nm, o + 3 * nm, o + 4 * nm, o + 5 * nm, o + 6 * nm,
                true);
    if (nproc) nproc[i] = std::accumulate(mem[k].n.begin(), mem[k].n.end(), 0LL);
  }
  // cells beyond the dense kernel's 8 atoms: the FP64 per-process form
  for (int i = 0; i < n; i++) {
    fp4_ctx *c = b->ctx[i];
    if (c->natoms <= kDenseMaxAtoms) continue;
    const size_t nm = (size_t)c->nb * c->nlist;
    std::vector<int> modes(nm);
    std::iota(modes.begin(), modes.end(), 1);
    double *o = rates[i];
    std::fill(o, o + 7 * nm, 0.0);
    int64_t np = 0;
    const int lo = (int)((long)part * c->nptk / nparts), hi = (int)((long)(part + 1) * c->nptk / nparts);
    if (hi <= lo) continue;
    int rc = fp4_rta_4ph_modes(c, FP4_PREC_FP64, (int)nm, modes.data(), lo + 1, hi, o, o + nm, o + 2 * nm, o + 3 * nm,
                               o + 4 * nm, o + 5 * nm, o + 6 * nm, &np);
    if (rc) return rc;
    if (nproc) nproc[i] = np;
  }
  if (kernel_ms) *kernel_ms = ms;
  if (launches) *launches = nl;
  return FP4_OK;
}

// ---- diagnostics (no counterpart in the reference) -----------------------
int fp4_ctx_kernel_stats(fp4_ctx *c, int32_t kind, double *total_ms, int64_t *launches) {
  if (!c || kind < 0 || kind >= KS_N) return fail(FP4_ERR_ARG, "bad arguments");
  if (total_ms) *total_ms = c->ks_ms[kind];
  if (launches) *launches = c->ks_n[kind];
  return FP4_OK;
}

int fp4_ctx_reset_stats(fp4_ctx *c) {
  if (!c) return fail(FP4_ERR_ARG, "null context");
  for (int k = 0; k < KS_N; k++) {
    c->ks_ms[k] = 0;
    c->ks_n[k] = 0;
  }
  return FP4_OK;
}

int fp4_ctx_dense_info(fp4_ctx *c, int64_t *triples, int64_t *items, int32_t *ngroups, int32_t *nfine) {
  if (!c || !triples || !items || !ngroups || !nfine) return fail(FP4_ERR_ARG, "null argument");
  long long t, it;
  int g, f;
  dense_info(c->dense, &t, &it, &g, &f);
  *triples = t;
  *items = it;
  *ngroups = g;
  *nfine = f;
  return FP4_OK;
}

int fp4_ctx_autotune_dense(fp4_ctx *c, int32_t precision, int32_t nmodes, const int32_t *modes, int32_t ii_lo,
                           int32_t ii_hi, int32_t *ib, int32_t *tq, int32_t *slots, double *best_ms) {
  int rc;
  if ((rc = check_ready(c, true))) return rc;
  if (precision != FP4_PREC_FP32 && precision != FP4_PREC_FP64_DENSE && precision != FP4_PREC_BF16X3 &&
      precision != FP4_PREC_FP16X3 && precision != FP4_PREC_FP16X2)
    return fail(FP4_ERR_ARG, "autotune: precision %d has no dense kernel", precision);
  if (ii_lo < 1 || ii_hi > c->nptk || ii_hi < ii_lo) return fail(FP4_ERR_ARG, "bad ii window %d..%d", ii_lo, ii_hi);
  Plan pl;
  if ((rc = subset_plan(c, nmodes, modes, pl))) return rc;
  // The kernels share work across the branches of one irreducible point
  // (IB modes per block), so the sample is widened to every branch of each
  // point it touches; a one-branch-per-point sample would favour small IB.
  {
    std::vector<int> all;
    for (int mm : pl.modes) {
      const int first = (mm - 1) / c->nb * c->nb + 1;
      if (all.empty() || all.back() < first)
        for (int b = 0; b < c->nb; b++) all.push_back(first + b);
    }
    pl.modes = all;
  }
  DenseRun dr;
  dr.m = mode_view(c);
  dr.roots = c->roots.as<double2>();
  dr.T = c->T;
  dr.onlyharmonic = c->onlyharmonic;
  dr.natoms = c->natoms;
  dr.precision = precision;
  dr.stream = c->stream;
  dr.tuning = nullptr;
  DenseTuning best;
  double ms = 0.0;
  int ncand = 0;
  rc = dense_autotune(c->dense, dr, pl.modes, ii_lo - 1, ii_hi - ii_lo + 1, &best, &ms, &ncand);
  if (rc) return fail(rc, "autotune: %s", dense_error());
  c->dense_tuning = best;
  int a, b, d;
  dense_shape(c->dense, &a, &b, &d);
  if (ib) *ib = d;
  if (tq) *tq = b;
  if (slots) *slots = best.slots;
  if (best_ms) *best_ms = ms;
  return FP4_OK;
}

int fp4_ctx_dense_shape(fp4_ctx *c, int32_t *mfma, int32_t *tq, int32_t *ib) {
  if (!c || !mfma || !tq || !ib) return fail(FP4_ERR_ARG, "null argument");
  int a, b, d;
  dense_shape(c->dense, &a, &b, &d);
  *mfma = a;
  *tq = b;
  *ib = d;
  return FP4_OK;
}

int fp4_ctx_dense_candidates(fp4_ctx *c, int32_t cap, int32_t *kg, int32_t *ib, int32_t *tq, int32_t *slots,
                             double *ms) {
  if (!c || cap < 0) return -FP4_ERR_ARG;
  std::vector<int> k(cap), i(cap), t(cap), sl(cap);
  const int n = dense_candidates(c->dense, cap, k.data(), i.data(), t.data(), sl.data(), ms);
  for (int x = 0; x < n && x < cap; x++) {
    if (kg) kg[x] = k[x];
    if (ib) ib[x] = i[x];
    if (tq) tq[x] = t[x];
    if (slots) slots[x] = sl[x];
  }
  return n;
}

int fp4_ctx_dense_phase_profile(fp4_ctx *c, int32_t cap, double *cycles) {
  if (!c || cap < 0 || (cap > 0 && !cycles)) return -FP4_ERR_ARG;
  const int n = dense_phase_profile(c->dense, cap, cycles);
  if (n < 0) fail(-n, "dense phase profile: this library was built without FP4_PHASE_PROF");
  return n;
}

int fp4_ctx_set_dense_shape(fp4_ctx *c, int32_t precision, int32_t kg, int32_t ib, int32_t slots) {
  if (!c) return fail(FP4_ERR_ARG, "null argument");
  if (!c->dense) return fail(FP4_ERR_STATE, "set_dense_shape: fourth-order constants not set");
  DenseTuning t;
  const int rc = dense_pick(c->dense, precision, c->nb, kg, ib, slots, &t);
  if (rc) return fail(rc, "%s", dense_error());
  c->dense_tuning = t;
  return FP4_OK;
}

int fp4_ctx_dense_kernel(fp4_ctx *c, int32_t *kg) {
  if (!c || !kg) return fail(FP4_ERR_ARG, "null argument");
  int a, b, d, k = 0;
  dense_shape(c->dense, &a, &b, &d, &k);
  *kg = k;
  return FP4_OK;
}

int fp4_selftest_mfma(int32_t device, double *max_err) {
  if (!max_err) return fail(FP4_ERR_ARG, "null argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return fail(FP4_ERR_DEVICE, "no device");
  HIPCHECK(hipSetDevice(device));
  const int rc = dense_selftest_mfma(max_err);
  return rc ? fail(rc, "%s", dense_error()) : FP4_OK;
}

int fp4_ctx_set_batch_capacity(fp4_ctx *c, int64_t processes) {
  if (!c || processes < 1024) return fail(FP4_ERR_ARG, "capacity must be at least 1024 processes");
  c->list_cap = processes;
  return FP4_OK;
}

}  // extern "C"
