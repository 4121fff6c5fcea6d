// mpk_pc.hip -- preconditioner kernels: Jacobi set up and apply, SOR, and the
// ILU / IC triangular solves, plus the mp_pc_* device exports.
//
// Triangular solves and SOR sweeps are sync-free (mpk_sweep.h): each row is a
// serial sum in the order the reference's loop visits it, so the factor apply
// gives the same bits as ILUApplyFactor_SeqAIJ / ICCApplyFactor_SeqAIJ /
// PCApply_SOR (src/ported.mojo:633-952) on the same factor.
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

#include "mpk_internal.h"
#include "mpk_ops.h"
#include "mpk_pc.h"
#include "mpk_sweep.h"
#include "mpk_progs.h"
#include "../../include/mpk.h"

namespace mpk {

// ------------------------------------------------------------ sweep plumbing

// Fills the published-value buffer with the sentinel and zeroes the ticket
// counter; the sweep that follows on the stream starts from both.
__global__ void k_sweep_prepare(i64 n, double *p, unsigned *ticket, const int64_t *stop) {
  if (stop && *stop) return;
  i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i == 0) {
    ticket[0] = 0u;  // block order of a block plan
    ticket[1] = 0u;  // progress of a ring plan
  }
  if (i < n) p[i] = __longlong_as_double((long long)MPK_SENTINEL);
}

// Which sweeps of a PC's plans run on the one-workgroup-per-block kernels
// and how (hdr[7] and hdr[5] of each: bits 0-15 the first sweep, 16-31 the
// second; a sweep's word is its slot count, plus for a wave or wave-block
// plan 16, 32 when none of its rows has over 2 entries, 64 times log2 of its
// waves, and 256 for wave blocks; 0 for other plans), with each sweep's
// block count, recorded on the host by PCView::schedule when it writes them,
// so an apply picks the kernel and its grid without reading the device.  A
// plan this process did not write counts as neither (block count 0).
struct PlanNote {
  int bits = 0;
  int nblk[2] = {0, 0};
};
static std::mutex g_narrow_mu;
static std::map<const char *, PlanNote> &narrow_plans() {
  static std::map<const char *, PlanNote> m;
  return m;
}
static void note_narrow(const char *plan, const PlanNote &note) {
  std::lock_guard<std::mutex> lk(g_narrow_mu);
  narrow_plans()[plan] = note;
}
// A sweep's word of the record above.
static int plan_narrow(const char *plan, int second) {
  std::lock_guard<std::mutex> lk(g_narrow_mu);
  auto it = narrow_plans().find(plan);
  return it == narrow_plans().end() ? 0 : (it->second.bits >> (second ? 16 : 0)) & 0xffff;
}
// A sweep's block count, or 0 when unknown.
static int plan_blocks(const char *plan, int second) {
  std::lock_guard<std::mutex> lk(g_narrow_mu);
  auto it = narrow_plans().find(plan);
  return it == narrow_plans().end() ? 0 : it->second.nblk[second ? 1 : 0];
}

// One sweep: over its block plan when the PC has one (`plan`, and `second`
// for the PC's second sweep), else in row order.  Narrow and wave plans need
// no sentinel fill (their kernels poll nothing): two launches for a narrow
// plan, one for a wave plan (its kernel stages the rows itself).
template <class Prog>
static void run_sweep(const Prog &prog, i64 n, bool forward, double *pub, const int64_t *stop,
                      const char *plan = nullptr, int second = 0, Staged *st = nullptr) {
  Ctx *c = ctx();
  if (!c || n <= 0) return;
  watch_install();
  const int form = plan ? plan_narrow(plan, second) : 0;
  const int narrow = form & 15;  // the plan's slot count, or 0
  if (form & 256) {  // wave blocks: published values, so the sentinel fill
    const int nw = 1 << ((form >> 6) & 3);
    // one workgroup per block (a launch past convergence then costs little);
    // without the count, blocks hold at least min(n, 1024) rows, so this grid
    // covers every block and the spare workgroups find none and return
    const int known = plan_blocks(plan, second);
    const unsigned nblk = known > 0 ? (unsigned)known : nblocks(n, 1024);
    MPK_LAUNCH(k_sweep_prepare, nblocks(n, 256), 256, n, pub, c->tickets, stop);
    if (narrow == 4 && (form & 32))
      MPK_LAUNCH((k_sweep_wblk<Prog, 4, 2>), nblk, MPK_WAVE_THREADS, prog, n, plan, second, st, pub, c->tickets,
                 nw, stop);
    else if (narrow == 4)
      MPK_LAUNCH((k_sweep_wblk<Prog, 4, 4>), nblk, MPK_WAVE_THREADS, prog, n, plan, second, st, pub, c->tickets,
                 nw, stop);
    else
      MPK_LAUNCH((k_sweep_wblk<Prog, 8, 8>), nblk, MPK_WAVE_THREADS, prog, n, plan, second, st, pub, c->tickets,
                 nw, stop);
    return;
  }
  if (form & 16) {
    const int nw = 1 << ((form >> 6) & 3);
    if (narrow == 4 && (form & 32))
      MPK_LAUNCH((k_sweep_wave<Prog, 4, 2>), 1, MPK_WAVE_THREADS, prog, n, plan, second, st, pub, nw, stop);
    else if (narrow == 4)
      MPK_LAUNCH((k_sweep_wave<Prog, 4, 4>), 1, MPK_WAVE_THREADS, prog, n, plan, second, st, pub, nw, stop);
    else
      MPK_LAUNCH((k_sweep_wave<Prog, 8, 8>), 1, MPK_WAVE_THREADS, prog, n, plan, second, st, pub, nw, stop);
    return;
  }
  if (narrow) {
    MPK_LAUNCH((k_sweep_stage<Prog, Staged>), nblocks(n, 256), 256, prog, n, plan, second, st, stop);
    if (narrow == 4)
      MPK_LAUNCH((k_sweep_narrow<Prog, 4>), 1, MPK_NARROW_THREADS, prog, n, plan, second,
                 (const Staged *)st, pub, stop);
    else
      MPK_LAUNCH((k_sweep_narrow<Prog, 8>), 1, MPK_NARROW_THREADS, prog, n, plan, second,
                 (const Staged *)st, pub, stop);
    return;
  }
  MPK_LAUNCH(k_sweep_prepare, nblocks(n, 256), 256, n, pub, c->tickets, stop);
  if (plan) {
    MPK_LAUNCH((k_sweep_stage<Prog, Staged>), nblocks(n, 256), 256, prog, n, plan, second, st, stop);
    // blocks hold at least min(n, 1024) rows (mpk_sweep_plan), so this grid
    // covers every block; the spare workgroups find no block and return
    MPK_LAUNCH((k_sweep_blk<Prog, SweepF64, PlanCoef, Staged>), nblocks(n, 1024), 256, prog, n, plan, second,
               (const Staged *)st, (const void *)nullptr, c->tickets, pub, stop);
  } else {
    MPK_LAUNCH(k_sweep<Prog>, nblocks(n, 64), 64, prog, n, forward ? 1 : 0, c->tickets, 0u, pub, stop);
  }
}

// The entries of one sweep on the host: off[n+1], col, coef (k_sweep_describe).
template <class Prog>
static bool describe_sweep(const Prog &prog, i64 n, std::vector<i64> &off, std::vector<int> &col,
                           std::vector<double> &coef) {
  int *cnt_d = nullptr, *col_d = nullptr;
  i64 *off_d = nullptr;
  double *coef_d = nullptr;
  std::vector<int> cnt(n);
  HIPCHK(hipMalloc(&cnt_d, n * sizeof(int)));
  if (failed()) return false;
  MPK_LAUNCH(k_sweep_describe<Prog>, nblocks(n, 256), 256, prog, n, 0, cnt_d, (const i64 *)nullptr,
             (int *)nullptr, (double *)nullptr);
  HIPCHK(hipMemcpyAsync(cnt.data(), cnt_d, n * sizeof(int), hipMemcpyDeviceToHost, stream()));
  HIPCHK(hipStreamSynchronize(stream()));
  off.assign(n + 1, 0);
  for (i64 i = 0; i < n; i++) off[i + 1] = off[i] + cnt[i];
  const i64 m = off[n], mm = m > 0 ? m : 1;
  col.assign(mm, 0);
  coef.assign(mm, 0.0);
  HIPCHK(hipMalloc(&off_d, (n + 1) * sizeof(i64)));
  HIPCHK(hipMalloc(&col_d, mm * sizeof(int)));
  HIPCHK(hipMalloc(&coef_d, mm * sizeof(double)));
  if (!failed()) {
    copy_htod(off_d, off.data(), (n + 1) * sizeof(i64));
    MPK_LAUNCH(k_sweep_describe<Prog>, nblocks(n, 256), 256, prog, n, 1, cnt_d, (const i64 *)off_d, col_d, coef_d);
    HIPCHK(hipMemcpyAsync(col.data(), col_d, mm * sizeof(int), hipMemcpyDeviceToHost, stream()));
    HIPCHK(hipMemcpyAsync(coef.data(), coef_d, mm * sizeof(double), hipMemcpyDeviceToHost, stream()));
    HIPCHK(hipStreamSynchronize(stream()));
  }
  HIPCHK(hipFree(cnt_d));
  HIPCHK(hipFree(off_d));
  HIPCHK(hipFree(col_d));
  HIPCHK(hipFree(coef_d));
  return !failed();
}

// The entries of one sweep on the device (k_sweep_describe; the offsets are
// scanned on the host from the counts), into buffers the caller frees.
template <class Prog>
static bool describe_sweep_dev(const Prog &prog, i64 n, DevEntries &E, std::vector<void *> &keep) {
  int *cnt_d = nullptr, *col_d = nullptr;
  i64 *off_d = nullptr;
  double *coef_d = nullptr;
  HIPCHK(hipMalloc(&cnt_d, n * sizeof(int)));
  if (failed()) return false;
  keep.push_back(cnt_d);
  MPK_LAUNCH(k_sweep_describe<Prog>, nblocks(n, 256), 256, prog, n, 0, cnt_d, (const i64 *)nullptr,
             (int *)nullptr, (double *)nullptr);
  std::vector<int> cnt(n);
  copy_dtoh(cnt.data(), cnt_d, n * sizeof(int));
  std::vector<i64> off(n + 1, 0);
  i64 widest = 0;
  for (i64 i = 0; i < n; i++) {
    off[i + 1] = off[i] + cnt[i];
    widest = cnt[i] > widest ? cnt[i] : widest;
  }
  const i64 m = off[n], mm = m > 0 ? m : 1;
  HIPCHK(hipMalloc(&off_d, (n + 1) * sizeof(i64)));
  HIPCHK(hipMalloc(&col_d, mm * sizeof(int)));
  HIPCHK(hipMalloc(&coef_d, mm * sizeof(double)));
  for (void *q : {(void *)off_d, (void *)col_d, (void *)coef_d})
    if (q) keep.push_back(q);
  if (failed()) return false;
  copy_htod(off_d, off.data(), (n + 1) * sizeof(i64));
  MPK_LAUNCH(k_sweep_describe<Prog>, nblocks(n, 256), 256, prog, n, 1, cnt_d, (const i64 *)off_d, col_d, coef_d);
  E = DevEntries{n, m, off_d, col_d, coef_d, widest};
  return !failed();
}

// The narrow word of a plan from its header's w (hdr[5]) and mark (hdr[7]),
// its longest row and widest level: narrow: the slot count; wave and wave
// blocks: the count plus 16, 32 when no row has over 2 entries, 64 log2 of
// its waves (one per 64 rows of its widest level), and 256 for wave blocks.
static int narrow_word(i64 w, i64 mark, int longest, int widest) {
  const int lw = widest <= 64 ? 0 : widest <= 128 ? 1 : 2;
  return mark == 1 ? (int)w
         : mark >= 2 ? ((int)w | 16 | (longest <= 2 ? 32 : 0) | (lw << 6) | (mark == 3 ? 256 : 0))
                     : 0;
}

// The block plan of one sweep into dst [D] (cap bytes, at least the bound);
// returns its size in bytes and the deepest block's level count, or -1.
// Block plans are built on the device (device_sweep_plan, byte for byte the
// host planner's); ring plans, and every plan with MPK_SWEEP_DEVICE_PLAN=0,
// by the host planner.
template <class Prog>
static i64 plan_sweep(const Prog &prog, i64 n, bool forward, char *dst, i64 cap, i64 *depth, int *narrow) {
  if (!plan_env_off("MPK_SWEEP_DEVICE_PLAN")) {
    DevEntries E{};
    std::vector<void *> keep;
    i64 bytes = -1, d = 0;
    int longest = 0, widest = 0, mark = 0;
    if (describe_sweep_dev(prog, n, E, keep)) {
      const i64 bound = mpk_sweep_plan_bound(n, E.m, MPK_BLK_ROWS);
      if (bound > cap)
        set_error(5, "sweep plan buffer too small");
      else
        bytes = device_sweep_plan(E, forward, MPK_BLK_ROWS, dst, cap, &d, &longest, &widest, &mark);
    }
    HIPCHK(hipStreamSynchronize(stream()));
    for (void *q : keep) HIPCHK(hipFree(q));
    if (bytes < 0 || failed()) return -1;
    if (bytes > 0) {
      i64 w = 0;
      copy_dtoh(&w, dst + 5 * sizeof(i64), sizeof w);
      *depth = d;
      *narrow = narrow_word(w, mark, longest, widest);
      return failed() ? -1 : bytes;
    }
  }
  std::vector<i64> off;
  std::vector<int> col;
  std::vector<double> coef;
  if (!describe_sweep(prog, n, off, col, coef)) return -1;
  const i64 bound = mpk_sweep_plan_bound(n, off[n], MPK_BLK_ROWS);
  if (bound > cap) {
    set_error(5, "sweep plan buffer too small");
    return -1;
  }
  // the planner writes every byte of its plan (it clears the buffer first)
  std::unique_ptr<int64_t[]> host(new int64_t[(size_t)(bound + 7) / 8]);
  const i64 bytes = mpk_sweep_plan(n, forward ? 1 : 0, (int64_t)off.data(), (int64_t)col.data(),
                                   (int64_t)coef.data(), MPK_BLK_ROWS, (int64_t)host.get(), bound);
  if (bytes < 0) {
    set_error(5, "sweep plan larger than its bound");
    return -1;
  }
  copy_htod(dst, host.get(), bytes);
  *depth = host[3];
  // narrow: its slot count; wave and wave blocks: the count plus 16, 32 when
  // no row has over 2 entries, 64 log2 of its waves (one per 64 rows of its
  // widest level), and 256 for wave blocks
  int longest = 0, widest = 0;
  if (host[7] == 2 || host[7] == 3) {
    const int *ip = reinterpret_cast<const int *>(host.get() + 8);
    const int *lvl = ip + 2 * (host[0] + 1), *pcnt = lvl + host[2] + n;
    for (i64 p = 0; p < n; p++) longest = pcnt[p] > longest ? pcnt[p] : longest;
    for (i64 l = 0; l + 1 < host[2]; l++) widest = lvl[l + 1] - lvl[l] > widest ? lvl[l + 1] - lvl[l] : widest;
  }
  *narrow = narrow_word(host[5], host[7], longest, widest);
  return failed() ? -1 : bytes;
}

// Entry count of one sweep (for the plan bound): the counting pass of
// k_sweep_describe only.
template <class Prog>
static i64 sweep_entries(const Prog &prog, i64 n) {
  int *cnt_d = nullptr;
  std::vector<int> cnt(n);
  HIPCHK(hipMalloc(&cnt_d, n * sizeof(int)));
  if (failed()) return -1;
  MPK_LAUNCH(k_sweep_describe<Prog>, nblocks(n, 256), 256, prog, n, 0, cnt_d, (const i64 *)nullptr,
             (int *)nullptr, (double *)nullptr);
  HIPCHK(hipMemcpyAsync(cnt.data(), cnt_d, n * sizeof(int), hipMemcpyDeviceToHost, stream()));
  HIPCHK(hipStreamSynchronize(stream()));
  HIPCHK(hipFree(cnt_d));
  if (failed()) return -1;
  i64 m = 0;
  for (i64 i = 0; i < n; i++) m += cnt[i];
  return m;
}

// ------------------------------------------------------------ row programs (mpk_progs.h)

// k_icc_late: icc_late_row for every row.
__global__ void k_icc_late(i64 n, const int *trp, const int *trow, const int *tpos, const int *tmid,
                           const double *ua, const double *Y, double *F, const int64_t *stop) {
  if (stop && *stop) return;
  i64 k = (i64)blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  icc_late_row(k, trp, trow, tpos, tmid, ua, Y, F);
}

// MatInvertDiagonalForSOR_SeqAIJ (ported.mojo:610-630)
__global__ void k_sor_idiag(i64 n, const double *a, const int *dmark, double *idiag, double *mdiag,
                            double omega, double fshift, const int64_t *stop) {
  if (stop && *stop) return;
  i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double d = a[dmark[i]];
  mdiag[i] = d;
  idiag[i] = (omega == 1.0 && fshift <= 0.0) ? 1.0 / d : omega / (fshift + d);
}

// ------------------------------------------------------------ applies

// Three published-value buffers of n, then (planned sweeps) the staged rows,
// 32 bytes each from a 16-byte boundary.
static double *sweep_scratch(i64 n, bool planned) {
  return scratch_vector((size_t)(planned ? 7 : 3) * (size_t)n + 1);
}
static Staged *staged_rows(double *S, i64 n) { return reinterpret_cast<Staged *>(S + 3 * n + (n & 1)); }
double *pc_sweep_scratch(i64 n) { return sweep_scratch(n, true); }
void *pc_staged_rows(double *S, i64 n) { return staged_rows(S, n); }
int plan_wave_form(const char *plan, int second) { return plan ? wave_form(plan_narrow(plan, second)) : -1; }

void ilu_apply(const IluFactor &f, const double *b, double *x, const int64_t *stop) {
  i64 n = f.n;
  double *S = sweep_scratch(n, f.plan != nullptr);
  if (!S || n <= 0) return;
  double *F = S, *B = S + n;
  Staged *st = staged_rows(S, n);
  run_sweep(IluFwd{f.bi, f.bj, f.r, f.ba, b}, n, true, F, stop, f.plan, 0, st);
  run_sweep(IluBwd{f.bj, f.bdiag, f.c, f.ba, F, x}, n, false, B, stop, f.plan, 1, st);
}

void icc_apply(const IccFactor &f, const double *b, double *x, const int64_t *stop) {
  i64 n = f.n;
  double *S = sweep_scratch(n, f.plan != nullptr);
  if (!S || n <= 0) return;
  double *Y = S, *F = S + n, *B = S + 2 * n;
  Staged *st = staged_rows(S, n);
  run_sweep(IccFwd{f.trp, f.trow, f.tpos, f.tmid, f.udiag, f.c, f.ua, b, F}, n, true, Y, stop, f.plan, 0, st);
  if (f.nlate) MPK_LAUNCH(k_icc_late, nblocks(n, 256), 256, n, f.trp, f.trow, f.tpos, f.tmid, f.ua, Y, F, stop);
  run_sweep(IccBwd{f.ui, f.uj, f.udiag, f.c, f.ua, F, x, n - 1}, n, false, B, stop, f.plan, 1, st);
}

void sor_apply(const SorData &s, const double *x, double *y, const int64_t *stop) {
  i64 n = s.n;
  double *S = sweep_scratch(n, s.plan != nullptr);
  if (!S || n <= 0) return;
  Staged *st = staged_rows(S, n);
  MPK_LAUNCH(k_sor_idiag, nblocks(n, 256), 256, n, s.a, s.dmark, s.idiag, s.mdiag, s.omega, s.fshift, stop);
  const bool fwd = (s.flag & (1 | 4)) != 0, bwd = (s.flag & (2 | 8)) != 0;
  double *YF = S, *YB = S + n;
  if (fwd)
    run_sweep(SorFwd{s.rp, s.ci, s.dmark, s.a, x, s.idiag, s.t, bwd ? nullptr : y, y}, n, true, YF, stop,
              s.plan, 0, st);
  if (bwd)
    run_sweep(SorBwd{s.rp, s.ci, s.dmark, s.a, fwd ? s.t : x, s.idiag, YF, y, s.omega, fwd ? 1 : 0}, n,
              false, YB, stop, s.plan, 1, st);
  // its * lits > 1 never reaches here: the apply is its = lits = 1 (capi.mojo:272-280)
}

void PCView::apply(const double *x, double *y, const int64_t *stop) const {
  switch (type) {
    case PCT_JACOBI:
      vpointwise_mult(y, x, jac_diag, n, stop);
      break;
    case PCT_ILU:
      ilu_apply(ilu, x, y, stop);
      break;
    case PCT_ICC:
      icc_apply(icc, x, y, stop);
      break;
    case PCT_SOR:
      sor_apply(sor, x, y, stop);
      break;
    default:
      vcopy(x, y, n);
  }
}

// The two sweeps of a PC as row programs (vectors unset: the plan only needs
// their entries).
template <class F>
static bool with_sweeps(const PCView &pc, F &&f) {
  switch (pc.type) {
    case PCT_ILU:
      f(IluFwd{pc.ilu.bi, pc.ilu.bj, pc.ilu.r, pc.ilu.ba, nullptr}, true);
      f(IluBwd{pc.ilu.bj, pc.ilu.bdiag, pc.ilu.c, pc.ilu.ba, nullptr, nullptr}, false);
      return true;
    case PCT_ICC:
      f(IccFwd{pc.icc.trp, pc.icc.trow, pc.icc.tpos, pc.icc.tmid, pc.icc.udiag, pc.icc.c, pc.icc.ua, nullptr,
               nullptr},
        true);
      f(IccBwd{pc.icc.ui, pc.icc.uj, pc.icc.udiag, pc.icc.c, pc.icc.ua, nullptr, nullptr, pc.n - 1}, false);
      return true;
    case PCT_SOR:
      f(SorFwd{pc.sor.rp, pc.sor.ci, pc.sor.dmark, pc.sor.a, nullptr, nullptr, nullptr, nullptr, nullptr}, true);
      f(SorBwd{pc.sor.rp, pc.sor.ci, pc.sor.dmark, pc.sor.a, nullptr, nullptr, nullptr, nullptr, 1.0, 0}, false);
      return true;
    default:
      return false;
  }
}

bool pc_sweep_entries(const PCView &pc, int which, std::vector<i64> &off, std::vector<int> &col,
                      std::vector<double> &coef) {
  int k = 0;
  bool ok = false;
  with_sweeps(pc, [&](auto prog, bool) {
    if (k++ == which) ok = describe_sweep(prog, pc.n, off, col, coef);
  });
  return ok;
}

i64 PCView::schedule_size() const {
  i64 total = 0;
  bool bad = false;
  with_sweeps(*this, [&](auto prog, bool) {
    const i64 m = sweep_entries(prog, n);
    if (m < 0) bad = true;
    total += mpk_sweep_plan_bound(n, m, MPK_BLK_ROWS);
  });
  return bad ? -1 : total;
}

i64 PCView::schedule(char *plan, i64 cap) const {
  i64 used = 0, depth = 0;
  bool bad = false;
  int which = 0;
  PlanNote note;
  const bool any = with_sweeps(*this, [&](auto prog, bool forward) {
    if (bad) return;
    i64 d = 0;
    int narrow = 0;
    const i64 bytes = plan_sweep(prog, n, forward, plan + used, cap - used, &d, &narrow);
    if (bytes < 0) {
      bad = true;
      return;
    }
    int64_t nblk = 0;  // the plan header's block count
    copy_dtoh(&nblk, plan + used, sizeof nblk);
    if (which < 2) note.nblk[which] = nblk > 0 && nblk < (1 << 30) ? (int)nblk : 0;
    note.bits |= narrow << (16 * which++);
    used += bytes;
    depth = d > depth ? d : depth;
  });
  if (!any) return 0;
  if (bad) note = PlanNote{};
  note_narrow(plan, note);
  return bad ? -1 : depth;
}

// PCApplySymmetric (ksp.mojo:201-208).  The Jacobi square root is read from
// the descriptor's aux1 slot, where the set up stores it.
void PCView::apply_symmetric(const double *x, double *y, const int64_t *stop) const {
  if (type == PCT_JACOBI)
    vpointwise_mult(y, x, jac_diagsqrt, n, stop);
  else
    apply(x, y, stop);
}

// ------------------------------------------------------------ ILU numeric

// Position of column c in row i of an ILU factor (the L part's columns
// ascending in [l0, l1), the strict U part's ascending in [u0, dg), the
// diagonal at dg), or -1 when the row's pattern has no such column.
__device__ __forceinline__ int ilu_pos(const int *__restrict__ bj, int l0, int l1, int u0, int dg, int i, int c) {
  if (c == i) return dg;
  int lo = c < i ? l0 : u0;
  const int end = c < i ? l1 : dg;
  int hi = end;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (bj[mid] < c)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < end && bj[lo] == c ? lo : -1;
}

// MatILUFactorNumeric_SeqAIJ (ported.mojo:957-1037) for the rows of one
// level, a thread per row, in place in ba: the row's own slots stand in for
// the dense work row (a column outside the row's pattern, which the
// reference writes to and never reads, is skipped).  The same operations in
// the same order: zero the row, load A's row r[i] through ic, then for each
// L pivot p ascending with a nonzero multiplier, mult = w[p] * (1/u_pp),
// w[c] = fma(-mult, u_pc, w[c]) along U's row p, and finally the inverted
// diagonal.  Pivot rows lie in earlier levels, finished by earlier launches.
__global__ void __launch_bounds__(256) k_ilu_numeric(i64 p0, i64 cnt, const int *__restrict__ lrow,
                                                     const int *__restrict__ rp, const int *__restrict__ ci,
                                                     const double *__restrict__ a, const int *__restrict__ r,
                                                     const int *__restrict__ ic, const int *__restrict__ bi,
                                                     const int *__restrict__ bj, const int *__restrict__ bd,
                                                     double *ba) {
  const i64 t = (i64)blockIdx.x * 256 + threadIdx.x;
  if (t >= cnt) return;
  const int i = lrow[p0 + t];
  const int l0 = bi[i], l1 = bi[i + 1], dg = bd[i], u0 = bd[i + 1] + 1;
  for (int j = l0; j < l1; j++) ba[j] = 0.0;
  for (int j = u0; j <= dg; j++) ba[j] = 0.0;
  const int ra = r[i];
  for (int j = rp[ra]; j < rp[ra + 1]; j++) {
    const int q = ilu_pos(bj, l0, l1, u0, dg, i, ic[ci[j]]);
    if (q >= 0) ba[q] = a[j];
  }
  for (int j = l0; j < l1; j++) {
    const int p = bj[j];
    const double pc = ba[j];
    if (pc != 0.0) {
      const double mult = pc * ba[bd[p]];
      ba[j] = mult;
      const int qe = bd[p];
      for (int q = bd[p + 1] + 1; q < qe; q++) {
        const int w = ilu_pos(bj, l0, l1, u0, dg, i, bj[q]);
        if (w >= 0) ba[w] = fma(-mult, ba[q], ba[w]);
      }
    }
  }
  ba[dg] = 1.0 / ba[dg];
}

// ------------------------------------------------------------ IC numeric

// Cell of column c in row k's work row: the slot of c among the strict
// columns (ascending in [s0, dg), some possibly below the diagonal), else the
// diagonal slot dg when c is its column (the pattern stores there the row's
// unpermuted index, which is k in the natural ordering), else -2 for column
// k itself when the diagonal slot holds another column, else -1 (a column
// outside the pattern).
__device__ __forceinline__ int icc_cell(const int *__restrict__ bj, int s0, int dg, int k, int c) {
  int lo = s0, hi = dg;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (bj[mid] < c)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo < dg && bj[lo] == c) return lo;
  if (c == bj[dg]) return dg;
  return c == k ? -2 : -1;
}

// MatICCFactorNumeric_SeqAIJ (ported.mojo:1040-1124) for the rows of one
// level, a thread per row.  The reference's dense work row becomes row k's
// own slots of ua, plus a register for column k when the diagonal slot holds
// another column (columns outside the pattern, which the reference writes to
// and never reads, are skipped).  The same operations in the same order:
// zero the row's slots, load A's row rip[k] on and above the diagonal, dk =
// column k's value, then for each visit (i, f) in the chains' order
// (mpk_icc_visits): uikdi = -u_f * (1/d_i), dk = fma(uikdi, u_f, dk), slot f
// takes uikdi, and row i's later slots (its diagonal slot included, as the
// reference's loop includes it) update the work row; finally the row's
// values and 1/dk.  Visits read a row's values as it finished them from u0,
// since rows of one level scale entries of a shared earlier row.
__global__ void __launch_bounds__(256) k_icc_numeric(i64 p0, i64 cnt, const int *__restrict__ lrow,
                                                     const int *__restrict__ rp, const int *__restrict__ ci,
                                                     const double *__restrict__ a, const int *__restrict__ rip,
                                                     const int *__restrict__ riip, const int *__restrict__ bi,
                                                     const int *__restrict__ bj, const int *__restrict__ vptr,
                                                     const int *__restrict__ vrow, const int *__restrict__ vpos,
                                                     double *u0, double *ua) {
  const i64 t = (i64)blockIdx.x * 256 + threadIdx.x;
  if (t >= cnt) return;
  const int k = lrow[p0 + t];
  const int s0 = bi[k], dg = bi[k + 1] - 1;
  for (int j = s0; j <= dg; j++) ua[j] = 0.0;
  double wk = 0.0;  // column k's cell when it has no slot (A's diagonal is then present)
  const int ra = rip[k];
  for (int j = rp[ra]; j < rp[ra + 1]; j++) {
    const int c = riip[ci[j]];
    if (c >= k) {
      const int q = icc_cell(bj, s0, dg, k, c);
      if (q >= 0)
        ua[q] = a[j];
      else if (q == -2)
        wk = a[j];
    }
  }
  const int qk = icc_cell(bj, s0, dg, k, k);
  double dk = qk >= 0 ? ua[qk] : wk;
  for (int v = vptr[k]; v < vptr[k + 1]; v++) {
    const int i = vrow[v], f = vpos[v];
    const double uf = u0[f];
    const int di = bi[i + 1] - 1;
    const double uikdi = -uf * u0[di];
    dk = fma(uikdi, uf, dk);
    ua[f] = uikdi;
    for (int j = f + 1; j <= di; j++) {
      const int q = icc_cell(bj, s0, dg, k, bj[j]);
      if (q >= 0) ua[q] = fma(uikdi, u0[j], ua[q]);
    }
  }
  for (int j = s0; j < dg; j++) u0[j] = ua[j];
  const double inv = 1.0 / dk;
  u0[dg] = inv;
  ua[dg] = inv;
}

// ------------------------------------------------------------ IC transpose map

// Transpose of the strict upper part of an IC factor: for column k, the rows
// i < k holding it, ascending, and the positions of those entries in ua.
static bool build_icc_map(const int *ui, const int *uj, const int *udiag, i64 n, int *trp_d, int *trow_d,
                          int *tpos_d, int *tmid_d, i64 *nlate) {
  Ctx *c = ctx();
  if (!c) return false;
  std::vector<int> hui(n + 1), hud(n);
  copy_dtoh(hui.data(), ui, (n + 1) * sizeof(int));
  copy_dtoh(hud.data(), udiag, n * sizeof(int));
  i64 nnz = n > 0 ? hui[n] : 0;
  std::vector<int> huj(nnz > 0 ? nnz : 1);
  if (nnz > 0) copy_dtoh(huj.data(), uj, nnz * sizeof(int));
  std::vector<int> trp(n + 1, 0);
  for (i64 i = 0; i < n; i++)
    for (int j = hui[i]; j < hud[i]; j++) trp[huj[j] + 1]++;
  for (i64 k = 0; k < n; k++) trp[k + 1] += trp[k];
  i64 m = trp[n];
  std::vector<int> trow(m > 0 ? m : 1), tpos(m > 0 ? m : 1), fill(trp.begin(), trp.end() - 1);
  for (i64 i = 0; i < n; i++)
    for (int j = hui[i]; j < hud[i]; j++) {
      int p = fill[huj[j]]++;
      trow[p] = (int)i;
      tpos[p] = j;
    }
  std::vector<int> tmid(n > 0 ? n : 1);
  i64 late = 0;
  for (i64 k = 0; k < n; k++) {
    int j = trp[k];
    while (j < trp[k + 1] && trow[j] < k) j++;
    tmid[k] = j;
    late += trp[k + 1] - j;
  }
  copy_htod(trp_d, trp.data(), (n + 1) * sizeof(int));
  if (m > 0) {
    copy_htod(trow_d, trow.data(), m * sizeof(int));
    copy_htod(tpos_d, tpos.data(), m * sizeof(int));
  }
  if (n > 0) copy_htod(tmid_d, tmid.data(), n * sizeof(int));
  if (nlate) *nlate = late;
  return !failed();
}

i64 icc_map_count(const int *ui, const int *udiag, i64 n) {
  std::vector<int> hui(n + 1), hud(n);
  copy_dtoh(hui.data(), ui, (n + 1) * sizeof(int));
  copy_dtoh(hud.data(), udiag, n * sizeof(int));
  i64 m = 0;
  for (i64 i = 0; i < n; i++) m += hud[i] - hui[i];
  return m;
}

bool PCView::parse(const int64_t *pci, const double *pcf, i64 n_) {
  n = n_;
  type = (int)pci[0];
  switch (type) {
    case PCT_JACOBI:
      jac_diag = (const double *)pci[1];
      jac_diagsqrt = (const double *)pci[2];
      return true;
    case PCT_SOR:
      sor = SorData{n, (const int *)pci[1], (const int *)pci[2], (const double *)pci[3], (const int *)pci[4],
                    (double *)pci[5], (double *)pci[6], (double *)pci[7], pcf[0], pcf[1], (int)pci[10],
                    (const char *)pci[11]};
      return true;
    case PCT_ILU:
      ilu = IluFactor{n, (const int *)pci[1], (const int *)pci[2], (const double *)pci[3], (const int *)pci[4],
                      (const int *)pci[5], (const int *)pci[6], (const char *)pci[11]};
      return true;
    case PCT_ICC: {
      icc.n = n;
      icc.ui = (const int *)pci[1];
      icc.uj = (const int *)pci[2];
      icc.ua = (const double *)pci[3];
      icc.udiag = (const int *)pci[4];
      icc.c = (const int *)pci[5];
      // the transpose of the strict part travels in aux5..aux8 and the count
      // of late entries in slot 10 (mpk_icc_transpose)
      icc.trp = (const int *)pci[6];
      icc.trow = (const int *)pci[7];
      icc.tpos = (const int *)pci[8];
      icc.tmid = (const int *)pci[9];
      icc.nlate = pci[10];
      icc.plan = (const char *)pci[11];
      if (!icc.trp || !icc.trow || !icc.tpos || !icc.tmid) {
        set_error(4, "IC descriptor without the transpose of the factor (aux5..aux8)");
        return false;
      }
      return true;
    }
    default:
      type = PCT_NONE;
      return true;
  }
}

// ------------------------------------------------------------ Jacobi set up

__device__ __forceinline__ double jacobi_fill(int jtype, i64 i, const int *rp, const int *ci,
                                              const double *v) {
  double s = 0.0;
  if (jtype == 0) {
    for (int j = rp[i]; j < rp[i + 1]; j++)
      if (ci[j] == i) {
        s = v[j];
        break;
      }
  } else if (jtype == 1) {
    for (int j = rp[i]; j < rp[i + 1]; j++) {
      double a = fabs(v[j]);
      s = a > s ? a : s;
    }
  } else if (jtype == 2) {
    for (int j = rp[i]; j < rp[i + 1]; j++) s = s + fabs(v[j]);
  } else {
    for (int j = rp[i]; j < rp[i + 1]; j++) s = s + v[j];
  }
  return s;
}

// PCSetUp_Jacobi (ported.mojo:502-541), one row per lane
__global__ void k_jacobi_setup(i64 n, int jtype, int useabs, int fixdiag, double scale, const int *rp,
                               const int *ci, const double *v, double *diag, double *didx) {
  i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double d = jacobi_fill(jtype, i, rp, ci, v);
  if (jtype == 2) {
    double dx = jacobi_fill(0, i, rp, ci, v);
    didx[i] = dx;
    if (dx < 0.0) d = -d;
    if (scale != 1.0) d = fma(scale, d - dx, dx);
  }
  d = d == 0.0 ? 0.0 : 1.0 / d;
  if (useabs) d = fabs(d);
  if (fixdiag && d == 0.0) d = 1.0;
  diag[i] = d;
}

// PCSetUp_Jacobi_Symmetric (ported.mojo:544-562)
__global__ void k_jacobi_setup_sym(i64 n, int jtype, const int *rp, const int *ci, const double *v,
                                   double *dsq) {
  i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double d = jacobi_fill(jtype, i, rp, ci, v);
  if (d < 0.0)
    d = 1.0 / sqrt(fabs(-d));
  else if (d > 0.0)
    d = 1.0 / sqrt(fabs(d));
  else
    d = 1.0;
  dsq[i] = d;
}

}  // namespace mpk

using namespace mpk;

extern "C" {

void mp_pc_jacobi_setup(int64_t diag, int64_t jtype, int64_t useabs, int64_t fixdiag, double scale, int64_t irow,
                        int64_t icol, int64_t data, int64_t nrow, int64_t didx) {
  if (nrow <= 0 || !ctx()) return;
  MPK_LAUNCH(k_jacobi_setup, nblocks(nrow, 256), 256, (i64)nrow, (int)jtype, (int)useabs, (int)fixdiag, scale,
             (const int *)irow, (const int *)icol, (const double *)data, (double *)diag, (double *)didx);
}

void mp_pc_jacobi_setup_symmetric(int64_t diagsqrt, int64_t jtype, int64_t irow, int64_t icol, int64_t data,
                                  int64_t nrow) {
  if (nrow <= 0 || !ctx()) return;
  MPK_LAUNCH(k_jacobi_setup_sym, nblocks(nrow, 256), 256, (i64)nrow, (int)jtype, (const int *)irow,
             (const int *)icol, (const double *)data, (double *)diagsqrt);
}

void mp_pc_jacobi_apply(int64_t y, int64_t x, int64_t diag, int64_t n) {
  vpointwise_mult((double *)y, (const double *)x, (const double *)diag, n, nullptr);
}

void mp_pc_sor_apply(int64_t x, int64_t y, int64_t irow, int64_t icol, int64_t data, int64_t dmark, int64_t idiag,
                     int64_t mdiag, int64_t t, double omega, double fshift, int64_t flag, int64_t nrow) {
  SorData s{nrow, (const int *)irow, (const int *)icol, (const double *)data, (const int *)dmark,
            (double *)idiag, (double *)mdiag, (double *)t, omega, fshift, (int)flag};
  sor_apply(s, (const double *)x, (double *)y, nullptr);
}

void mp_pc_ilu_apply(int64_t b, int64_t x, int64_t tmp, int64_t bi, int64_t bj, int64_t ba, int64_t bdiag,
                     int64_t r, int64_t c, int64_t nrow) {
  (void)tmp;  // the sweeps keep their published values in the library's scratch
  IluFactor f{nrow, (const int *)bi, (const int *)bj, (const double *)ba, (const int *)bdiag, (const int *)r,
              (const int *)c};
  ilu_apply(f, (const double *)b, (double *)x, nullptr);
}

int64_t mp_pc_icc_apply(int64_t b, int64_t x, int64_t ui, int64_t uj, int64_t ua, int64_t udiag, int64_t c,
                        int64_t scratch, int64_t nrow) {
  // the reference signature carries no transpose, so one is built for the
  // call; mpk_icc_transpose + mpk_pc_icc_apply_t keep it instead
  (void)scratch;
  if (nrow <= 0 || !ctx()) return 0;
  i64 m = icc_map_count((const int *)ui, (const int *)udiag, nrow);
  int *trp = nullptr, *trow = nullptr, *tpos = nullptr, *tmid = nullptr;
  i64 nlate = 0;
  HIPCHK(hipMalloc(&trp, (nrow + 1) * sizeof(int)));
  HIPCHK(hipMalloc(&trow, (m > 0 ? m : 1) * sizeof(int)));
  HIPCHK(hipMalloc(&tpos, (m > 0 ? m : 1) * sizeof(int)));
  HIPCHK(hipMalloc(&tmid, nrow * sizeof(int)));
  if (build_icc_map((const int *)ui, (const int *)uj, (const int *)udiag, nrow, trp, trow, tpos, tmid, &nlate)) {
    IccFactor f{nrow, (const int *)ui, (const int *)uj, (const double *)ua, (const int *)udiag, (const int *)c,
                trp, trow, tpos, tmid, nlate};
    icc_apply(f, (const double *)b, (double *)x, nullptr);
  }
  HIPCHK(hipStreamSynchronize(stream()));
  HIPCHK(hipFree(trp));
  HIPCHK(hipFree(trow));
  HIPCHK(hipFree(tpos));
  HIPCHK(hipFree(tmid));
  return 0;
}

// mp_pc_icc_apply with the transpose of the strict part supplied.
int64_t mpk_pc_icc_apply_t(int64_t b, int64_t x, int64_t ui, int64_t uj, int64_t ua, int64_t udiag, int64_t c,
                           int64_t trp, int64_t trow, int64_t tpos, int64_t tmid, int64_t nlate, int64_t nrow) {
  if (nrow <= 0 || !ctx()) return 0;
  IccFactor f{nrow, (const int *)ui, (const int *)uj, (const double *)ua, (const int *)udiag, (const int *)c,
              (const int *)trp, (const int *)trow, (const int *)tpos, (const int *)tmid, nlate};
  icc_apply(f, (const double *)b, (double *)x, nullptr);
  return 0;
}

// The transpose of an IC factor's strict part, for the forward gather:
// trp[n + 1], trow[m], tpos[m], tmid[n] (device, int32) with
// m = mpk_icc_transpose_size.  Returns the number of late entries (>= 0) or
// -1 on failure.
int64_t mpk_icc_transpose_size(int64_t ui, int64_t udiag, int64_t nrow) {
  if (nrow <= 0 || !ctx()) return 0;
  return icc_map_count((const int *)ui, (const int *)udiag, nrow);
}

int64_t mpk_icc_transpose(int64_t ui, int64_t uj, int64_t udiag, int64_t nrow, int64_t trp, int64_t trow,
                          int64_t tpos, int64_t tmid) {
  if (nrow <= 0) return 0;
  i64 nlate = 0;
  return build_icc_map((const int *)ui, (const int *)uj, (const int *)udiag, nrow, (int *)trp, (int *)trow,
                       (int *)tpos, (int *)tmid, &nlate)
             ? nlate
             : -1;
}

// MatILUFactorNumeric_SeqAIJ on the device, one launch per level of the
// factor's L part (lrow / lptr from mpk_ilu_levels).
int64_t mpk_ilu_numeric(int64_t n, int64_t irow, int64_t icol, int64_t data, int64_t r, int64_t ic, int64_t bi,
                        int64_t bj, int64_t bdiag, int64_t lrow, int64_t lptr, int64_t nlev, int64_t ba) {
  if (n <= 0) return 0;
  if (!ctx()) return -1;
  const int64_t *lp = (const int64_t *)lptr;
  if (nlev <= 0 || lp[0] != 0 || lp[nlev] != n) {
    set_error(2, "mpk_ilu_numeric: the level bounds do not cover the rows");
    return -1;
  }
  for (int64_t l = 0; l < nlev; l++) {
    const i64 cnt = lp[l + 1] - lp[l];
    if (cnt < 0) {
      set_error(2, "mpk_ilu_numeric: level bounds out of order");
      return -1;
    }
    if (cnt == 0) continue;
    MPK_LAUNCH(k_ilu_numeric, nblocks(cnt, 256), 256, (i64)lp[l], cnt, (const int *)lrow, (const int *)irow,
               (const int *)icol, (const double *)data, (const int *)r, (const int *)ic, (const int *)bi,
               (const int *)bj, (const int *)bdiag, (double *)ba);
  }
  return failed() ? -1 : 0;
}

// MatICCFactorNumeric_SeqAIJ on the device, one launch per level of the
// chains' visits (vptr / vrow / vpos, lrow / lptr from mpk_icc_visits); u0
// [D] float64[bi[n]] is scratch for the values as each row finished them.
int64_t mpk_icc_numeric(int64_t n, int64_t irow, int64_t icol, int64_t data, int64_t rip, int64_t riip, int64_t bi,
                        int64_t bj, int64_t vptr, int64_t vrow, int64_t vpos, int64_t lrow, int64_t lptr,
                        int64_t nlev, int64_t u0, int64_t ua) {
  if (n <= 0) return 0;
  if (!ctx()) return -1;
  const int64_t *lp = (const int64_t *)lptr;
  if (nlev <= 0 || lp[0] != 0 || lp[nlev] != n) {
    set_error(2, "mpk_icc_numeric: the level bounds do not cover the rows");
    return -1;
  }
  for (int64_t l = 0; l < nlev; l++) {
    const i64 cnt = lp[l + 1] - lp[l];
    if (cnt < 0) {
      set_error(2, "mpk_icc_numeric: level bounds out of order");
      return -1;
    }
    if (cnt == 0) continue;
    MPK_LAUNCH(k_icc_numeric, nblocks(cnt, 256), 256, (i64)lp[l], cnt, (const int *)lrow, (const int *)irow,
               (const int *)icol, (const double *)data, (const int *)rip, (const int *)riip, (const int *)bi,
               (const int *)bj, (const int *)vptr, (const int *)vrow, (const int *)vpos, (double *)u0,
               (double *)ua);
  }
  return failed() ? -1 : 0;
}

// Block plans of the descriptor's two sweeps (see include/mpk.h).
int64_t mpk_pc_schedule_size(int64_t n, int64_t pci, int64_t pcf) {
  if (n <= 0 || !ctx()) return 0;
  PCView pc;
  if (!pc.parse((const int64_t *)pci, (const double *)pcf, n)) return -1;
  return pc.schedule_size();
}

int64_t mpk_pc_schedule(int64_t n, int64_t pci, int64_t pcf, int64_t plan, int64_t cap) {
  if (n <= 0 || !ctx()) return 0;
  PCView pc;
  if (!pc.parse((const int64_t *)pci, (const double *)pcf, n)) return -1;
  return pc.schedule((char *)plan, cap);
}

// y = M^-1 x for the preconditioner a descriptor describes (the apply the
// solvers use, including its level orders).
void mpk_pc_apply(int64_t n, int64_t pci, int64_t pcf, int64_t x, int64_t y) {
  if (n <= 0 || !ctx()) return;
  PCView pc;
  if (!pc.parse((const int64_t *)pci, (const double *)pcf, n)) return;
  pc.apply((const double *)x, (double *)y, nullptr);
}

}  // extern "C"
