// mpk_pc.h -- device preconditioner descriptors shared by the PC exports and
// the Krylov drivers.
#pragma once

#include <vector>

#include "mpk_internal.h"

namespace mpk {

// PC type ordinals of the descriptor block (petscksp.h PCType, PC.py:23-27)
enum { PCT_NONE = 0, PCT_JACOBI = 2, PCT_SOR = 4, PCT_ILU = 5, PCT_ICC = 6 };

// ILU factor in the reference layout (src/factor.mojo:31-67, 335-381): L rows
// ascending, U rows in reverse row order each ending on its inverted diagonal.
struct IluFactor {
  i64 n;
  const int *bi, *bj;
  const double *ba;
  const int *bdiag, *r, *c;
  const char *plan = nullptr;  // block plans of the two sweeps (mpk_pc_schedule), or null
};

// IC factor in the reference layout (factor.mojo:125-273): row i holds its
// strict columns then the diagonal slot (1/d_i); plus the transpose of the
// strict part for the forward gather.  With level of fill >= 1 the
// reference's pattern can hold columns below the diagonal in a row; their
// forward contributions land after the target row is final ("late" entries,
// trow > k, the tail of each transposed column from tmid[k]); nlate counts
// them.
struct IccFactor {
  i64 n;
  const int *ui, *uj;
  const double *ua;
  const int *udiag, *c;
  const int *trp, *trow, *tpos, *tmid;
  i64 nlate;
  const char *plan = nullptr;  // block plans of the two sweeps (mpk_pc_schedule), or null
};

struct SorData {
  i64 n;
  const int *rp, *ci;
  const double *a;
  const int *dmark;
  double *idiag, *mdiag, *t;
  double omega, fshift;
  int flag;
  const char *plan = nullptr;  // block plans of the two sweeps (mpk_pc_schedule), or null
};

void ilu_apply(const IluFactor &f, const double *b, double *x, const int64_t *stop);
void icc_apply(const IccFactor &f, const double *b, double *x, const int64_t *stop);
void sor_apply(const SorData &s, const double *x, double *y, const int64_t *stop);

// A preconditioner as the Krylov drivers see it: parsed from the descriptor
// block `pci` (device addresses, int32 index arrays) and `pcf` (omega, fshift).
struct PCView {
  int type = PCT_NONE;
  i64 n = 0;
  const double *jac_diag = nullptr, *jac_diagsqrt = nullptr;
  SorData sor{};
  IluFactor ilu{};
  IccFactor icc{};
  bool parse(const int64_t *pci, const double *pcf, i64 n);
  void apply(const double *x, double *y, const int64_t *stop) const;            // y = M^-1 x
  void apply_symmetric(const double *x, double *y, const int64_t *stop) const;  // half of M^-1
  // block plans of the PC's two sweeps: bytes needed, then the plans into
  // plan (cap bytes; returns the deepest block's level count, 0 for a PC
  // without sweeps, -1 on error)
  i64 schedule_size() const;
  i64 schedule(char *plan, i64 cap) const;
};

// The one-workgroup form of sweep `second` of a PC's plans, as the process
// recorded it when it wrote them (wave_form in mpk_sweep.h: 0..2 for a wave
// plan, -1 otherwise or for a plan it did not write).
int plan_wave_form(const char *plan, int second);
// The scratch of a PC's planned sweeps (published values, then the Staged
// records): 7 n + 1 doubles, and where its Staged records start.
double *pc_sweep_scratch(i64 n);
void *pc_staged_rows(double *S, i64 n);

// The host planner's shape rules (mpk_factor.cpp): rows per block, fixed
// slots per position, plan bytes, the MPK_SWEEP_PLAN mode (0 auto, 1 blocks,
// 2 ring), the ring's level-width limit, an environment switch set to "0",
// and a wave record's bytes.
int64_t plan_rows_rule(int64_t n, int64_t bmax);
int64_t plan_w_rule(int64_t widest, int64_t n, int64_t m);
int64_t plan_bytes_rule(int64_t n, int64_t nblk, int64_t nlvl, int64_t slots);
int plan_mode_rule();
int64_t plan_ring_width_rule();
bool plan_env_off(const char *name);
int64_t wave_rec_bytes_rule(int64_t w);

// A sweep's entries on the device, in each row's own order: off [n+1] int64,
// col int32, coef float64 (m entries), and the longest row.
struct DevEntries {
  i64 n, m;
  const i64 *off;
  const int *col;
  const double *coef;
  i64 widest;
};
// The plan of one sweep built on the device into dst [D] (cap bytes),
// byte for byte mpk_sweep_plan's, for the block form.  Returns its bytes
// (and depth, longest row, widest level, hdr[7] through the pointers), 0
// when the sweep takes a ring plan (which the host planner builds), -1 on
// error.
i64 device_sweep_plan(const DevEntries &E, bool forward, i64 bmax, char *dst, i64 cap, i64 *depth,
                      int *longest, int *widest_level, int *mark);

// Every entry of sweep `which` (0: forward / first, 1: second) of a PC with
// sweeps, in each row's own order, on the host: off[n+1], col, coef.
bool pc_sweep_entries(const PCView &pc, int which, std::vector<i64> &off, std::vector<int> &col,
                      std::vector<double> &coef);

}  // namespace mpk
