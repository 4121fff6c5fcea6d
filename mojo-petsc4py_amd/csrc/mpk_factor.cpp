// mpk_factor.cpp -- host-side set up of the incomplete factorisations.
//
// The symbolic phases are integer work on the host.  They must produce the
// reference's fill pattern entry for entry, so they follow its steps
// (src/factor.mojo, with the sorted fill list of src/ported.mojo:759-861):
// the same list walk, the same level rule, the same drain order.  The numeric
// factorisations run here too and give the same bits as
// MatILUFactorNumeric_SeqAIJ / MatICCFactorNumeric_SeqAIJ (ported.mojo:957-1124);
// the factor is then uploaded for the device sweeps.
//
// All arrays in this file are host arrays; index arrays are int64, as the
// reference's Python layer allocates them (PC.py:376-492).
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <vector>

#include "../../include/mpk.h"

namespace {

typedef int64_t I;

// The sorted fill list: a singly linked list threaded through `next`, headed
// at slot `head` (= n), with a fill level per entry and a membership bit.
struct FillList {
  I *next, *lvl, *mark;
  I head;

  // PetscLLInsertLocation: walk from `from` to the insertion point of `entry`
  // and splice it in; returns the new resume point (the entry itself).
  I splice(I from, I entry) {
    I loc = from, succ;
    for (;;) {
      succ = next[loc];
      if (entry > succ)
        loc = succ;
      else
        break;
    }
    next[loc] = entry;
    next[entry] = succ;
    return entry;
  }
  // mark `entry`, reporting whether it was already present
  bool test_and_mark(I entry) {
    bool was = mark[entry] != 0;
    mark[entry] = 1;
    return was;
  }
  // IncompleteLLInit: the columns of a row of A, permuted, at level 0.
  I seed(I count, const I *cols, const I *perm) {
    I added = 0, resume = head;
    for (I k = 0; k < count; k++) {
      I e = perm[cols[k]];
      if (!test_and_mark(e)) {
        I from = (k != 0 && e < resume) ? head : resume;
        resume = splice(from, e);
        added++;
        lvl[e] = 0;
      }
    }
    return added;
  }
  // IncompleteLLAdd with sorted input: merge a pivot row's entries whose
  // level lvl_k + offset + 1 stays within `levels`; keep the smaller level of
  // an entry already present.
  I merge(I count, const I *cols, const I *clvl, I levels, I offset) {
    I added = 0, resume = head;
    for (I k = 0; k < count; k++) {
      I inc = clvl[k] + offset + 1;
      if ((double)inc <= (double)levels) {
        I e = cols[k];
        if (!test_and_mark(e)) {
          resume = splice(resume, e);
          added++;
          lvl[e] = inc;
        } else if (lvl[e] > inc) {
          lvl[e] = inc;
        }
      }
    }
    return added;
  }
  // IncompleteLLClean: write out `count` entries in list order, reset.
  void drain(I count, I *out, I *outlvl) {
    I at = head;
    for (I j = 0; j < count; j++) {
      at = next[at];
      out[j] = at;
      outlvl[j] = lvl[at];
      lvl[at] = -1;
      mark[at] = 0;
    }
    next[head] = head;
  }
};

}  // namespace

// count entries from src to dst; a count below one copies nothing (the
// reference's copies are `while k < count` loops, factor.mojo:31-67,335-381,
// which a row without a diagonal drives negative)
static inline void copy_ids(I *dst, const I *src, I count) {
  if (count > 0) memcpy(dst, src, sizeof(I) * (size_t)count);
}

// hdr[7] marks the plans the device runs with one workgroup per block,
// every level of a block in turn:
//   1, narrow (k_sweep_narrow): one block (a single-block plan or a ring), 4
//      or 8 fixed slots, at most 256 rows a level and 4095 levels, and no
//      entry reading another block's published value (d in [-n, 0));
//   2, wave (k_sweep_wave): as narrow, with 12288 levels at most and no
//      stale entry (d < -n); one to four waves run every level, a row per
//      lane (a wave per 64 rows of the widest level);
//   3, wave blocks (k_sweep_wblk): as wave, for plans of several blocks or
//      with entries reading published values (earlier blocks' rows, or a
//      ring's old ones): the wave kernel per block, with those values loaded
//      ahead.
// Plans 2 and 3 carry records after the plan (at plan_bytes, a multiple of
// 16): per position, the w coefficients (-0.0 past the row's count), the w
// dependencies (edep, kWaveZero past the count: an LDS slot that holds 0.0,
// so that a padded entry adds fma(-0.0, 0.0, acc) == acc exactly), the row
// and the count, in wave_rec_bytes(w) bytes.
// MPK_SWEEP_NARROW=0 leaves every plan unmarked; MPK_SWEEP_WAVE=0 marks no
// plan 2 or 3.  Returns the plan's size with the records.
static constexpr int kWaveZero = 8192;
static int64_t wave_rec_bytes(int64_t w) { return (12 * w + 8 + 15) & ~(int64_t)15; }
static bool env_off(const char *name) {
  const char *e = std::getenv(name);
  return e && e[0] == '0' && !e[1];
}
static int64_t mark_plan(char *out, int64_t n, int64_t bytes, int64_t cap) {
  int64_t *hdr = (int64_t *)out;
  hdr[7] = 0;
  if (env_off("MPK_SWEEP_NARROW")) return bytes;
  const int64_t nblk = hdr[0], nlvl = hdr[2], w = hdr[5];
  if (nblk < 1 || (w != 4 && w != 8)) return bytes;
  const int *blk_pos = (const int *)(out + 64), *blk_lev = blk_pos + nblk + 1, *lvl = blk_lev + nblk + 1;
  int widest = 0;
  int64_t levels = 0;  // a block's most
  for (int64_t b = 0; b < nblk; b++) {
    for (int64_t l = blk_lev[b]; l + 1 < blk_lev[b + 1]; l++) widest = std::max(widest, lvl[l + 1] - lvl[l]);
    levels = std::max<int64_t>(levels, blk_lev[b + 1] - 1 - blk_lev[b]);
  }
  const int64_t ints = 2 * (nblk + 1) + nlvl + 3 * n + 1;
  const double *ecoef = (const double *)(out + 64 + ((4 * ints + 7) & ~(int64_t)7));
  const int *edep = (const int *)(ecoef + w * n);
  const int *prow = lvl + nlvl, *pcnt = prow + n;
  bool stale = false, published = false;
  for (int64_t p = 0; p < n; p++)
    for (int64_t k = 0; k < pcnt[p]; k++) {
      const int d = edep[k * n + p];
      published |= d < 0 && d >= -n;
      stale |= d < -n;
    }
  const bool narrow = nblk == 1 && !published && widest <= 256 && levels < 4096;  // levels + 1 bounds held
  const int64_t rb = wave_rec_bytes(w), wbytes = bytes + n * rb;
  const bool wave = widest <= 256 && levels <= 12288 && !stale && wbytes <= cap && !env_off("MPK_SWEEP_WAVE");
  if (!wave) {
    hdr[7] = narrow ? 1 : 0;
    return bytes;
  }
  std::memset(out + bytes, 0, (size_t)(wbytes - bytes));  // the records' padding too
  for (int64_t p = 0; p < n; p++) {
    char *r = out + bytes + p * rb;
    double *c = (double *)r;
    int *d = (int *)(c + w);
    for (int64_t e = 0; e < w; e++) {
      const bool used = e < pcnt[p];
      c[e] = used ? ecoef[e * n + p] : -0.0;
      d[e] = used ? edep[e * n + p] : kWaveZero;
    }
    d[w] = prow[p];
    d[w + 1] = pcnt[p];
  }
  hdr[4] = wbytes;
  hdr[7] = nblk > 1 || published ? 3 : 2;
  return wbytes;
}

extern "C" {

// factor.mojo:31-67 -- ILU(0), natural ordering: L = strict lower part of A,
// U rows (strict upper then the diagonal) stored from the last row down.
int64_t mp_ilu_symbolic_ilu0(int64_t bi_, int64_t bj_, int64_t bdiag_, int64_t adiag_, int64_t irow_,
                             int64_t icol_, int64_t nrow) {
  I *bi = (I *)bi_, *bj = (I *)bj_, *bdiag = (I *)bdiag_;
  const I *adiag = (const I *)adiag_, *rp = (const I *)irow_, *ci = (const I *)icol_;
  I w = 0;
  bi[0] = 0;
  for (I i = 0; i < nrow; i++) {
    I nl = adiag[i] - rp[i];
    bi[i + 1] = bi[i] + nl;
    copy_ids(bj + w, ci + rp[i], nl);
    w += nl > 0 ? nl : 0;
  }
  bdiag[nrow] = bi[nrow] - 1;
  for (I i = nrow - 1; i >= 0; i--) {
    // a row without a diagonal (adiag at the row end) has nu = -1: nothing
    // is copied and its diagonal slot is the next row's (factor.mojo:57-66)
    I nu = rp[i + 1] - adiag[i] - 1;
    copy_ids(bj + w, ci + adiag[i] + 1, nu);
    w += nu > 0 ? nu : 0;
    bj[w++] = i;
    bdiag[i] = bdiag[i + 1] + nu + 1;
  }
  return w;
}

// factor.mojo:70-122 -- ILU(k): rows staged contiguously (L, diagonal, U),
// bdiag[i] = number of L entries; negative return = capacity needed.
int64_t mp_ilu_symbolic(int64_t bi_, int64_t bj_, int64_t bjlvl_, int64_t bdiag_, int64_t irow_, int64_t icol_,
                        int64_t r_, int64_t ic_, int64_t levels, int64_t cap, int64_t scratch_, int64_t nrow) {
  I *bi = (I *)bi_, *bj = (I *)bj_, *bjlvl = (I *)bjlvl_, *bdiag = (I *)bdiag_;
  const I *rp = (const I *)irow_, *ci = (const I *)icol_, *r = (const I *)r_, *ic = (const I *)ic_;
  I *scr = (I *)scratch_;
  const I n1 = nrow + 1;
  FillList L{scr, scr + n1, scr + 2 * n1, nrow};
  L.next[nrow] = nrow;
  bi[0] = 0;
  bdiag[0] = 0;
  for (I i = 0; i < nrow; i++) {
    const I row = r[i];
    L.next[i] = -1;  // marker: the diagonal has not been seen yet
    I nz = L.seed(rp[row + 1] - rp[row], ci + rp[row], ic);
    I npiv = 0;
    for (I p = L.next[nrow]; p < i; p = L.next[p]) {
      const I nl = bdiag[p];
      const I start = bi[p] + nl + 1;
      const I cnt = bi[p + 1] - bi[p] - nl - 1;
      nz += L.merge(cnt, bj + start, bjlvl + start, levels, L.lvl[p]);
      npiv++;
    }
    bdiag[i] = npiv;
    if (bi[i] + nz > cap) return -(bi[i] + nz);
    bi[i + 1] = bi[i] + nz;
    L.drain(nz, bj + bi[i], bjlvl + bi[i]);
  }
  return bi[nrow];
}

// factor.mojo:125-273 -- IC(k).  Each row is drained into the staging block
// at its start and then written out with the strict columns ascending and
// the diagonal last.  Pivot rows are read back from the staging block at the
// offset the row's position implies (factor.mojo:204-206), which holds the
// most recently drained row: the reference's pattern is defined by that read.
int64_t mp_icc_symbolic(int64_t ui_, int64_t uj_, int64_t ujlvl_, int64_t udiag_, int64_t irow_, int64_t icol_,
                        int64_t rip_, int64_t riip_, int64_t levels, int64_t cap, int64_t scratch_, int64_t nrow,
                        int64_t scratchlen) {
  I *ui = (I *)ui_, *uj = (I *)uj_, *ujlvl = (I *)ujlvl_, *udiag = (I *)udiag_;
  const I *rp = (const I *)irow_, *ci = (const I *)icol_, *rip = (const I *)rip_, *riip = (const I *)riip_;
  const I n1 = nrow + 1;
  if (6 * n1 + 2 * cap > scratchlen) return -(6 * n1 + 2 * cap);
  I *scr = (I *)scratch_;
  FillList L{scr, scr + n1, scr + 2 * n1, nrow};
  I *chain = scr + 3 * n1;   // jl: rows waiting on a column, linked
  I *first = scr + 4 * n1;   // il: next unused entry of a pivot row
  I *upper = scr + 5 * n1;   // the strictly upper columns of A's row
  I *stage = scr + 6 * n1, *stagelvl = stage + cap;
  for (I i = 0; i < nrow; i++) {
    chain[i] = nrow;
    first[i] = 0;
  }
  L.next[nrow] = nrow;
  ui[0] = 0;
  for (I k = 0; k < nrow; k++) {
    const I rk = rip[k];
    I nu = 0;
    for (I j = rp[rk]; j < rp[rk + 1]; j++)
      if (riip[ci[j]] > k) upper[nu++] = ci[j];
    I nz = L.seed(nu, upper, riip);
    if (L.mark[rk] == 0) {
      L.splice(nrow, rk);  // an unsorted insert walks from the head
      nz++;
      L.lvl[rk] = 0;
      L.mark[rk] = 1;
    }
    const I uk = ui[k];
    for (I p = chain[k]; p < k;) {
      const I pnext = chain[p];
      const I jmin = first[p] + 1, jmax = ui[p + 1];
      const I off = jmin - ui[p];
      const I *cols = stage + off;
      const I *clvl = stagelvl + off;
      if (jmax - jmin > 0) nz += L.merge(jmax - jmin, cols, clvl, levels, clvl[-1]);
      if (jmin < jmax) {
        first[p] = jmin;
        const I cj = cols[0];
        chain[p] = chain[cj];
        chain[cj] = p;
      }
      p = pnext;
    }
    if (uk + nz > cap) return -(uk + nz);
    L.drain(nz, stage, stagelvl);
    if (nz > 1) {
      const I cj = stage[1];
      chain[k] = chain[cj];
      chain[cj] = k;
      first[k] = uk + 1;
    }
    I at = -1;
    for (I c = 0; c < nz; c++)
      if (stage[c] == rk) {
        at = c;
        break;
      }
    I w = 0;
    for (I c = 0; c < nz; c++)
      if (c != at) {
        uj[uk + w] = stage[c];
        ujlvl[uk + w] = stagelvl[c];
        w++;
      }
    // ascending strict columns (insertion sort: the drained order is only
    // sorted when every merge source was)
    for (I c = 1; c < w; c++) {
      const I key = uj[uk + c];
      I q = c - 1;
      while (q >= 0 && uj[uk + q] > key) {
        uj[uk + q + 1] = uj[uk + q];
        q--;
      }
      uj[uk + q + 1] = key;
    }
    uj[uk + nz - 1] = rk;
    ujlvl[uk + nz - 1] = stagelvl[at];
    udiag[k] = uk + nz - 1;
    ui[k + 1] = uk + nz;
  }
  return ui[nrow];
}

// factor.mojo:279-332 -- reverse Cuthill-McKee by breadth-first sweeps that
// visit neighbours in index order, one reversed component at a time.
void mp_ordering_rcm(int64_t perm_, int64_t level_, int64_t visited_, int64_t queue_, int64_t irow_,
                     int64_t icol_, int64_t nrow) {
  I *perm = (I *)perm_, *level = (I *)level_, *seen = (I *)visited_, *q = (I *)queue_;
  const I *rp = (const I *)irow_, *ci = (const I *)icol_;
  for (I i = 0; i < nrow; i++) {
    level[i] = rp[i + 1] - rp[i];
    seen[i] = 0;
  }
  I out = 0;
  for (I s = 0; s < nrow; s++) {
    if (seen[s]) continue;
    seen[s] = 1;
    q[0] = s;
    I head = 0, tail = 1;
    while (head < tail) {
      const I v = q[head++];
      for (I j = rp[v]; j < rp[v + 1]; j++)
        if (!seen[ci[j]]) {
          seen[ci[j]] = 1;
          q[tail++] = ci[j];
        }
    }
    for (I a = 0, b = tail - 1; a < b; a++, b--) {
      I t = q[a];
      q[a] = q[b];
      q[b] = t;
    }
    copy_ids(perm + out, q, tail);
    out += tail;
  }
}

// factor.mojo:335-381 -- from staged rows to the factor layout: L rows
// ascending, U rows (strict part, then the diagonal) from the end backwards.
void mp_ilu_contiguous_lu(int64_t bjsrc_, int64_t bi_, int64_t bdiag_, int64_t bj_, int64_t n) {
  const I *src = (const I *)bjsrc_;
  I *bi = (I *)bi_, *bdiag = (I *)bdiag_, *bj = (I *)bj_;
  I utop = bi[n];
  I at = 0;
  I prev_l = n ? bdiag[0] : 0;
  for (I row = 0; row < n; row++) {
    const I len = bi[row + 1] - bi[row];
    bi[row] = row == 0 ? 0 : bi[row - 1] + prev_l;
    const I nl = bdiag[row];
    prev_l = nl;
    copy_ids(bj + bi[row], src + at, nl);
    bdiag[row] = utop - 1;
    bj[utop - 1] = row;
    const I nu = len - nl;
    utop -= nu;
    copy_ids(bj + utop, src + at + nl + 1, nu - 1);
    at += len;
  }
  if (n) {
    bi[n] = bi[n - 1] + prev_l;
    bdiag[n] = bdiag[n - 1] - 1;
  }
}

// factor.mojo:384-412 -- copy the staged IC rows out and set the row starts;
// bdiag[n] carries the spare entry the numeric pass reads for the last row.
void mp_icc_contiguous_cholesky(int64_t ujsrc_, int64_t ui_, int64_t udiag_, int64_t bi_, int64_t bdiag_,
                                int64_t bj_, int64_t n) {
  const I *src = (const I *)ujsrc_, *ui = (const I *)ui_, *udiag = (const I *)udiag_;
  I *bi = (I *)bi_, *bdiag = (I *)bdiag_, *bj = (I *)bj_;
  for (I row = 0; row < n; row++)
    for (I j = ui[row]; j <= udiag[row]; j++) bj[j] = src[j];
  if (n) {
    bi[0] = 0;
    for (I row = 0; row < n; row++) bi[row + 1] = udiag[row] + 1;
    bdiag[n] = bi[n];
  }
}

// ported.mojo:957-1037 -- row-wise IKJ elimination with a dense work row.
int64_t mp_pc_ilu_numeric(int64_t ba_, int64_t rtmp_, int64_t irow_, int64_t icol_, int64_t data_, int64_t bi_,
                          int64_t bj_, int64_t bdiag_, int64_t r_, int64_t ic_, int64_t nrow) {
  double *ba = (double *)ba_, *w = (double *)rtmp_;
  const I *rp = (const I *)irow_, *ci = (const I *)icol_, *bi = (const I *)bi_, *bj = (const I *)bj_,
          *bd = (const I *)bdiag_, *r = (const I *)r_, *ic = (const I *)ic_;
  const double *a = (const double *)data_;
  for (I i = 0; i < nrow; i++) {
    const I l0 = bi[i], l1 = bi[i + 1];     // L part of row i
    const I u0 = bd[i + 1] + 1, u1 = bd[i];  // U part (strict), diagonal at bd[i]
    for (I j = l0; j < l1; j++) w[bj[j]] = 0.0;
    for (I j = u0; j <= u1; j++) w[bj[j]] = 0.0;  // strict U and the diagonal slot
    for (I j = rp[r[i]]; j < rp[r[i] + 1]; j++) w[ic[ci[j]]] = a[j];
    for (I j = l0; j < l1; j++) {
      const I p = bj[j];
      const double pc = w[p];
      if (pc != 0.0) {
        const double mult = pc * ba[bd[p]];
        w[p] = mult;
        for (I q = bd[p + 1] + 1; q < bd[p]; q++) w[bj[q]] = fma(-mult, ba[q], w[bj[q]]);
      }
    }
    for (I j = l0; j < l1; j++) ba[j] = w[bj[j]];
    for (I j = u0; j < u1; j++) ba[j] = w[bj[j]];
    ba[bd[i]] = 1.0 / w[i];
  }
  return bd[0] + 1;
}

// Level sets of an ILU factor's L part, for the level-by-level numeric
// factorisation on the device: row i's level is one past the deepest level
// among the rows its L entries name (0 with none), so every pivot row of a
// level lies in an earlier one.  lrow [H] int32[n] lists the rows level by
// level (ascending inside a level); lptr [H] int64[nlev + 1] bounds the
// levels.  Returns the level count, or -2 when the layout is not one the
// device pass reproduces: each row must hold its L columns strictly
// ascending below i, its strict U columns strictly ascending above i, and a
// diagonal slot of its own (bdiag[i] > bdiag[i + 1], bj[bdiag[i]] = i).  The
// level-of-fill pattern of some nonsymmetric matrices leaves rows without
// one; the reference's dense work row then carries values from row to row,
// which only the serial pass (mp_pc_ilu_numeric) reproduces.
int64_t mpk_ilu_levels(int64_t bi_, int64_t bj_, int64_t bdiag_, int64_t n, int64_t lrow_, int64_t lptr_) {
  const I *bi = (const I *)bi_, *bj = (const I *)bj_, *bd = (const I *)bdiag_;
  int32_t *lrow = (int32_t *)lrow_;
  I *lptr = (I *)lptr_;
  if (n <= 0) {
    lptr[0] = 0;
    return 0;
  }
  for (I i = 0; i < n; i++) {
    if (bd[i] <= bd[i + 1] || bj[bd[i]] != i) return -2;
    for (I j = bi[i]; j < bi[i + 1]; j++)
      if (bj[j] < 0 || bj[j] >= i || (j > bi[i] && bj[j] <= bj[j - 1])) return -2;
    for (I j = bd[i + 1] + 1; j < bd[i]; j++)
      if (bj[j] <= i || bj[j] >= n || (j > bd[i + 1] + 1 && bj[j] <= bj[j - 1])) return -2;
  }
  std::vector<int32_t> lev(n);
  int32_t top = 0;
  for (I i = 0; i < n; i++) {
    int32_t l = 0;
    for (I j = bi[i]; j < bi[i + 1]; j++) l = std::max(l, lev[bj[j]] + 1);
    lev[i] = l;
    top = std::max(top, l);
  }
  const I nlev = (I)top + 1;
  std::vector<I> cnt(nlev + 1, 0);
  for (I i = 0; i < n; i++) cnt[lev[i] + 1]++;
  for (I l = 0; l < nlev; l++) cnt[l + 1] += cnt[l];
  for (I l = 0; l <= nlev; l++) lptr[l] = cnt[l];
  for (I i = 0; i < n; i++) lrow[cnt[lev[i]]++] = (int32_t)i;
  return nlev;
}

// ported.mojo:1040-1124 -- up-looking IC with the row chains of each column.
void mp_pc_icc_numeric(int64_t ua_, int64_t rtmp_, int64_t il_, int64_t c2r_, int64_t irow_, int64_t icol_,
                       int64_t data_, int64_t ui_, int64_t uj_, int64_t udiag_, int64_t r_, int64_t riip_,
                       int64_t nrow) {
  double *ua = (double *)ua_, *w = (double *)rtmp_;
  I *first = (I *)il_, *chain = (I *)c2r_;
  const I *rp = (const I *)irow_, *ci = (const I *)icol_, *bi = (const I *)ui_, *bj = (const I *)uj_,
          *bd = (const I *)udiag_, *r = (const I *)r_, *riip = (const I *)riip_;
  const double *a = (const double *)data_;
  for (I i = 0; i < nrow; i++) chain[i] = nrow;
  if (nrow) first[0] = 0;
  for (I k = 0; k < nrow; k++) {
    for (I j = bi[k]; j < bi[k + 1]; j++) w[bj[j]] = 0.0;
    I slot = bi[k];
    for (I j = rp[r[k]]; j < rp[r[k] + 1]; j++) {
      const I col = riip[ci[j]];
      if (col >= k) {
        w[col] = a[j];
        ua[slot++] = 0.0;
      }
    }
    double dk = w[k];
    for (I i = chain[k]; i < k;) {
      const I inext = chain[i];
      const I f = first[i];
      const double uikdi = -ua[f] * ua[bd[i]];
      dk = fma(uikdi, ua[f], dk);
      ua[f] = uikdi;
      const I jmin = f + 1, jmax = bi[i + 1];
      if (jmin < jmax) {
        for (I j = jmin; j < jmax; j++) w[bj[j]] = fma(uikdi, ua[j], w[bj[j]]);
        first[i] = jmin;
        const I cj = bj[jmin];
        chain[i] = chain[cj];
        chain[cj] = i;
      }
      i = inext;
    }
    const I jmin = bi[k], jmax = bi[k + 1] - 1;
    if (jmin < jmax) {
      for (I j = jmin; j < jmax; j++) ua[j] = w[bj[j]];
      first[k] = jmin;
      const I cj = bj[jmin];
      chain[k] = chain[cj];
      chain[cj] = k;
    }
    ua[bd[k]] = 1.0 / dk;
  }
}

// The up-looking IC pass (MatICCFactorNumeric_SeqAIJ, ported.mojo:1040-1124)
// replayed without arithmetic, for the level-by-level numeric pass on the
// device: its row chains (il / c2r) decide which rows i < k hand row k an
// update, and in what order; that order is integer work, done here.  bi, bj
// [H] int64: the factor layout of mp_icc_contiguous_cholesky (row k's
// strict columns, then its diagonal slot at bi[k + 1] - 1); irow, icol [H]
// int64: A; rip, riip [H] int64: the permutation and its inverse.  Out
// (host): vptr int32[n + 1], and vrow / vpos int32[bi[n]]: the visits of
// row k in order, each a row i and the position of (i, k) in ua; lrow
// int32[n] and lptr int64[nlev + 1]: rows level by level, a row one level
// past the deepest row that updates it.  Returns nlev, or -2 when the layout
// is not one the device pass reproduces: every row must hold its strict
// columns strictly ascending (level-of-fill can place some below the
// diagonal; the replay follows the chains through them as the reference
// does) and a diagonal slot; A's entries on or above the diagonal of the
// last rows must not run past the factor's end; and where the diagonal slot holds
// another column than k (the unpermuted index, under a reordering) A's row
// must hold the diagonal, which is then the only write to column k's cell
// before it is read.
int64_t mpk_icc_visits(int64_t bi_, int64_t bj_, int64_t irow_, int64_t icol_, int64_t rip_, int64_t riip_,
                       int64_t n, int64_t vptr_, int64_t vrow_, int64_t vpos_, int64_t lrow_, int64_t lptr_) {
  const I *bi = (const I *)bi_, *bj = (const I *)bj_, *rp = (const I *)irow_, *ci = (const I *)icol_,
          *rip = (const I *)rip_, *riip = (const I *)riip_;
  int32_t *vptr = (int32_t *)vptr_, *vrow = (int32_t *)vrow_, *vpos = (int32_t *)vpos_, *lrow = (int32_t *)lrow_;
  I *lptr = (I *)lptr_;
  if (n <= 0) {
    vptr[0] = 0;
    lptr[0] = 0;
    return 0;
  }
  if (bi[n] >= ((I)1 << 31)) return -2;
  for (I k = 0; k < n; k++) {
    const I s0 = bi[k], dg = bi[k + 1] - 1;
    if (dg < s0 || bj[dg] < 0 || bj[dg] >= n) return -2;
    for (I j = s0; j < dg; j++)
      if (bj[j] < 0 || bj[j] >= n || (j > s0 && bj[j] <= bj[j - 1])) return -2;
    I up = 0, has_diag = 0;
    for (I j = rp[rip[k]]; j < rp[rip[k] + 1]; j++) {
      up += riip[ci[j]] >= k;
      has_diag |= riip[ci[j]] == k;
    }
    // the reference zeroes one slot per such entry from the row's start; past
    // the row's end those are later rows' slots, which they rewrite
    if (s0 + up > bi[n]) return -2;
    if (bj[dg] != k && !has_diag) {
      bool slot = false;  // column k may still have a strict slot
      for (I j = s0; j < dg; j++) slot |= bj[j] == k;
      if (!slot) return -2;
    }
  }
  std::vector<I> chain(n, n), first(n, 0);
  std::vector<int32_t> lev(n, 0);
  I nv = 0;
  int32_t top = 0;
  for (I k = 0; k < n; k++) {
    vptr[k] = (int32_t)nv;
    int32_t l = 0;
    for (I i = chain[k]; i < k;) {
      const I inext = chain[i];
      const I f = first[i];
      if (nv >= bi[n]) return -2;  // more visits than entries: a chain loops
      vrow[nv] = (int32_t)i;
      vpos[nv] = (int32_t)f;
      nv++;
      l = std::max(l, lev[i] + 1);
      const I jmin = f + 1, jmax = bi[i + 1];
      if (jmin < jmax) {
        first[i] = jmin;
        const I cj = bj[jmin];
        chain[i] = chain[cj];
        chain[cj] = i;
      }
      i = inext;
    }
    lev[k] = l;
    top = std::max(top, l);
    const I jmin = bi[k], jmax = bi[k + 1] - 1;
    if (jmin < jmax) {
      first[k] = jmin;
      const I cj = bj[jmin];
      chain[k] = chain[cj];
      chain[cj] = k;
    }
  }
  vptr[n] = (int32_t)nv;
  const I nlev = (I)top + 1;
  std::vector<I> cnt(nlev + 1, 0);
  for (I k = 0; k < n; k++) cnt[lev[k] + 1]++;
  for (I l = 0; l < nlev; l++) cnt[l + 1] += cnt[l];
  for (I l = 0; l <= nlev; l++) lptr[l] = cnt[l];
  for (I k = 0; k < n; k++) lrow[cnt[lev[k]]++] = (int32_t)k;
  return nlev;
}

// Block plan of a triangular sweep (host part; the device part is
// k_sweep_blk in mpk_sweep.h, which documents the layout and encoding).  off [H]
// int64[n+1], col [H] int32 and coef [H] float64 hold every entry of every
// row in the row's own order (the order its serial sum takes).  Forward
// sweeps visit rows 0..n-1, backward sweeps n-1..0; an entry is a dependency
// when its row comes earlier in the sweep.  Blocks are bmax consecutive
// positions when n > bmax, sized so that there are about 256 of them (but not
// under 1024 rows); a smaller sweep is one block.
static int64_t plan_rows(int64_t n, int64_t bmax) {
  // tuning override: MPK_SWEEP_BLOCK_ROWS=k fixes the block size (capped by bmax)
  if (const char *e = std::getenv("MPK_SWEEP_BLOCK_ROWS")) {
    const int64_t k = std::atoll(e);
    if (k > 0) return std::min(k, std::max<int64_t>(n, 1) < bmax ? std::max<int64_t>(n, 1) : bmax);
  }
  if (n <= bmax) return n > 0 ? n : 1;
  int64_t b = (n + 255) / 256;
  b = b < 1024 ? 1024 : b;
  return b > bmax ? bmax : b;
}
// Entries sit in w fixed slots per position: w the power of two from 4 to 32
// that holds the longest row, else the multiple of 8 up to 64 that holds it
// when n w stays within 4 times the entry count; other sweeps keep CSR form
// (w = 0).
static int64_t plan_w(int64_t widest, int64_t n, int64_t m) {
  int64_t w = 4;
  while (w < widest && w < 32) w *= 2;
  if (w >= widest) return w;
  w = (widest + 7) / 8 * 8;
  return w <= 64 && n * w <= 4 * m ? w : 0;
}
// (a multiple of 16, so that plans placed one after another keep the wave
// records 16-byte aligned)
static int64_t plan_bytes(int64_t n, int64_t nblk, int64_t nlvl, int64_t slots) {
  const int64_t ints = 2 * (nblk + 1) + nlvl + 3 * n + 1;
  return (64 + ((4 * ints + 7) & ~(int64_t)7) + ((12 * slots + 7) & ~(int64_t)7) + 15) & ~(int64_t)15;
}

// Upper bound of the plan size in bytes for n rows and m entries (with the
// records of a wave plan).
int64_t mpk_sweep_plan_bound(int64_t n, int64_t m, int64_t bmax) {
  const int64_t B = plan_rows(n, bmax), nblk = n > 0 ? (n + B - 1) / B : 0;
  return plan_bytes(n, nblk, n + nblk, std::max(4 * m, 32 * n)) + n * wave_rec_bytes(8);
}

// Ring plans.  A sweep too long for one block whose DAG is deep and narrow
// (a few hundred rows per level) spends its time waiting on other blocks'
// rows; such a sweep runs instead as one block over global levels, its LDS
// holding the values of the last W positions (slot = position mod W, W the
// largest power of two <= bmax).  Rows sit at their latest possible level
// (the number of levels is the DAG's depth), which keeps each value close to
// its first reader.  A value less than W positions behind the end of the
// reading row's level is read from its slot; an older one (its slot may
// have been rewritten) from the published values, like a row of an earlier
// block.  Chosen when the average level holds at most kRingWidth rows and
// at most one entry in 64 reads an old value; MPK_SWEEP_PLAN=blocks or ring
// forces the choice (a ring only when n > bmax).
static const int64_t kRingWidth = 1024;
static int plan_mode() {
  const char *e = std::getenv("MPK_SWEEP_PLAN");
  if (!e) return 0;
  if (!std::strcmp(e, "blocks")) return 1;
  if (!std::strcmp(e, "ring")) return 2;
  return 0;
}

// Pencil blocks.  A sweep whose every dependency lies at one of three
// offsets 1 < a < b (a | b, b | n) reads as the natural ordering of an
// a x (b / a) x (n / b) grid with x fastest (7-point and lighter stencils).
// Blocks of consecutive positions are then slabs of the (x, y) planes: each
// block waits on the one below for every row, and the sweep's end moves one
// publish-to-poll hop per slab.  Blocks of t x t columns of the (x, y) plane
// spanning every z, taken in row-major tile order (reversed for a backward
// sweep), put about (a + b / a) / t hops on that path instead; every row's
// z chain stays inside its block.  t is the largest with t^2 (n / b) <= bmax.
// The order is kept only when it is topological for the sweep (every
// dependency at an earlier position), checked here; MPK_SWEEP_PENCILS=0
// turns it off.  ord[s] is the row at sweep position s, bstart the blocks'
// first positions (nblk + 1 of them).
static bool pencil_order(int64_t n, bool forward, const int64_t *off, const int *col, int64_t bmax,
                         std::vector<int> &ord, std::vector<int64_t> &bstart) {
  if (env_off("MPK_SWEEP_PENCILS") || n <= bmax) return false;
  int64_t d[3] = {0, 0, 0};
  int nd = 0;
  for (int64_t i = 0; i < n; i++)
    for (int64_t e = off[i]; e < off[i + 1]; e++) {
      const int64_t k = col[e], dist = forward ? i - k : k - i;
      if (dist <= 0) continue;
      if (dist != d[0] && dist != d[1] && dist != d[2]) {
        if (nd == 3) return false;
        d[nd++] = dist;
      }
    }
  if (nd != 3) return false;
  std::sort(d, d + 3);
  if (d[0] != 1 || d[2] % d[1] || n % d[2]) return false;
  const int64_t nx = d[1], ny = d[2] / d[1], nz = n / d[2];
  int64_t t = 1;
  while ((t + 1) * (t + 1) * nz <= bmax) t++;
  const int64_t tx = std::min(t, nx), ty = std::min(t, ny);
  const int64_t mx = (nx + tx - 1) / tx, my = (ny + ty - 1) / ty;
  if (tx * ty < 4 || mx * my < 2) return false;
  ord.clear();
  ord.reserve((size_t)n);
  bstart.assign(1, 0);
  for (int64_t by = 0; by < my; by++)
    for (int64_t bx = 0; bx < mx; bx++) {
      for (int64_t z = 0; z < nz; z++)
        for (int64_t y = by * ty; y < std::min(ny, (by + 1) * ty); y++)
          for (int64_t x = bx * tx; x < std::min(nx, (bx + 1) * tx); x++) ord.push_back((int)(z * d[2] + y * nx + x));
      bstart.push_back((int64_t)ord.size());
    }
  if (!forward) {
    std::reverse(ord.begin(), ord.end());
    std::vector<int64_t> rb(bstart.size());
    for (size_t j = 0; j < bstart.size(); j++) rb[j] = n - bstart[bstart.size() - 1 - j];
    bstart.swap(rb);
  }
  std::vector<int> rank((size_t)n);
  for (int64_t s = 0; s < n; s++) rank[(size_t)ord[(size_t)s]] = (int)s;
  for (int64_t i = 0; i < n; i++)
    for (int64_t e = off[i]; e < off[i + 1]; e++) {
      const int64_t k = col[e];
      if ((forward ? k < i : k > i) && rank[(size_t)k] >= rank[(size_t)i]) return false;
    }
  return true;
}

// Writes the plan into out [H] (cap bytes).  Returns its size in bytes, or
// minus the size needed when cap is too small.
int64_t mpk_sweep_plan(int64_t n, int64_t forward, int64_t off_, int64_t col_, int64_t coef_, int64_t bmax,
                       int64_t out_, int64_t cap) {
  const int64_t *off = (const int64_t *)off_;
  const int *col = (const int *)col_;
  const double *coef = (const double *)coef_;
  char *out = (char *)out_;
  const int64_t m = n > 0 ? off[n] : 0;
  auto spos = [&](int64_t i) { return forward ? i : n - 1 - i; };  // sweep position of row i
  auto earlier = [&](int64_t k, int64_t i) { return forward ? k < i : k > i; };
  int64_t widest = 0;
  for (int64_t i = 0; i < n; i++) widest = std::max(widest, off[i + 1] - off[i]);
  const int64_t w = plan_w(widest, n, m), slots = w ? w * n : m;
  const int mode = plan_mode();
  if (n > bmax && mode != 1) {
    int64_t W = 1;
    while (2 * W <= bmax) W *= 2;
    // earliest levels (their count is the depth), then latest levels
    std::vector<int> lev((size_t)n, 0);
    int64_t nl = 0;
    for (int64_t s = 0; s < n; s++) {
      const int64_t i = forward ? s : n - 1 - s;
      int l = 0;
      for (int64_t e = off[i]; e < off[i + 1]; e++)
        if (earlier(col[e], i)) l = std::max(l, lev[col[e]] + 1);
      lev[i] = l;
      nl = std::max<int64_t>(nl, l + 1);
    }
    // a ring needs the average level to hold at most kRingWidth rows: decided
    // by the depth alone, before the latest levels are worked out
    if (mode != 2 && n > kRingWidth * nl) goto blocks;
    std::fill(lev.begin(), lev.end(), (int)(nl - 1));
    for (int64_t s = n - 1; s >= 0; s--) {
      const int64_t i = forward ? s : n - 1 - s;
      for (int64_t e = off[i]; e < off[i + 1]; e++)
        if (earlier(col[e], i)) lev[col[e]] = std::min(lev[col[e]], lev[i] - 1);
    }
    std::vector<int64_t> start((size_t)nl + 1, 0);
    for (int64_t i = 0; i < n; i++) start[(size_t)lev[i] + 1]++;
    for (int64_t l = 0; l < nl; l++) start[(size_t)l + 1] += start[(size_t)l];
    std::vector<int> pos((size_t)n), prow_h((size_t)n);
    {
      std::vector<int64_t> next(start.begin(), start.end() - 1);
      for (int64_t s = 0; s < n; s++) {
        const int64_t i = forward ? s : n - 1 - s;
        const int64_t p = next[(size_t)lev[i]]++;
        pos[(size_t)i] = (int)p;
        prow_h[(size_t)p] = (int)i;
      }
    }
    auto in_window = [&](int64_t i, int64_t k) { return start[(size_t)lev[i] + 1] - pos[(size_t)k] <= W; };
    int64_t old = 0;
    for (int64_t i = 0; i < n; i++)
      for (int64_t e = off[i]; e < off[i + 1]; e++)
        if (earlier(col[e], i) && !in_window(i, col[e])) old++;
    if (mode == 2 || (n <= kRingWidth * nl && 64 * old <= m)) {
      const int64_t nlvl = nl + 1, bytes = plan_bytes(n, 1, nlvl, slots);
      if (bytes > cap) return -bytes;
      std::memset(out, 0, (size_t)bytes);
      int64_t *hdr = (int64_t *)out;
      hdr[0] = 1;
      hdr[1] = m;
      hdr[2] = nlvl;
      hdr[3] = nl;
      hdr[4] = bytes;
      hdr[5] = w;
      hdr[6] = W;
      int *blk_pos = (int *)(out + 64), *blk_lev = blk_pos + 2, *lvl = blk_lev + 2;
      int *prow = lvl + nlvl, *pcnt = prow + n, *pent = pcnt + n;
      const int64_t ints = 2 * 2 + nlvl + 3 * n + 1;
      double *ecoef = (double *)(out + 64 + ((4 * ints + 7) & ~(int64_t)7));
      int *edep = (int *)(ecoef + slots);
      blk_pos[0] = 0;
      blk_pos[1] = (int)n;
      blk_lev[0] = 0;
      blk_lev[1] = (int)nlvl;
      for (int64_t l = 0; l <= nl; l++) lvl[l] = (int)start[(size_t)l];
      int64_t e_out = 0;
      for (int64_t p = 0; p < n; p++) {
        const int64_t i = prow_h[(size_t)p];
        prow[p] = (int)i;
        pent[p] = w ? 0 : (int)e_out;
        pcnt[p] = (int)(off[i + 1] - off[i]);
        for (int64_t e = off[i]; e < off[i + 1]; e++, e_out++) {
          const int64_t k = col[e], q = w ? (e - off[i]) * n + p : e_out;  // slot-major with fixed slots
          ecoef[q] = coef[e];
          if (!earlier(k, i))
            edep[q] = (int)(-k - n - 1);
          else if (in_window(i, k))
            edep[q] = (int)(pos[(size_t)k] & (W - 1));
          else
            edep[q] = (int)(-k - 1);
        }
      }
      pent[n] = w ? 0 : (int)e_out;
      return mark_plan(out, n, bytes, cap);
    }
  }
blocks:
  // the blocks: pencils, or runs of B consecutive sweep positions
  std::vector<int> ord;
  std::vector<int64_t> bstart;
  if (mode != 2 && pencil_order(n, forward, off, col, bmax, ord, bstart)) {
  } else {
    const int64_t B = plan_rows(n, bmax), nb = n > 0 ? (n + B - 1) / B : 0;
    ord.resize((size_t)n);
    for (int64_t s = 0; s < n; s++) ord[(size_t)s] = (int)spos(s);
    bstart.resize((size_t)nb + 1);
    for (int64_t b = 0; b <= nb; b++) bstart[(size_t)b] = std::min(n, b * B);
  }
  const int64_t nblk = (int64_t)bstart.size() - 1;
  std::vector<int> rblk(n > 0 ? n : 1);  // each row's block
  for (int64_t b = 0; b < nblk; b++)
    for (int64_t s = bstart[(size_t)b]; s < bstart[(size_t)b + 1]; s++) rblk[(size_t)ord[(size_t)s]] = (int)b;
  // local levels
  std::vector<int> lev(n > 0 ? n : 1, 0);
  std::vector<int> nlev(nblk > 0 ? nblk : 1, 0);
  for (int64_t s = 0; s < n; s++) {
    const int64_t i = ord[(size_t)s], b = rblk[(size_t)i];
    int l = 0;
    for (int64_t e = off[i]; e < off[i + 1]; e++) {
      const int64_t k = col[e];
      if (earlier(k, i) && rblk[(size_t)k] == b) l = std::max(l, lev[k] + 1);
    }
    lev[i] = l;
    nlev[b] = std::max(nlev[b], l + 1);
  }
  int64_t nlvl = 0, depth = 0;
  for (int64_t b = 0; b < nblk; b++) {
    nlvl += nlev[b] + 1;
    depth = std::max<int64_t>(depth, nlev[b]);
  }
  const int64_t bytes = plan_bytes(n, nblk, nlvl, slots);
  if (bytes > cap) return -bytes;
  std::memset(out, 0, (size_t)bytes);
  int64_t *hdr = (int64_t *)out;
  hdr[0] = nblk;
  hdr[1] = m;
  hdr[2] = nlvl;
  hdr[3] = depth;
  hdr[4] = bytes;
  hdr[5] = w;
  int *blk_pos = (int *)(out + 64), *blk_lev = blk_pos + nblk + 1, *lvl = blk_lev + nblk + 1;
  int *prow = lvl + nlvl, *pcnt = prow + n, *pent = pcnt + n;
  const int64_t ints = 2 * (nblk + 1) + nlvl + 3 * n + 1;
  double *ecoef = (double *)(out + 64 + ((4 * ints + 7) & ~(int64_t)7));
  int *edep = (int *)(ecoef + slots);
  // positions: blocks in sweep order, rows by level inside a block (sweep
  // order within a level)
  std::vector<int> pos(n > 0 ? n : 1);
  int64_t li = 0;
  for (int64_t b = 0; b < nblk; b++) {
    const int64_t s0 = bstart[(size_t)b], s1 = bstart[(size_t)b + 1];
    std::vector<int64_t> start((size_t)nlev[b] + 1, 0);
    for (int64_t s = s0; s < s1; s++) start[(size_t)lev[ord[(size_t)s]] + 1]++;
    for (size_t l = 0; l + 1 < start.size(); l++) start[l + 1] += start[l];
    blk_pos[b] = (int)s0;
    blk_lev[b] = (int)li;
    for (int64_t l = 0; l < nlev[b]; l++) lvl[li + l] = (int)(s0 + start[(size_t)l]);
    lvl[li + nlev[b]] = (int)s1;
    li += nlev[b] + 1;
    for (int64_t s = s0; s < s1; s++) {
      const int64_t i = ord[(size_t)s];
      const int64_t p = s0 + start[(size_t)lev[i]]++;
      prow[p] = (int)i;
      pos[i] = (int)p;
    }
  }
  blk_pos[nblk] = (int)n;
  blk_lev[nblk] = (int)li;
  // entries in position order (fixed slots: entry e of position p at e n + p,
  // so that the rows of a level read each slot from consecutive addresses)
  int64_t e_out = 0;
  for (int64_t p = 0; p < n; p++) {
    const int64_t i = prow[p], b = rblk[(size_t)i];
    pent[p] = (int)e_out;
    pcnt[p] = (int)(off[i + 1] - off[i]);
    for (int64_t e = off[i]; e < off[i + 1]; e++, e_out++) {
      const int64_t k = col[e], q = w ? (e - off[i]) * n + p : e_out;
      ecoef[q] = coef[e];
      if (!earlier(k, i))
        edep[q] = (int)(-k - n - 1);
      else if (rblk[(size_t)k] == b)
        edep[q] = pos[k] - blk_pos[b];
      else
        edep[q] = (int)(-k - 1);
    }
  }
  pent[n] = (int)(w ? 0 : e_out);
  if (w)
    for (int64_t p = 0; p <= n; p++) pent[p] = 0;  // unused with fixed slots
  return mark_plan(out, n, bytes, cap);
}

}  // extern "C"

namespace mpk {

// The planner's shape rules, shared with the device planner (mpk_plan.hip),
// which must lay plans out exactly as mpk_sweep_plan does.
int64_t plan_rows_rule(int64_t n, int64_t bmax) { return plan_rows(n, bmax); }
int64_t plan_w_rule(int64_t widest, int64_t n, int64_t m) { return plan_w(widest, n, m); }
int64_t plan_bytes_rule(int64_t n, int64_t nblk, int64_t nlvl, int64_t slots) {
  return plan_bytes(n, nblk, nlvl, slots);
}
int plan_mode_rule() { return plan_mode(); }
int64_t plan_ring_width_rule() { return kRingWidth; }
bool plan_env_off(const char *name) { return env_off(name); }
int64_t wave_rec_bytes_rule(int64_t w) { return wave_rec_bytes(w); }

}  // namespace mpk
