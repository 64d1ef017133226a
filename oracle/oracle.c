/*
 * oracle.c - CPU restatement of the reference's SpMV arithmetic and CPU
 * format builders. TEST INFRASTRUCTURE ONLY: the library never links, loads
 * or calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg use it, as the checker.
 *
 * Each function names the reference file:line it restates (paths relative
 * to the reference repository root). Written in plain C99 (gcc), no
 * floating-point contraction (-ffp-contract=off), so every product is
 * rounded before it is added, as in the reference's x86 build.
 *
 * Parity pins (tests/test_oracle.py): golden fixtures from the compiled
 * reference dbsr.h (oracle/_ref), closed forms for stencils with x == 1,
 * dense reconstruction of every blocked format, brute-force dense SpMV on
 * small inputs, invariants of the DASP layout.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* SpMV: src/ref.cpp:26-37 (FP64), :14-24 (FP32), :3-13 (FP16).            */

/* y_i = sum_j val[j]*x[col[j]], row by row, left to right (ref.cpp:29-36). */
void orc_spmv_f64(int64_t nrows, const int64_t* rowptr, const int32_t* col, const double* val,
                  const double* x, double* y) {
  for (int64_t i = 0; i < nrows; i++) {
    double sum = 0;
    for (int64_t j = rowptr[i]; j < rowptr[i + 1]; j++) sum += val[j] * x[col[j]];
    y[i] = sum;
  }
}

typedef struct {
  int64_t r0, r1;
  const int64_t* rowptr;
  const int32_t* col;
  const double* val;
  const double* x;
  double* y;
} spmv_job;

static void* spmv_worker(void* arg) {
  spmv_job* j = (spmv_job*)arg;
  for (int64_t i = j->r0; i < j->r1; i++) {
    double sum = 0;
    for (int64_t k = j->rowptr[i]; k < j->rowptr[i + 1]; k++) sum += j->val[k] * j->x[j->col[k]];
    j->y[i] = sum;
  }
  return NULL;
}

/* The same loop with rows split statically over `threads` threads. Each row
 * keeps its summation order, so the result is bit-identical to
 * orc_spmv_f64. */
void orc_spmv_f64_par(int64_t nrows, const int64_t* rowptr, const int32_t* col, const double* val,
                      const double* x, double* y, int threads) {
  if (threads <= 1 || nrows < 1024) {
    orc_spmv_f64(nrows, rowptr, col, val, x, y);
    return;
  }
  if (threads > 256) threads = 256;
  pthread_t th[256];
  spmv_job jobs[256];
  int64_t chunk = (nrows + threads - 1) / threads;
  int started = 0;
  for (int t = 0; t < threads; t++) {
    int64_t a = t * chunk, b = a + chunk < nrows ? a + chunk : nrows;
    if (a >= b) break;
    jobs[t] = (spmv_job){a, b, rowptr, col, val, x, y};
    if (pthread_create(&th[t], NULL, spmv_worker, &jobs[t]) != 0) {
      spmv_worker(&jobs[t]);
      continue;
    }
    started = t + 1;
  }
  for (int t = 0; t < started; t++) pthread_join(th[t], NULL);
}

/* ref.cpp:14-24: all-FP32 loop. */
void orc_spmv_f32(int64_t nrows, const int64_t* rowptr, const int32_t* col, const float* val,
                  const float* x, float* y) {
  for (int64_t i = 0; i < nrows; i++) {
    float sum = 0;
    for (int64_t j = rowptr[i]; j < rowptr[i + 1]; j++) sum += val[j] * x[col[j]];
    y[i] = sum;
  }
}

/* IEEE binary16 <-> binary32, round to nearest even (textbook bit
 * manipulation; pinned against numpy.float16 in the tests). */
uint16_t orc_f32_to_f16(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t absx = x & 0x7fffffffu;
  if (absx >= 0x7f800000u) /* inf or nan */
    return (uint16_t)(sign | 0x7c00u | (absx > 0x7f800000u ? 0x200u | ((absx >> 13) & 0x3ffu) : 0));
  if (absx >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* rounds to >= 65520 -> inf */
  if (absx < 0x38800000u) {                                    /* result subnormal or zero */
    if (absx < 0x33000000u) return (uint16_t)sign;             /* < 2^-25: rounds to 0 */
    uint32_t e = absx >> 23;
    uint32_t m = (absx & 0x7fffffu) | 0x800000u;
    uint32_t shift = 126 - e; /* 14 + (112 - e) ... value = m * 2^(e-150); unit 2^-24 */
    uint32_t q = m >> shift, rem = m & ((1u << shift) - 1), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (q & 1))) q++;
    return (uint16_t)(sign | q);
  }
  uint32_t h = ((absx - 0x38000000u) >> 13);
  uint32_t rem = absx & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1))) h++;
  return (uint16_t)(sign | h);
}

float orc_f16_to_f32(uint16_t h) {
  uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu, x;
  if (e == 0) {
    if (m == 0) {
      x = sign;
    } else { /* subnormal: m * 2^-24 */
      float f = (float)m * 5.9604644775390625e-08f;
      memcpy(&x, &f, 4);
      x |= sign;
    }
  } else if (e == 31) {
    x = sign | 0x7f800000u | (m << 13);
  } else {
    x = sign | ((e + 112) << 23) | (m << 13);
  }
  float f;
  memcpy(&f, &x, 4);
  return f;
}

/* ref.cpp:3-13: values and x rounded to half, products and sum in FP32,
 * y rounded to half (returned as float). */
void orc_spmv_f16(int64_t nrows, const int64_t* rowptr, const int32_t* col, const float* val,
                  const float* x, float* y) {
  for (int64_t i = 0; i < nrows; i++) {
    float sum = 0;
    for (int64_t j = rowptr[i]; j < rowptr[i + 1]; j++)
      sum += orc_f16_to_f32(orc_f32_to_f16(val[j])) * orc_f16_to_f32(orc_f32_to_f16(x[col[j]]));
    y[i] = orc_f16_to_f32(orc_f32_to_f16(sum));
  }
}

/* ------------------------------------------------------------------------ */
/* Short/long split: src/spmv.cu:131-196 (`preprocess`). Rows with more than
 * `threshold` nonzeros go to the long list, the rest to the short list, both
 * in ascending row order (the index maps of the two sub-CSRs). */
void orc_split(int64_t nrows, const int64_t* rowptr, int64_t threshold, int32_t* short_map,
               int64_t* nshort, int32_t* long_map, int64_t* nlong) {
  int64_t s = 0, l = 0;
  for (int64_t i = 0; i < nrows; i++) {
    if (rowptr[i + 1] - rowptr[i] > threshold)
      long_map[l++] = (int32_t)i;
    else
      short_map[s++] = (int32_t)i;
  }
  *nshort = s;
  *nlong = l;
}

/* ------------------------------------------------------------------------ */
/* BCSR: utils/utils.hpp:9-124 (csrToBcsr<T,Y>) for any TILE_M x TILE_K, plus
 * the AmgT 4x4 extras of src/amgT.cu:585-1200 (blcMap, balance arrays,
 * avg_nnz, stand).
 *
 * The reference finds occupied blocks with a dense rowRegions x colRegions
 * flag array; this restatement walks one block row at a time with a
 * colRegions-sized marker, which yields the same blocks in the same
 * (block row, ascending block column) order. Values are placed row-major
 * at (row%TM)*TK + col%TK (utils.hpp:356); duplicates: last write wins. */
typedef struct {
  int64_t brows, bcols, nblocks;
  int32_t tm, tk;
  int32_t* rowptr; /* brows+1 */
  int32_t* colidx; /* nblocks */
  double* val;     /* nblocks*tm*tk */
  uint16_t* map;   /* nblocks+1 (4x4 only): bit r*4+c per stored entry, amgT.cu:781-834 */
  int32_t warpnum;       /* sum ceil(blocks_in_row/64), amgT.cu:523-545,1174-1200 */
  int32_t* rowptr_by_warp; /* brows+1 */
  int32_t* rowidx_by_warp; /* warpnum */
  double avg_nnz, stand;   /* amgT.cu:1144-1157 */
} orc_bcsr;

orc_bcsr* orc_bcsr_build(int64_t nrows, int64_t ncols, const int64_t* rowptr, const int32_t* col,
                         const double* val, int tm, int tk) {
  orc_bcsr* b = (orc_bcsr*)calloc(1, sizeof(orc_bcsr));
  b->tm = tm;
  b->tk = tk;
  b->brows = (nrows + tm - 1) / tm;
  b->bcols = (ncols + tk - 1) / tk;
  b->rowptr = (int32_t*)calloc(b->brows + 1, sizeof(int32_t));
  int32_t* mark = (int32_t*)malloc(sizeof(int32_t) * (b->bcols ? b->bcols : 1));
  for (int64_t i = 0; i < b->bcols; i++) mark[i] = -1;
  /* pass 1: count distinct block columns per block row */
  for (int64_t br = 0; br < b->brows; br++) {
    int32_t cnt = 0;
    for (int64_t r = br * tm; r < (br + 1) * tm && r < nrows; r++)
      for (int64_t j = rowptr[r]; j < rowptr[r + 1]; j++) {
        int64_t bc = col[j] / tk;
        if (mark[bc] != (int32_t)br) {
          mark[bc] = (int32_t)br;
          cnt++;
        }
      }
    b->rowptr[br + 1] = b->rowptr[br] + cnt;
  }
  b->nblocks = b->rowptr[b->brows];
  b->colidx = (int32_t*)malloc(sizeof(int32_t) * (b->nblocks ? b->nblocks : 1));
  b->val = (double*)calloc((size_t)(b->nblocks ? b->nblocks : 1) * tm * tk, sizeof(double));
  b->map = (uint16_t*)calloc(b->nblocks + 1, sizeof(uint16_t));
  /* pass 2: ascending block columns, then values */
  int32_t* slot = (int32_t*)malloc(sizeof(int32_t) * (b->bcols ? b->bcols : 1));
  for (int64_t i = 0; i < b->bcols; i++) mark[i] = -1;
  for (int64_t br = 0; br < b->brows; br++) {
    int32_t base = b->rowptr[br], cnt = 0;
    for (int64_t r = br * tm; r < (br + 1) * tm && r < nrows; r++)
      for (int64_t j = rowptr[r]; j < rowptr[r + 1]; j++) {
        int64_t bc = col[j] / tk;
        if (mark[bc] != (int32_t)br) {
          mark[bc] = (int32_t)br;
          b->colidx[base + cnt++] = (int32_t)bc;
        }
      }
    /* insertion sort of this block row's columns (few per row) */
    for (int32_t a = 1; a < cnt; a++) {
      int32_t v = b->colidx[base + a], k = a - 1;
      while (k >= 0 && b->colidx[base + k] > v) {
        b->colidx[base + k + 1] = b->colidx[base + k];
        k--;
      }
      b->colidx[base + k + 1] = v;
    }
    for (int32_t a = 0; a < cnt; a++) slot[b->colidx[base + a]] = base + a;
    for (int64_t r = br * tm; r < (br + 1) * tm && r < nrows; r++)
      for (int64_t j = rowptr[r]; j < rowptr[r + 1]; j++) {
        int64_t c = col[j];
        int64_t blk = slot[c / tk];
        int64_t off = (r % tm) * tk + c % tk;
        b->val[blk * tm * tk + off] = val[j];
        if (tm == 4 && tk == 4) b->map[blk] |= (uint16_t)(1u << off);
      }
  }
  free(slot);
  free(mark);
  /* AmgT balance arrays with 64 blocks per warp (amgT.cu:523-545 with
   * WARP_CAPACITY 64 from include/common.h:7). */
  b->rowptr_by_warp = (int32_t*)calloc(b->brows + 1, sizeof(int32_t));
  for (int64_t br = 0; br < b->brows; br++)
    b->rowptr_by_warp[br + 1] =
        b->rowptr_by_warp[br] + (b->rowptr[br + 1] - b->rowptr[br] + 63) / 64;
  b->warpnum = b->rowptr_by_warp[b->brows];
  b->rowidx_by_warp = (int32_t*)malloc(sizeof(int32_t) * (b->warpnum ? b->warpnum : 1));
  for (int64_t br = 0; br < b->brows; br++)
    for (int32_t w = b->rowptr_by_warp[br]; w < b->rowptr_by_warp[br + 1]; w++)
      b->rowidx_by_warp[w] = (int32_t)br;
  int64_t nnz = nrows ? rowptr[nrows] : 0;
  b->avg_nnz = b->nblocks ? (double)nnz / (double)b->nblocks : 0.0;
  double avg_len = b->brows ? (double)b->nblocks / (double)b->brows : 0.0, acc = 0.0;
  for (int64_t br = 0; br < b->brows; br++) {
    double d = (double)(b->rowptr[br + 1] - b->rowptr[br]) - avg_len;
    acc += d * d;
  }
  b->stand = b->brows ? sqrt(acc / (double)b->brows) : 0.0;
  return b;
}

void orc_bcsr_free(orc_bcsr* b) {
  if (!b) return;
  free(b->rowptr);
  free(b->colidx);
  free(b->val);
  free(b->map);
  free(b->rowptr_by_warp);
  free(b->rowidx_by_warp);
  free(b);
}

/* y = A*x from a BCSR (block by block, row-major within the block). */
void orc_bcsr_spmv(const orc_bcsr* b, int64_t nrows, const double* x, int64_t ncols, double* y) {
  for (int64_t r = 0; r < nrows; r++) y[r] = 0.0;
  for (int64_t br = 0; br < b->brows; br++)
    for (int32_t k = b->rowptr[br]; k < b->rowptr[br + 1]; k++)
      for (int i = 0; i < b->tm; i++) {
        int64_t r = br * b->tm + i;
        if (r >= nrows) break;
        for (int j = 0; j < b->tk; j++) {
          int64_t c = (int64_t)b->colidx[k] * b->tk + j;
          if (c < ncols) y[r] += b->val[(int64_t)k * b->tm * b->tk + i * b->tk + j] * x[c];
        }
      }
}

/* ------------------------------------------------------------------------ */
/* DBSR: include/dbsr.h:208-253 (csrTodbsrWithRecorder), :10-93 (csrTodbsr),
 * :94-169 (balanceDbsr with WARP_CAPACITY 64, include/common.h:7). */
typedef struct {
  int64_t nrows, ncols, nnz, brows, nblocks;
  int32_t* recorder; /* nrows: recorderMap[new] = old */
  int32_t* blcptr;   /* brows+1 */
  int32_t* colidx;   /* nblocks*16 */
  double* val;       /* nblocks*16 */
  uint16_t* map;     /* nblocks */
  int32_t warpnum;
  int32_t* prefix; /* warpnum+1 */
  /* 1 when the reference's second pass would index past the warpcount
   * entries its first pass allocated (dbsr.h:135-155, undefined behaviour
   * there); the oracle then extends the warp list and warpnum counts it. */
  int32_t overflow;
} orc_dbsr;

static int cmp_len_idx(const void* a, const void* b) {
  const int64_t* x = (const int64_t*)a;
  const int64_t* y = (const int64_t*)b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  return x[1] < y[1] ? -1 : (x[1] > y[1]);
}

/* sort_rows = 1: csrTodbsrWithRecorder (dbsr.h:208-253), rows sorted first.
 * sort_rows = 0: csrTodbsr (dbsr.h:10-93) on the CSR as given, then
 * balanceDbsr (:94-169); recorder is then the identity. */
orc_dbsr* orc_dbsr_build_ex(int64_t nrows, int64_t ncols, const int64_t* rowptr,
                            const int32_t* col, const double* val, int sort_rows) {
  orc_dbsr* d = (orc_dbsr*)calloc(1, sizeof(orc_dbsr));
  d->nrows = nrows;
  d->ncols = ncols;
  d->nnz = nrows ? rowptr[nrows] : 0;
  /* dbsr.h:213-226: rows sorted by (len, index) ascending; recorder[new] = old */
  int64_t* key = (int64_t*)malloc(sizeof(int64_t) * 2 * (nrows ? nrows : 1));
  for (int64_t i = 0; i < nrows; i++) {
    key[2 * i] = sort_rows ? rowptr[i + 1] - rowptr[i] : 0;
    key[2 * i + 1] = i;
  }
  if (sort_rows) qsort(key, (size_t)nrows, 2 * sizeof(int64_t), cmp_len_idx);
  d->recorder = (int32_t*)malloc(sizeof(int32_t) * (nrows ? nrows : 1));
  for (int64_t i = 0; i < nrows; i++) d->recorder[i] = (int32_t)key[2 * i + 1];
  free(key);
  /* dbsr.h:14-28: brow = ceil(row/4); width of a 4-row group = max ceil(len/4) */
  d->brows = (nrows + 3) / 4;
  d->blcptr = (int32_t*)calloc(d->brows + 1, sizeof(int32_t));
  for (int64_t i = 0; i < nrows; i++) {
    int64_t o = d->recorder[i];
    int32_t w = (int32_t)((rowptr[o + 1] - rowptr[o] + 3) / 4);
    if (w > d->blcptr[i / 4 + 1]) d->blcptr[i / 4 + 1] = w;
  }
  for (int64_t i = 1; i <= d->brows; i++) d->blcptr[i] += d->blcptr[i - 1];
  d->nblocks = d->blcptr[d->brows];
  size_t nb = (size_t)(d->nblocks ? d->nblocks : 1);
  d->colidx = (int32_t*)calloc(nb * 16, sizeof(int32_t));
  d->val = (double*)calloc(nb * 16, sizeof(double));
  d->map = (uint16_t*)calloc(nb, sizeof(uint16_t));
  /* dbsr.h:65-80: element z of sorted row i -> block blcptr[i/4] + z/4,
   * slot (i%4)*4 + z%4 */
  for (int64_t i = 0; i < nrows; i++) {
    int64_t o = d->recorder[i];
    int64_t z = 0;
    for (int64_t j = rowptr[o]; j < rowptr[o + 1]; j++, z++) {
      int64_t blk = d->blcptr[i / 4] + z / 4;
      int64_t slot = (i % 4) * 4 + z % 4;
      d->val[blk * 16 + slot] = val[j];
      d->colidx[blk * 16 + slot] = col[j];
      d->map[blk] |= (uint16_t)(1u << slot);
    }
  }
  /* dbsr.h:94-169, two passes as written: the second pass starts from the
   * `now` the first pass ended with, and opening a warp for a short row does
   * not reset `now` to that row's length. */
  const int C = 64;
  int now = 0, warpcount = 0;
  for (int64_t i = 0; i < d->brows; i++) {
    int len = d->blcptr[i + 1] - d->blcptr[i];
    if (now + len <= C) {
      now += len;
    } else {
      if (now != 0) {
        warpcount++;
        now = 0;
      }
      if (len <= C)
        now = len;
      else
        warpcount += (len + C - 1) / C;
    }
  }
  if (now) warpcount++;
  int64_t cap = warpcount + 2;
  for (int64_t i = 0; i < d->brows; i++) cap += (d->blcptr[i + 1] - d->blcptr[i] + C - 1) / C + 1;
  int32_t* cnt = (int32_t*)calloc((size_t)cap, sizeof(int32_t));
  int now_warp = 0;
  for (int64_t i = 0; i < d->brows; i++) {
    int len = d->blcptr[i + 1] - d->blcptr[i];
    if (now + len <= C) {
      cnt[now_warp]++;
      now += len;
    } else {
      if (now != 0) now = 0;
      if (len <= C) {
        now_warp++;
        cnt[now_warp]++;
      } else {
        for (int j = 1; j <= (len + C - 1) / C; j++) {
          now_warp++;
          cnt[now_warp]++;
        }
      }
    }
  }
  d->overflow = now_warp >= warpcount && d->brows > 0;
  if (d->overflow) warpcount = now_warp + 1;
  d->warpnum = warpcount;
  d->prefix = (int32_t*)calloc((size_t)warpcount + 1, sizeof(int32_t));
  for (int w = 0; w < warpcount; w++) d->prefix[w + 1] = d->prefix[w] + cnt[w];
  free(cnt);
  return d;
}

orc_dbsr* orc_dbsr_build(int64_t nrows, int64_t ncols, const int64_t* rowptr, const int32_t* col,
                         const double* val) {
  return orc_dbsr_build_ex(nrows, ncols, rowptr, col, val, 1);
}

void orc_dbsr_free(orc_dbsr* d) {
  if (!d) return;
  free(d->recorder);
  free(d->blcptr);
  free(d->colidx);
  free(d->val);
  free(d->map);
  free(d->prefix);
  free(d);
}

/* ------------------------------------------------------------------------ */
/* DASP: src/dasp.cu:814-1503 (host part of spmv_all) with MMA_M 8, MMA_K 4,
 * BlockSize 8, warpNum_short 4, loopNum_short 4, warpNum_long 4,
 * loopNum_long 2, groupNum 1, rowloop 1 (dasp.cu:27-71, 867). The dead
 * superX and short3 arrays (dasp.cu:1045-1239, 1311-1327) are not built. */
typedef struct {
  int32_t row_long, row_block, row_zero;
  int32_t short_row_1, short_row_2, short_row_3, short_row_4, common_13;
  int32_t fill0_nnz_short13, fill0_nnz_short34, fill0_nnz_short22, fill0_nnz_short;
  int32_t threadblock13, threadblock34, threadblock22;
  int32_t blocknum, fill0_nnz_reg, nnz_irreg;
  int32_t warp_number, fill0_nnz_long, nnz_long;
  int32_t* order_rid;   /* nrows */
  double* short_val;    /* fill0_nnz_short */
  int32_t* short_cid;
  int32_t* long_rpt;    /* row_long+1: warp offset of each long row (long_rpt_new) */
  double* long_val;     /* fill0_nnz_long */
  int32_t* long_cid;
  int32_t* rid_by_warp; /* warp_number (padded entries 0) */
  int32_t* block_ptr;   /* blocknum+1 */
  int32_t* irreg_rpt;   /* row_block+1 */
  double* reg_val;      /* fill0_nnz_reg */
  int32_t* reg_cid;
  double* irreg_val;    /* nnz_irreg */
  int32_t* irreg_cid;
} orc_dasp;

/* dasp.cu:214-246 + 282-289: LSD base-10 radix sort whose passes are stable
 * and descending, i.e. a stable sort by length, longest first. */
static void dasp_radix_desc(int32_t* arr, int32_t* idx, int32_t len) {
  if (len <= 0) return;
  int32_t mx = arr[0];
  for (int32_t i = 1; i < len; i++)
    if (arr[i] > mx) mx = arr[i];
  int32_t* ta = (int32_t*)malloc(sizeof(int32_t) * len);
  int32_t* ti = (int32_t*)malloc(sizeof(int32_t) * len);
  for (int64_t exp = 1; mx / exp > 0; exp *= 10) {
    int32_t buckets[10] = {0};
    for (int32_t i = 0; i < len; i++) buckets[(arr[i] / exp) % 10]++;
    for (int b = 1; b < 10; b++) buckets[b] += buckets[b - 1];
    for (int32_t i = 0; i < len; i++) {
      int d = (int)((arr[i] / exp) % 10);
      int32_t off = len - (buckets[d] - 1) - 1;
      ta[off] = arr[i];
      ti[off] = idx[i];
      buckets[d]--;
    }
    memcpy(arr, ta, sizeof(int32_t) * len);
    memcpy(idx, ti, sizeof(int32_t) * len);
  }
  free(ta);
  free(ti);
}

orc_dasp* orc_dasp_build(int32_t rowA, const int32_t* rpt, const int32_t* cid, const double* val,
                         double threshold, int32_t block_longest) {
  enum { BS = 8, MM = 8, MK = 4, WNS = 4, LNS = 4, WNL = 4, LNL = 2, ROWLOOP = 1 };
  orc_dasp* D = (orc_dasp*)calloc(1, sizeof(orc_dasp));
  int32_t s1 = 0, s2 = 0, s3 = 0, s4 = 0, rl = 0, rb = 0, rz = 0;
  /* dasp.cu:827-859 classification */
  for (int32_t i = 0; i < rowA; i++) {
    int32_t L = rpt[i + 1] - rpt[i];
    if (L == 1) s1++;
    else if (L == 3) s3++;
    else if (L == 2) s2++;
    else if (L == 0) rz++;
    else if (L == 4) s4++;
    else if (L >= block_longest) rl++;
    else rb++;
  }
  int32_t* rid1 = (int32_t*)malloc(sizeof(int32_t) * (s1 + 1));
  int32_t* rid2 = (int32_t*)malloc(sizeof(int32_t) * (s2 + 1));
  int32_t* rid3 = (int32_t*)malloc(sizeof(int32_t) * (s3 + 1));
  int32_t* rid4 = (int32_t*)malloc(sizeof(int32_t) * (s4 + 1));
  int32_t* ridl = (int32_t*)malloc(sizeof(int32_t) * (rl + 1));
  int32_t* ridz = (int32_t*)malloc(sizeof(int32_t) * (rz + 1));
  int32_t* ridA = (int32_t*)malloc(sizeof(int32_t) * (rb + 1));
  int32_t* rptA = (int32_t*)calloc(rb + 1, sizeof(int32_t));
  int32_t* long_rpt = (int32_t*)calloc(rl + 1, sizeof(int32_t));
  {
    int32_t f1 = 0, f2 = 0, f3 = 0, f4 = 0, fl = 0, fz = 0, fb = 0;
    for (int32_t i = 0; i < rowA; i++) { /* dasp.cu:883-924 */
      int32_t L = rpt[i + 1] - rpt[i];
      if (L == 1) rid1[f1++] = i;
      else if (L == 3) rid3[f3++] = i;
      else if (L == 2) rid2[f2++] = i;
      else if (L == 0) ridz[fz++] = i;
      else if (L == 4) rid4[f4++] = i;
      else if (L >= block_longest) { long_rpt[fl] = L; ridl[fl++] = i; }
      else { rptA[fb] = L; ridA[fb++] = i; }
    }
  }
  int32_t nnz_short = s1 + s3 * 3 + s2 * 2 + s4 * 4;
  /* dasp.cu:927-937 */
  int32_t c13 = s1 < s3 ? s1 : s3;
  if (c13 / BS >= 16) {
    c13 = BS * (c13 / BS);
    s1 -= c13;
    s3 -= c13;
  } else {
    c13 = 0;
  }
  /* dasp.cu:939-960 */
  int32_t short_block13 = (c13 + BS - 1) / BS;
  int32_t half_short_row_2 = (s2 + 1) / 2;
  int32_t short_block22 = (half_short_row_2 + BS - 1) / BS;
  int32_t short_row_34 = s3 + s4;
  int32_t short_block34 = (short_row_34 + BS - 1) / BS;
  int32_t b13_tb = WNS * 1 * 2, b22_tb = WNS * 1 * 2, b34_tb = WNS * 1 * LNS;
  int32_t tb13 = (short_block13 + b13_tb - 1) / b13_tb;
  int32_t tb22 = (short_block22 + b22_tb - 1) / b22_tb;
  int32_t tb34 = (short_block34 + b34_tb - 1) / b34_tb;
  int32_t f13 = tb13 * b13_tb * MM * MK, f34 = tb34 * b34_tb * MM * MK, f22 = tb22 * b22_tb * MM * MK;
  int32_t fshort = s1 + f13 + f34 + f22;
  D->short_val = (double*)calloc(fshort + 1, sizeof(double));
  D->short_cid = (int32_t*)calloc(fshort + 1, sizeof(int32_t));
  /* dasp.cu:969-974 singles */
  for (int32_t i = 0; i < s1; i++) {
    D->short_val[i] = val[rpt[rid1[i]]];
    D->short_cid[i] = cid[rpt[rid1[i]]];
  }
  /* dasp.cu:976-994 1&3 tiles */
  for (int32_t i = 0; i < short_block13; i++) {
    double* cv = D->short_val + s1 + i * MM * MK;
    int32_t* cc = D->short_cid + s1 + i * MM * MK;
    for (int32_t j = 0; j < BS && i * BS + j < c13; j++) {
      int32_t r1 = rid1[s1 + i * BS + j], r3 = rid3[i * BS + j];
      cv[j * MK] = val[rpt[r1]];
      cc[j * MK] = cid[rpt[r1]];
      for (int t = 0; t < 3; t++) {
        cv[j * MK + 1 + t] = val[rpt[r3] + t];
        cc[j * MK + 1 + t] = cid[rpt[r3] + t];
      }
    }
  }
  /* dasp.cu:996-1028 len-3 then len-4 rows, 4 slots each */
  for (int32_t i = 0; i < s3; i++) {
    double* cv = D->short_val + s1 + f13 + i * MK;
    int32_t* cc = D->short_cid + s1 + f13 + i * MK;
    int32_t r = rid3[c13 + i];
    for (int t = 0; t < 3; t++) { cv[t] = val[rpt[r] + t]; cc[t] = cid[rpt[r] + t]; }
  }
  for (int32_t i = 0; i < s4; i++) {
    double* cv = D->short_val + s1 + f13 + (s3 + i) * MK;
    int32_t* cc = D->short_cid + s1 + f13 + (s3 + i) * MK;
    int32_t r = rid4[i];
    for (int t = 0; t < 4; t++) { cv[t] = val[rpt[r] + t]; cc[t] = cid[rpt[r] + t]; }
  }
  /* dasp.cu:1030-1043 2&2 tiles */
  for (int32_t i = 0; i < short_block22; i++) {
    double* cv = D->short_val + s1 + f13 + f34 + i * MM * MK;
    int32_t* cc = D->short_cid + s1 + f13 + f34 + i * MM * MK;
    for (int32_t j = 0; j < BS * 2 && i * BS * 2 + j < s2; j++) {
      int32_t r = rid2[i * BS * 2 + j];
      int32_t o = j % BS * MK + (j / BS) * 2;
      cv[o] = val[rpt[r]];
      cv[o + 1] = val[rpt[r] + 1];
      cc[o] = cid[rpt[r]];
      cc[o + 1] = cid[rpt[r] + 1];
    }
  }
  /* dasp.cu:1244 stable descending sort of the row-block rows by length */
  dasp_radix_desc(rptA, ridA, rb);
  /* dasp.cu:1284-1287 */
  {
    int32_t old = rptA[0];
    rptA[0] = 0;
    for (int32_t i = 1; i <= rb; i++) { int32_t nv = rptA[i]; rptA[i] = old + rptA[i - 1]; old = nv; }
    old = long_rpt[0];
    long_rpt[0] = 0;
    for (int32_t i = 1; i <= rl; i++) { int32_t nv = long_rpt[i]; long_rpt[i] = old + long_rpt[i - 1]; old = nv; }
  }
  int32_t nnz_long = long_rpt[rl];
  /* dasp.cu:1290-1306 order_rid */
  D->order_rid = (int32_t*)calloc(rowA + 1, sizeof(int32_t));
  {
    int32_t* o = D->order_rid;
    memcpy(o, ridl, sizeof(int32_t) * rl);
    memcpy(o + rl, ridA, sizeof(int32_t) * rb);
    memcpy(o + rl + rb, rid1, sizeof(int32_t) * s1);
    for (int32_t i = 0; i < short_block13; i++) {
      int32_t* co = o + rl + rb + s1 + i * BS * 2;
      for (int32_t j = 0; j < BS; j++) {
        co[j] = rid1[s1 + i * BS + j];
        co[BS + j] = rid3[i * BS + j];
      }
    }
    int32_t base = rl + rb + s1 + c13 * 2;
    memcpy(o + base, rid3 + c13, sizeof(int32_t) * s3);
    memcpy(o + base + s3, rid4, sizeof(int32_t) * s4);
    memcpy(o + base + s3 + s4, rid2, sizeof(int32_t) * s2);
    memcpy(o + base + s3 + s4 + s2, ridz, sizeof(int32_t) * rz);
  }
  /* dasp.cu:1330-1369 long rows: 64-nonzero warps, CTA-rounded */
  D->long_rpt = (int32_t*)calloc(rl + 1, sizeof(int32_t));
  for (int32_t i = 0; i < rl; i++) {
    int32_t nn = long_rpt[i + 1] - long_rpt[i];
    D->long_rpt[i] = (nn + MM * MK * LNL - 1) / (MM * MK * LNL);
  }
  {
    int32_t old = D->long_rpt[0];
    D->long_rpt[0] = 0;
    for (int32_t i = 1; i <= rl; i++) { int32_t nv = D->long_rpt[i]; D->long_rpt[i] = old + D->long_rpt[i - 1]; old = nv; }
  }
  int32_t warp_number = D->long_rpt[rl];
  int32_t BlockNum_long = (warp_number + WNL - 1) / WNL;
  int32_t fill0_nnz_long = BlockNum_long * WNL * LNL * MM * MK;
  warp_number = BlockNum_long * WNL;
  D->long_val = (double*)calloc(fill0_nnz_long + 1, sizeof(double));
  D->long_cid = (int32_t*)calloc(fill0_nnz_long + 1, sizeof(int32_t));
  D->rid_by_warp = (int32_t*)calloc(warp_number + 1, sizeof(int32_t));
  for (int32_t i = 0; i < rl; i++) {
    double* cv = D->long_val + (int64_t)D->long_rpt[i] * LNL * MM * MK;
    int32_t* cc = D->long_cid + (int64_t)D->long_rpt[i] * LNL * MM * MK;
    int32_t r = ridl[i];
    for (int32_t j = 0; j < long_rpt[i + 1] - long_rpt[i]; j++) {
      cv[j] = val[rpt[r] + j];
      cc[j] = cid[rpt[r] + j];
    }
    for (int32_t j = D->long_rpt[i]; j < D->long_rpt[i + 1]; j++) D->rid_by_warp[j] = i;
  }
  /* dasp.cu:1374-1416 row blocks: regular slabs of 4 columns while the slab
   * holds at least threshold*32 nonzeros, the rest of each row irregular */
  int32_t blocknum = (rb + BS - 1) / BS;
  blocknum = ((blocknum + ROWLOOP * 4 - 1) / (ROWLOOP * 4)) * ROWLOOP * 4;
  D->block_ptr = (int32_t*)calloc(blocknum + 1, sizeof(int32_t));
  D->irreg_rpt = (int32_t*)calloc(rb + 1, sizeof(int32_t));
  for (int32_t i = 0; i < blocknum; i++) {
    int32_t row_start = i * BS;
    int32_t row_end = (i + 1) * BS >= rb ? rb : (i + 1) * BS;
    int32_t k = 1;
    while (1) {
      int32_t block_nnz = 0;
      for (int32_t r = row_start; r < row_end; r++) {
        int32_t L = rptA[r + 1] - rptA[r];
        if (L / MK >= k) block_nnz += MK;
        else if (L / MK == k - 1) block_nnz += L % MK;
      }
      if (block_nnz >= threshold * MK * MM) {
        D->block_ptr[i] += MK * MM;
      } else {
        for (int32_t r = row_start; r < row_end; r++) {
          int32_t L = rptA[r + 1] - rptA[r];
          D->irreg_rpt[r] = L - (k - 1) * MK > 0 ? L - (k - 1) * MK : 0;
        }
        break;
      }
      k++;
    }
  }
  {
    int32_t old = D->block_ptr[0];
    D->block_ptr[0] = 0;
    for (int32_t i = 1; i <= blocknum; i++) { int32_t nv = D->block_ptr[i]; D->block_ptr[i] = old + D->block_ptr[i - 1]; old = nv; }
    old = D->irreg_rpt[0];
    D->irreg_rpt[0] = 0;
    for (int32_t i = 1; i <= rb; i++) { int32_t nv = D->irreg_rpt[i]; D->irreg_rpt[i] = old + D->irreg_rpt[i - 1]; old = nv; }
  }
  int32_t fill0_nnz_reg = D->block_ptr[blocknum];
  int32_t nnz_irreg = D->irreg_rpt[rb];
  /* dasp.cu:1424-1436 irregular: the last irreg_len entries of each row */
  D->irreg_val = (double*)calloc(nnz_irreg + 1, sizeof(double));
  D->irreg_cid = (int32_t*)calloc(nnz_irreg + 1, sizeof(int32_t));
  for (int32_t i = 0; i < rb; i++) {
    int32_t r = ridA[i], off = D->irreg_rpt[i], il = D->irreg_rpt[i + 1] - off;
    for (int32_t j = 0; j < il; j++) {
      D->irreg_val[off + j] = val[rpt[r + 1] - il + j];
      D->irreg_cid[off + j] = cid[rpt[r + 1] - il + j];
    }
  }
  /* dasp.cu:1439-1487 regular: first blocklen entries of each row, zero
   * padded, then slab-major reorder */
  D->reg_val = (double*)calloc(fill0_nnz_reg + 1, sizeof(double));
  D->reg_cid = (int32_t*)calloc(fill0_nnz_reg + 1, sizeof(int32_t));
  double* tv = (double*)malloc(sizeof(double) * (fill0_nnz_reg + 1));
  int32_t* tc = (int32_t*)malloc(sizeof(int32_t) * (fill0_nnz_reg + 1));
  for (int32_t bid = 0; bid < blocknum; bid++) {
    int32_t nnz_block = D->block_ptr[bid + 1] - D->block_ptr[bid];
    int32_t blocklen = nnz_block / BS;
    for (int32_t rowid = bid * BS; rowid < (bid + 1) * BS; rowid++) {
      int32_t st = D->block_ptr[bid] + blocklen * (rowid - bid * BS);
      if (rowid < rb) {
        int32_t r = ridA[rowid], a0 = rpt[r], L = rpt[r + 1] - a0;
        for (int32_t i = 0; i < blocklen; i++) {
          D->reg_val[st + i] = i < L ? val[a0 + i] : 0.0;
          D->reg_cid[st + i] = i < L ? cid[a0 + i] : 0;
        }
      } else {
        for (int32_t i = 0; i < blocklen; i++) {
          D->reg_val[st + i] = 0.0;
          D->reg_cid[st + i] = 0;
        }
      }
    }
    double* cv = D->reg_val + D->block_ptr[bid];
    int32_t* cc = D->reg_cid + D->block_ptr[bid];
    for (int32_t i = 0; i < nnz_block; i++) {
      int32_t nid = ((i % blocklen) / MK) * BS * MK + (i / blocklen) * MK + i % MK;
      tv[nid] = cv[i];
      tc[nid] = cc[i];
    }
    memcpy(cv, tv, sizeof(double) * nnz_block);
    memcpy(cc, tc, sizeof(int32_t) * nnz_block);
  }
  free(tv);
  free(tc);
  D->row_long = rl;
  D->row_block = rb;
  D->row_zero = rz;
  D->short_row_1 = s1;
  D->short_row_2 = s2;
  D->short_row_3 = s3;
  D->short_row_4 = s4;
  D->common_13 = c13;
  D->fill0_nnz_short13 = f13;
  D->fill0_nnz_short34 = f34;
  D->fill0_nnz_short22 = f22;
  D->fill0_nnz_short = fshort;
  D->threadblock13 = tb13;
  D->threadblock34 = tb34;
  D->threadblock22 = tb22;
  D->blocknum = blocknum;
  D->fill0_nnz_reg = fill0_nnz_reg;
  D->nnz_irreg = nnz_irreg;
  D->warp_number = warp_number;
  D->fill0_nnz_long = fill0_nnz_long;
  D->nnz_long = nnz_long;
  (void)nnz_short;
  free(rid1); free(rid2); free(rid3); free(rid4); free(ridl); free(ridz); free(ridA);
  free(rptA); free(long_rpt);
  return D;
}

void orc_dasp_free(orc_dasp* D) {
  if (!D) return;
  free(D->order_rid); free(D->short_val); free(D->short_cid); free(D->long_rpt);
  free(D->long_val); free(D->long_cid); free(D->rid_by_warp); free(D->block_ptr);
  free(D->irreg_rpt); free(D->reg_val); free(D->reg_cid); free(D->irreg_val);
  free(D->irreg_cid);
  free(D);
}

/* ------------------------------------------------------------------------ */
/* FP64 oracles for the SpMM path (no counterpart in the reference's CPU
 * code: the reference checks nothing for SpMM, bench/spmmfp16.cpp:75-79).
 * Textbook definitions, row-major dense operands with F columns. */

/* Y = A X: Y[i,:] = sum_j A_ij X[j,:], entries in CSR order. */
void orc_csr_spmm_f64(int64_t nrows, const int64_t* rowptr, const int32_t* col, const double* val,
                      const double* X, int64_t F, double* Y) {
  for (int64_t i = 0; i < nrows; i++) {
    double* y = Y + i * F;
    for (int64_t f = 0; f < F; f++) y[f] = 0.0;
    for (int64_t j = rowptr[i]; j < rowptr[i + 1]; j++) {
      const double a = val[j];
      const double* x = X + (int64_t)col[j] * F;
      for (int64_t f = 0; f < F; f++) y[f] += a * x[f];
    }
  }
}

/* dX = A^T dY: dX[c,:] = sum over entries (i,c) of A_ic dY[i,:]. */
void orc_csr_spmm_t_f64(int64_t nrows, int64_t ncols, const int64_t* rowptr, const int32_t* col,
                        const double* val, const double* dY, int64_t F, double* dX) {
  for (int64_t k = 0; k < ncols * F; k++) dX[k] = 0.0;
  for (int64_t i = 0; i < nrows; i++)
    for (int64_t j = rowptr[i]; j < rowptr[i + 1]; j++) {
      const double a = val[j];
      double* d = dX + (int64_t)col[j] * F;
      const double* g = dY + i * F;
      for (int64_t f = 0; f < F; f++) d[f] += a * g[f];
    }
}

/* dvals[j] = dY[row_j,:] . X[col_j,:] for every entry (CSR order). */
void orc_sddmm_f64(int64_t nrows, const int64_t* rowptr, const int32_t* col, const double* dY,
                   const double* X, int64_t F, double* dvals) {
  for (int64_t i = 0; i < nrows; i++)
    for (int64_t j = rowptr[i]; j < rowptr[i + 1]; j++) {
      const double* g = dY + i * F;
      const double* x = X + (int64_t)col[j] * F;
      double s = 0.0;
      for (int64_t f = 0; f < F; f++) s += g[f] * x[f];
      dvals[j] = s;
    }
}

/* Block-sparse C = A B with A in BSR (bs x bs dense blocks, row-major,
 * `brows` block rows) and dense B (ncols x N, row-major): FP64 block GEMM. */
void orc_bsr_spmm_f64(int64_t brows, int32_t bs, const int32_t* browptr, const int32_t* bcol,
                      const double* bval, const double* B, int64_t N, double* Cm) {
  for (int64_t k = 0; k < brows * bs * N; k++) Cm[k] = 0.0;
  for (int64_t br = 0; br < brows; br++)
    for (int32_t b = browptr[br]; b < browptr[br + 1]; b++) {
      const double* a = bval + (int64_t)b * bs * bs;
      for (int32_t i = 0; i < bs; i++) {
        double* c = Cm + (br * bs + i) * N;
        for (int32_t k = 0; k < bs; k++) {
          const double av = a[i * bs + k];
          const double* brow = B + ((int64_t)bcol[b] * bs + k) * N;
          for (int64_t n = 0; n < N; n++) c[n] += av * brow[n];
        }
      }
    }
}

/* ------------------------------------------------------------------------ */
/* Block statistics: bench/blockinfo.cpp:96-119. `blocks` counts the distinct
 * (row / b, col / b) keys; `packed` counts the distinct (row / b, k / b)
 * keys, k being an entry's position within its row (the "recorder" count).
 * Keys are 64-bit here (the reference's int key row/b*1000005 + col/b
 * overflows once row/b reaches 2148). */
static int cmp_i64(const void* a, const void* b) {
  const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
  return (x > y) - (x < y);
}

static int64_t distinct_i64(int64_t* k, int64_t n) {
  if (n == 0) return 0;
  qsort(k, (size_t)n, sizeof(int64_t), cmp_i64);
  int64_t d = 1;
  for (int64_t i = 1; i < n; i++) d += k[i] != k[i - 1];
  return d;
}

void orc_blockinfo(int64_t nrows, int64_t ncols, const int64_t* rowptr, const int32_t* col,
                   int32_t b, int64_t* blocks, int64_t* packed) {
  const int64_t nnz = rowptr[nrows];
  const int64_t span = ncols / b + 2;
  int64_t* k = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nnz ? nnz : 1));
  for (int64_t i = 0; i < nrows; i++)
    for (int64_t j = rowptr[i]; j < rowptr[i + 1]; j++) k[j] = (i / b) * span + col[j] / b;
  *blocks = distinct_i64(k, nnz);
  for (int64_t i = 0; i < nrows; i++)
    for (int64_t j = rowptr[i]; j < rowptr[i + 1]; j++)
      k[j] = (i / b) * span + (j - rowptr[i]) / b;
  *packed = distinct_i64(k, nnz);
  free(k);
}
