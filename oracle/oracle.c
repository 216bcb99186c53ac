/*
 * oracle.c — CPU restatement of AccSpMM's host algorithms, for TESTING ONLY.
 *
 * Test infrastructure: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, never by the product library.  Every
 * function follows the reference file:line given in its comment; the checks
 * that pin each one are listed in oracle.h.
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_CW 8u
#define ORC_PAD 0xFFFFFFFFu

void orc_free(void* p) { free(p); }

/* ---------------------------------------------------------------------- */
/* Matrix Market: include/mmio.hpp:24-72 (banner), :74-93 (size line),     */
/* :106-168 (entries + symmetric expansion), :171-210 (driver).            */
/* ---------------------------------------------------------------------- */
enum { ORC_OK = 0, ORC_LOAD_ERR = -1, ORC_BANNER_ERR = -2, ORC_READ_ERR = 3 };

/* typecode slots: [0] M, [1] C/A, [2] R/C/P/I, [3] G/S/H/K
 * (include/common.hpp:114-140).  Returns 0 when accepted. */
static int orc_banner(FILE* f, char code[4]) {
  char line[1025], banner[64], mtx[64], crd[64], dt[64], st[64];
  code[0] = code[1] = code[2] = ' ';
  code[3] = 'G'; /* mm_clear_typecode */
  if (!fgets(line, sizeof line, f)) return 12;
  if (sscanf(line, "%63s %63s %63s %63s %63s", banner, mtx, crd, dt, st) != 5)
    return 12;
  if (strcmp(banner, "%%MatrixMarket") != 0) return 14;
  if (strcmp(mtx, "matrix") != 0) return 15;
  code[0] = 'M';
  /* an unknown storage token is accepted silently (mmio.hpp:54-55) */
  if (strcmp(crd, "coordinate") == 0) code[1] = 'C';
  else if (strcmp(crd, "array") == 0) code[1] = 'A';
  if (strcmp(dt, "real") == 0) code[2] = 'R';
  else if (strcmp(dt, "complex") == 0) code[2] = 'C';
  else if (strcmp(dt, "pattern") == 0) code[2] = 'P';
  else if (strcmp(dt, "integer") == 0) code[2] = 'I';
  else return 15;
  if (strcmp(st, "general") == 0) code[3] = 'G';
  else if (strcmp(st, "symmetric") == 0) code[3] = 'S';
  else if (strcmp(st, "Hermitian") == 0) code[3] = 'H';
  else if (strcmp(st, "skew-symmetric") == 0) code[3] = 'K';
  else return 15;
  return 0;
}

static int orc_size_line(FILE* f, uint32_t* m, uint32_t* n, uint32_t* nz) {
  char line[1025];
  *m = *n = *nz = 0;
  do {
    if (!fgets(line, sizeof line, f)) return 12;
  } while (line[0] == '%');
  if (sscanf(line, "%u %u %u", m, n, nz) == 3) return 0;
  for (;;) {
    int k = fscanf(f, "%u %u %u", m, n, nz);
    if (k == EOF) return 12;
    if (k == 3) return 0;
  }
}

int orc_read_mtx(const char* path, uint32_t* rows, uint32_t* cols,
                 uint32_t* nnz, uint32_t** r_out, uint32_t** c_out,
                 float** v_out) {
  FILE* f = fopen(path, "r");
  if (!f) return ORC_LOAD_ERR;
  char code[4];
  if (orc_banner(f, code) != 0) {
    fclose(f);
    return ORC_BANNER_ERR;
  }
  int isPattern = code[2] == 'P', isReal = !isPattern && code[2] == 'R';
  int isComplex = !isPattern && !isReal && code[2] == 'C';
  int isInteger = !isPattern && !isReal && !isComplex && code[2] == 'I';
  uint32_t m, n, nz;
  if (orc_size_line(f, &m, &n, &nz) != 0) {
    fclose(f);
    return ORC_READ_ERR;
  }
  int sym = code[3] == 'S' || code[3] == 'H';
  uint32_t* tr = (uint32_t*)malloc(sizeof(uint32_t) * (nz ? nz : 1));
  uint32_t* tc = (uint32_t*)malloc(sizeof(uint32_t) * (nz ? nz : 1));
  float* tv = (float*)malloc(sizeof(float) * (nz ? nz : 1));
  uint32_t row_id = 0, col_id = 0, diag = 0;
  double data_f = 0.0, data_im = 0.0;
  for (uint32_t i = 0; i < nz; ++i) {
    /* fscanf results are ignored as in mmio.hpp:122-125: a short read keeps
     * the previous values. */
    if (isReal) {
      (void)!fscanf(f, "%u %u %lg\n", &row_id, &col_id, &data_f);
    } else if (isComplex) {
      (void)!fscanf(f, "%u %u %lg %lg\n", &row_id, &col_id, &data_f, &data_im);
    } else if (isInteger) {
      /* The reference reads %d into a double (defect D11: garbage values);
       * this restatement parses the integer, as the product does. */
      int iv = 0;
      if (fscanf(f, "%u %u %d\n", &row_id, &col_id, &iv) == 3) data_f = iv;
    } else if (isPattern) {
      (void)!fscanf(f, "%u %u\n", &row_id, &col_id);
      data_f = 1.0;
    }
    tr[i] = row_id - 1;
    tc[i] = col_id - 1;
    tv[i] = (float)data_f;
    if (row_id == col_id) ++diag;
  }
  fclose(f);
  uint32_t total = sym ? nz * 2 - diag : nz;
  uint32_t* R = (uint32_t*)malloc(sizeof(uint32_t) * (total ? total : 1));
  uint32_t* C = (uint32_t*)malloc(sizeof(uint32_t) * (total ? total : 1));
  float* V = (float*)malloc(sizeof(float) * (total ? total : 1));
  uint32_t k = 0;
  for (uint32_t i = 0; i < nz; ++i) {
    R[k] = tr[i]; C[k] = tc[i]; V[k] = tv[i]; ++k;
    if (sym && tr[i] != tc[i]) {
      R[k] = tc[i]; C[k] = tr[i]; V[k] = tv[i]; ++k;
    }
  }
  free(tr); free(tc); free(tv);
  *rows = m; *cols = n; *nnz = total;
  *r_out = R; *c_out = C; *v_out = V;
  return ORC_OK;
}

/* include/format.hpp:110-142 */
void orc_coo_to_csr(uint32_t rows, uint32_t nnz, const uint32_t* r,
                    const uint32_t* c, const float* v, uint32_t* row_ptr,
                    uint32_t* col_idx, float* val) {
  memset(row_ptr, 0, sizeof(uint32_t) * (rows + 1));
  for (uint32_t i = 0; i < nnz; ++i) row_ptr[r[i] + 1]++;
  for (uint32_t i = 0; i < rows; ++i) row_ptr[i + 1] += row_ptr[i];
  for (uint32_t i = 0; i < nnz; ++i) {
    uint32_t dest = row_ptr[r[i]]++;
    val[dest] = v[i];
    col_idx[dest] = c[i];
  }
  for (uint32_t i = rows; i > 0; --i) row_ptr[i] = row_ptr[i - 1];
  row_ptr[0] = 0;
}

/* ---------------------------------------------------------------------- */
/* METCF / AdpBME                                                          */
/* ---------------------------------------------------------------------- */
static int cmp_u32(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return x < y ? -1 : x > y;
}

/* growable arrays */
typedef struct { void* p; size_t n, cap, es; } vec_t;
static void vinit(vec_t* v, size_t es) { v->p = NULL; v->n = v->cap = 0; v->es = es; }
static void* vgrow(vec_t* v, size_t add) {
  if (v->n + add > v->cap) {
    size_t nc = v->cap ? v->cap : 64;
    while (nc < v->n + add) nc *= 2;
    v->p = realloc(v->p, nc * v->es);
    v->cap = nc;
  }
  void* at = (char*)v->p + v->n * v->es;
  v->n += add;
  return at;
}

/* Shared tile construction of METCFBit::convertFromCSR (format.hpp:215-274)
 * and AdpBME::CSR2AdpBME (:340-429); the two produce identical
 * sparseA2B / tcOffset / tcLocalBit / data.  When `adp` is set the greedy
 * grouping of :366-382 and :418-423 and rowIndices (:364) are added. */
static orc_metcf* orc_build(uint32_t rw, uint32_t rows,
                            const uint32_t* row_ptr, const uint32_t* col_idx,
                            const float* val, int adp, uint32_t t_load,
                            uint32_t t_comp, uint32_t t_write,
                            uint32_t target_num) {
  if (rw != 8 && rw != 16) return NULL;
  const uint32_t words = rw / 8; /* tc_tmp_len (format.hpp:219) */
  orc_metcf* m = (orc_metcf*)calloc(1, sizeof(orc_metcf));
  m->rw = rw;
  vec_t rwo, tco, bits, a2b, dat, grp, rix;
  vinit(&rwo, 4); vinit(&tco, 4); vinit(&bits, 8); vinit(&a2b, 4);
  vinit(&dat, 4); vinit(&grp, 4); vinit(&rix, 4);
  *(uint32_t*)vgrow(&rwo, 1) = 0;
  *(uint32_t*)vgrow(&tco, 1) = 0; /* the leading 0 of format.hpp:272 */
  if (adp) *(uint32_t*)vgrow(&grp, 1) = 0;
  const uint32_t target = target_num * (t_load + t_comp) + t_write;
  uint32_t gtime = 0, gsize = 0;
  uint32_t* uniq = NULL;
  size_t uniq_cap = 0;
  for (uint32_t iter = 0; iter < rows; iter += rw) {
    uint32_t wend = iter + rw < rows ? iter + rw : rows;
    uint32_t bs = row_ptr[iter], be = row_ptr[wend];
    uint32_t len = be - bs;
    /* extractUniqueColumns (format.hpp:179-199): sort + unique */
    if (len + ORC_CW > uniq_cap) {
      uniq_cap = (len + ORC_CW) * 2;
      uniq = (uint32_t*)realloc(uniq, uniq_cap * sizeof(uint32_t));
    }
    memcpy(uniq, col_idx + bs, len * sizeof(uint32_t));
    qsort(uniq, len, sizeof(uint32_t), cmp_u32);
    uint32_t u = 0;
    for (uint32_t i = 0; i < len; ++i)
      if (u == 0 || uniq[u - 1] != uniq[i]) uniq[u++] = uniq[i];
    /* padListLength (format.hpp:23-30) */
    uint32_t padded = (u + ORC_CW - 1) / ORC_CW * ORC_CW;
    for (uint32_t i = u; i < padded; ++i) uniq[i] = ORC_PAD;
    memcpy(vgrow(&a2b, padded), uniq, padded * sizeof(uint32_t));
    uint32_t ntc = padded / ORC_CW;
    uint32_t base = ((uint32_t*)rwo.p)[rwo.n - 1];
    *(uint32_t*)vgrow(&rwo, 1) = base + ntc;
    if (adp) {
      for (uint32_t i = 0; i < ntc; ++i) *(uint32_t*)vgrow(&rix, 1) = iter;
      for (uint32_t i = 0; i < ntc; ++i) {
        uint32_t add = i == 0 ? t_load + t_comp + t_write : t_load + t_comp;
        if (gsize > 0 && gtime + add > target) {
          *(uint32_t*)vgrow(&grp, 1) = gsize;
          gtime = 0;
          gsize = 0;
        }
        gtime += add;
        gsize++;
      }
    }
    /* per-tile counts, bitmaps, values in CSR order (format.hpp:237-269) */
    uint32_t* cnt = (uint32_t*)vgrow(&tco, ntc);
    memset(cnt, 0, ntc * sizeof(uint32_t));
    uint64_t* bm = (uint64_t*)vgrow(&bits, (size_t)ntc * words);
    memset(bm, 0, (size_t)ntc * words * sizeof(uint64_t));
    uint32_t* tcount = (uint32_t*)calloc(ntc + 1, sizeof(uint32_t));
    for (uint32_t r = iter; r < wend; ++r)
      for (uint32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
        /* the hash map lookup of :241 == position in the sorted unique list */
        uint32_t lo = 0, hi = u;
        while (lo < hi) {
          uint32_t mid = (lo + hi) / 2;
          if (uniq[mid] < col_idx[e]) lo = mid + 1; else hi = mid;
        }
        uint32_t ci = lo, t = ci / ORC_CW;
        cnt[t]++;
        tcount[t]++;
        uint32_t local = (r % rw) * ORC_CW + ci % ORC_CW;
        bm[(size_t)t * words + local / 64] |= 1ULL << (local % 64);
      }
    /* datatmp: values grouped by tile, CSR order inside a tile */
    uint32_t* start = (uint32_t*)calloc(ntc + 1, sizeof(uint32_t));
    for (uint32_t t = 0; t < ntc; ++t) start[t + 1] = start[t] + tcount[t];
    float* wd = (float*)vgrow(&dat, len);
    for (uint32_t r = iter; r < wend; ++r)
      for (uint32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
        uint32_t lo = 0, hi = u;
        while (lo < hi) {
          uint32_t mid = (lo + hi) / 2;
          if (uniq[mid] < col_idx[e]) lo = mid + 1; else hi = mid;
        }
        wd[start[lo / ORC_CW]++] = val[e];
      }
    free(start);
    free(tcount);
  }
  free(uniq);
  if (adp) {
    if (gsize != 0) *(uint32_t*)vgrow(&grp, 1) = gsize;
    uint32_t* g = (uint32_t*)grp.p;
    for (size_t i = 0; i + 1 < grp.n; ++i) g[i + 1] += g[i];
  }
  uint32_t* tc = (uint32_t*)tco.p;
  for (size_t i = 0; i + 1 < tco.n; ++i) tc[i + 1] += tc[i];
  m->num_windows = (uint32_t)(rwo.n - 1);
  m->num_tiles = (uint32_t)(tco.n - 1);
  m->nnz = (uint32_t)dat.n;
  m->rowWindowOffset = (uint32_t*)rwo.p;
  m->tcOffset = (uint32_t*)tco.p;
  m->tcLocalBit = (uint64_t*)bits.p;
  m->sparseA2B = (uint32_t*)a2b.p;
  m->data = (float*)dat.p;
  if (adp) {
    m->num_groups = (uint32_t)(grp.n - 1);
    m->groupOffset = (uint32_t*)grp.p;
    m->rowIndices = (uint32_t*)rix.p;
  }
  return m;
}

orc_metcf* orc_metcf_build(uint32_t rw, uint32_t rows, const uint32_t* row_ptr,
                           const uint32_t* col_idx, const float* val) {
  return orc_build(rw, rows, row_ptr, col_idx, val, 0, 0, 0, 0, 0);
}

orc_metcf* orc_adpbme_build(uint32_t rw, uint32_t rows,
                            const uint32_t* row_ptr, const uint32_t* col_idx,
                            const float* val, uint32_t t_load,
                            uint32_t t_comp, uint32_t t_write,
                            uint32_t target_num) {
  return orc_build(rw, rows, row_ptr, col_idx, val, 1, t_load, t_comp, t_write,
                   target_num);
}

void orc_metcf_free(orc_metcf* m) {
  if (!m) return;
  free(m->rowWindowOffset); free(m->tcOffset); free(m->tcLocalBit);
  free(m->sparseA2B); free(m->data); free(m->groupOffset);
  free(m->rowIndices); free(m);
}

/* include/utils.hpp:75-85.  Unsigned arithmetic as in the reference: the
 * difference is formed in 32-bit unsigned, cast to int, abs'd, summed in
 * 32-bit unsigned; windows 0..W-2 only (the last window is excluded). */
uint32_t orc_ibd(const uint32_t* rwo, uint32_t W, uint32_t T,
                 uint32_t threshold, int* balanced) {
  if (W == 0) { if (balanced) *balanced = 0; return 0; }
  uint32_t avg = T / W, sum = 0;
  for (uint32_t i = 1; i < W; ++i) {
    uint32_t cur = rwo[i] - rwo[i - 1];
    int32_t d = (int32_t)(uint32_t)(cur - avg);
    sum += (uint32_t)(d < 0 ? -d : d);
  }
  uint32_t ibd = sum / W;
  if (balanced) *balanced = ibd > threshold;
  return ibd;
}

uint32_t orc_metcf_decode(const orc_metcf* m, uint32_t* r, uint32_t* c,
                          float* v) {
  uint32_t k = 0, words = m->rw / 8;
  for (uint32_t w = 0; w < m->num_windows; ++w)
    for (uint32_t t = m->rowWindowOffset[w]; t < m->rowWindowOffset[w + 1];
         ++t) {
      uint32_t rank = 0;
      for (uint32_t b = 0; b < m->rw * 8; ++b) {
        uint64_t word = m->tcLocalBit[(size_t)t * words + b / 64];
        if (!((word >> (b % 64)) & 1)) continue;
        r[k] = w * m->rw + b / 8;
        c[k] = m->sparseA2B[(size_t)t * 8 + b % 8];
        v[k] = m->data[m->tcOffset[t] + rank];
        ++rank;
        ++k;
      }
    }
  return k;
}

/* ---------------------------------------------------------------------- */
/* SpMM oracles and the CPU baseline                                       */
/* ---------------------------------------------------------------------- */
void orc_spmm_f64(uint32_t rows, const uint32_t* row_ptr,
                  const uint32_t* col_idx, const float* val, const float* B,
                  uint32_t ldb, uint32_t K, double* C) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < (int64_t)rows; ++i) {
    double* out = C + (size_t)i * K;
    for (uint32_t k = 0; k < K; ++k) out[k] = 0.0;
    for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
      double a = val[e];
      const float* b = B + (size_t)col_idx[e] * ldb;
      for (uint32_t k = 0; k < K; ++k) out[k] += a * (double)b[k];
    }
  }
}

float orc_tf32_rna(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7F800000u) == 0x7F800000u) return x;  /* infinity, NaN */
  /* 13 mantissa bits are dropped: add half of the lowest kept bit to the
   * magnitude (sign-magnitude, so ties go away from zero), then truncate */
  u = (u + 0x1000u) & 0xFFFFE000u;
  memcpy(&x, &u, 4);
  return x;
}

void orc_spmm_tf32_f64(uint32_t rows, const uint32_t* row_ptr,
                       const uint32_t* col_idx, const float* val, const float* B,
                       uint32_t ldb, uint32_t K, double* C) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < (int64_t)rows; ++i) {
    double* out = C + (size_t)i * K;
    for (uint32_t k = 0; k < K; ++k) out[k] = 0.0;
    for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
      const double a = orc_tf32_rna(val[e]);
      const float* b = B + (size_t)col_idx[e] * ldb;
      for (uint32_t k = 0; k < K; ++k) out[k] += a * (double)orc_tf32_rna(b[k]);
    }
  }
}

void orc_spmm_abs_f64(uint32_t rows, const uint32_t* row_ptr,
                      const uint32_t* col_idx, const float* val,
                      const float* B, uint32_t ldb, uint32_t K, double* C) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < (int64_t)rows; ++i) {
    double* out = C + (size_t)i * K;
    for (uint32_t k = 0; k < K; ++k) out[k] = 0.0;
    for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
      double a = fabs((double)val[e]);
      const float* b = B + (size_t)col_idx[e] * ldb;
      for (uint32_t k = 0; k < K; ++k) out[k] += a * fabs((double)b[k]);
    }
  }
}

void orc_errors(uint64_t rows, uint32_t K, const float* C, uint32_t ldc,
                const double* C64, const double* absAB, double* max_rel,
                double* fro_rel) {
  double mx = 0.0, num = 0.0, den = 0.0;
  for (uint64_t i = 0; i < rows; ++i)
    for (uint32_t k = 0; k < K; ++k) {
      double c = C[i * ldc + k], e = C64[i * K + k], s = absAB[i * K + k];
      double d = fabs(c - e);
      num += d * d;
      den += e * e;
      if (s == 0.0) {
        if (c != 0.0 || c != c) mx = INFINITY;
      } else {
        double r = d / s;
        if (r > mx || r != r) mx = r != r ? INFINITY : r;
      }
    }
  *max_rel = mx;
  *fro_rel = den > 0 ? sqrt(num) / sqrt(den) : sqrt(num);
}

int orc_spmm_f32_omp(uint32_t rows, const uint32_t* row_ptr,
                     const uint32_t* col_idx, const float* val,
                     const float* B, uint32_t ldb, uint32_t K, float* C,
                     int threads) {
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
  for (int64_t i = 0; i < (int64_t)rows; ++i) {
    float* out = C + (size_t)i * K;
    for (uint32_t k = 0; k < K; ++k) out[k] = 0.0f;
    for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
      const float a = val[e];
      const float* __restrict__ b = B + (size_t)col_idx[e] * ldb;
#pragma omp simd
      for (uint32_t k = 0; k < K; ++k) out[k] += a * b[k];
    }
  }
  return threads;
}
