/*
 * groupby_oracle.c -- CPU restatement of libcudf's hash group-by, used ONLY as
 * the test oracle (pytest, smoke check and the CPU leg of the benchmark).  The
 * product path never links or calls this file.
 *
 * Each function cites the reference file:line whose behaviour it restates
 * (paths relative to the cudf source tree, version 24.10).
 *
 * Pinning (what checks each function against something other than itself):
 *   orc_murmur3_x86_32   smhasher's published verification value and the
 *                        reference's Python docstring row hashes
 *                        (python/cudf/cudf/core/indexed_frame.py:2894-2923)
 *   orc_row_hash         the same docstring vectors (single and two-column)
 *   orc_groupby          the known-answer tests of cpp/tests/groupby/
 *                        {sum,count,min,max,mean,var,keys}_tests.cpp, brute
 *                        force over tiny inputs, and pandas group-by on the
 *                        same seeded data (tests/test_oracle.py); NUNIQUE
 *                        by cpp/tests/groupby/nunique_tests.cpp and a
 *                        set-based brute force; NTH_ELEMENT by
 *                        cpp/tests/groupby/nth_element_tests.cpp and a
 *                        per-group list brute force
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* type ids as in cudf/types.hpp:203-235 */
enum {
  T_INT8 = 1, T_INT16, T_INT32, T_INT64, T_UINT8, T_UINT16, T_UINT32, T_UINT64,
  T_FLOAT32, T_FLOAT64, T_BOOL8, T_TS_D, T_TS_S, T_TS_MS, T_TS_US, T_TS_NS,
  T_DUR_D, T_DUR_S, T_DUR_MS, T_DUR_US, T_DUR_NS, T_DICT32, T_STRING = 23
};
/* aggregation kinds as in cudf/aggregation.hpp:86-122 */
enum {
  A_SUM = 0, A_PRODUCT = 1, A_MIN = 2, A_MAX = 3, A_COUNT_VALID = 4, A_COUNT_ALL = 5,
  A_SUM_OF_SQUARES = 8, A_MEAN = 9, A_VARIANCE = 11, A_STD = 12, A_ARGMAX = 15, A_ARGMIN = 16,
  A_NUNIQUE = 17, A_NTH_ELEMENT = 18
};

typedef struct orc_col {
  int32_t type_id;
  int32_t size;
  const void* data;       /* elements, or chars for strings          */
  const uint8_t* valid;   /* one byte per row (1 = valid); NULL = all */
  const int64_t* offsets; /* strings: size+1 offsets into data        */
} orc_col;

static int type_width(int t)
{
  switch (t) {
    case T_INT8: case T_UINT8: case T_BOOL8: return 1;
    case T_INT16: case T_UINT16: return 2;
    case T_INT32: case T_UINT32: case T_FLOAT32: case T_TS_D: case T_DUR_D: case T_DICT32:
      return 4;
    case T_INT64: case T_UINT64: case T_FLOAT64: case T_TS_S: case T_TS_MS: case T_TS_US:
    case T_TS_NS: case T_DUR_S: case T_DUR_MS: case T_DUR_US: case T_DUR_NS: return 8;
    default: return 0;
  }
}
static int is_float(int t) { return t == T_FLOAT32 || t == T_FLOAT64; }
static int is_signed_int(int t)
{
  return t == T_INT8 || t == T_INT16 || t == T_INT32 || t == T_INT64 || (t >= T_TS_D && t <= T_DUR_NS);
}
static int is_unsigned_int(int t) { return t == T_UINT8 || t == T_UINT16 || t == T_UINT32 || t == T_UINT64 || t == T_BOOL8; }

/* ---- MurmurHash3_x86_32: hashing/detail/murmurhash3_x86_32.cuh:40-130 ---- */
static uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
static uint32_t fmix32(uint32_t h)
{
  h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
  return h;
}

uint32_t orc_murmur3_x86_32(const uint8_t* data, int64_t len, uint32_t seed)
{
  uint32_t h = seed;
  int64_t nblocks = len / 4;
  for (int64_t i = 0; i < nblocks; i++) {
    const uint8_t* b = data + 4 * i;
    uint32_t k = (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
    k *= 0xcc9e2d51u; k = rotl32(k, 15); k *= 0x1b873593u;
    h ^= k; h = rotl32(h, 13); h = h * 5 + 0xe6546b64u;
  }
  const uint8_t* tail = data + 4 * nblocks;
  uint32_t k = 0;
  switch (len & 3) {
    case 3: k ^= (uint32_t)tail[2] << 16; /* fall through */
    case 2: k ^= (uint32_t)tail[1] << 8;  /* fall through */
    case 1: k ^= tail[0]; k *= 0xcc9e2d51u; k = rotl32(k, 15); k *= 0x1b873593u; h ^= k;
  }
  h ^= (uint32_t)len;
  return fmix32(h);
}

/* hashing/detail/hashing.hpp:85-88 (boost hash_combine) */
uint32_t orc_hash_combine(uint32_t lhs, uint32_t rhs)
{
  return lhs ^ (rhs + 0x9e3779b9u + (lhs << 6) + (lhs >> 2));
}

static int row_valid(const orc_col* c, int64_t i) { return c->valid == NULL || c->valid[i]; }

/* Element hash: element_hasher (table/experimental/row_operators.cuh:1824-1830)
 * with floats normalised first (hash_functions.cuh:28-47: -0.0 -> 0.0, any
 * NaN -> quiet NaN), bool hashed as one byte, strings over their bytes. */
static uint32_t element_hash(const orc_col* c, int64_t i, uint32_t seed)
{
  if (!row_valid(c, i)) return 0xFFFFFFFFu;
  if (c->type_id == T_STRING) {
    int64_t b = c->offsets[i], e = c->offsets[i + 1];
    return orc_murmur3_x86_32((const uint8_t*)c->data + b, e - b, seed);
  }
  if (c->type_id == T_FLOAT32) {
    float v = ((const float*)c->data)[i];
    if (v == 0.0f) v = 0.0f;
    if (isnan(v)) v = NAN;
    return orc_murmur3_x86_32((const uint8_t*)&v, 4, seed);
  }
  if (c->type_id == T_FLOAT64) {
    double v = ((const double*)c->data)[i];
    if (v == 0.0) v = 0.0;
    if (isnan(v)) v = (double)NAN;
    return orc_murmur3_x86_32((const uint8_t*)&v, 8, seed);
  }
  if (c->type_id == T_BOOL8) {
    uint8_t v = ((const uint8_t*)c->data)[i] ? 1 : 0;
    return orc_murmur3_x86_32(&v, 1, seed);
  }
  int w = type_width(c->type_id);
  return orc_murmur3_x86_32((const uint8_t*)c->data + (int64_t)w * i, w, seed);
}

/* device_row_hasher::operator() (row_operators.cuh:1869-1883): accumulate
 * hash_combine over the columns starting from the seed.  `elem_seed` is the
 * seed given to each element hash (the reference passes the same seed). */
uint32_t orc_row_hash(const orc_col* cols, int32_t ncols, int64_t row, uint32_t seed, uint32_t elem_seed)
{
  uint32_t h = seed;
  for (int c = 0; c < ncols; c++) h = orc_hash_combine(h, element_hash(&cols[c], row, elem_seed));
  return h;
}

void orc_hash_rows(const orc_col* cols, int32_t ncols, uint32_t seed, uint32_t elem_seed, uint32_t* out)
{
  if (ncols == 0) return;
  for (int64_t r = 0; r < cols[0].size; r++) out[r] = orc_row_hash(cols, ncols, r, seed, elem_seed);
}

/* Element equality (row_operators.cuh:1429-1448, 1285-1298): nulls equal each
 * other, NaN equals NaN, -0.0 equals 0.0, strings compare bytes. */
static int element_equal(const orc_col* c, int64_t a, int64_t b)
{
  int va = row_valid(c, a), vb = row_valid(c, b);
  if (!va || !vb) return va == vb;
  if (c->type_id == T_STRING) {
    int64_t la = c->offsets[a + 1] - c->offsets[a], lb = c->offsets[b + 1] - c->offsets[b];
    if (la != lb) return 0;
    return memcmp((const char*)c->data + c->offsets[a], (const char*)c->data + c->offsets[b], la) == 0;
  }
  if (c->type_id == T_FLOAT32) {
    float x = ((const float*)c->data)[a], y = ((const float*)c->data)[b];
    if (isnan(x) && isnan(y)) return 1;
    return x == y;
  }
  if (c->type_id == T_FLOAT64) {
    double x = ((const double*)c->data)[a], y = ((const double*)c->data)[b];
    if (isnan(x) && isnan(y)) return 1;
    return x == y;
  }
  if (c->type_id == T_BOOL8) {
    return (((const uint8_t*)c->data)[a] != 0) == (((const uint8_t*)c->data)[b] != 0);
  }
  int w = type_width(c->type_id);
  return memcmp((const char*)c->data + (int64_t)w * a, (const char*)c->data + (int64_t)w * b, w) == 0;
}

static int rows_equal(const orc_col* cols, int ncols, int64_t a, int64_t b)
{
  for (int c = 0; c < ncols; c++)
    if (!element_equal(&cols[c], a, b)) return 0;
  return 1;
}

/* Read element i of a numeric column as long double / int64 / uint64. */
static long double as_ld(const orc_col* c, int64_t i)
{
  const void* d = c->data;
  switch (c->type_id) {
    case T_INT8: return ((const int8_t*)d)[i];
    case T_INT16: return ((const int16_t*)d)[i];
    case T_INT32: case T_TS_D: case T_DUR_D: return ((const int32_t*)d)[i];
    case T_INT64: case T_TS_S: case T_TS_MS: case T_TS_US: case T_TS_NS:
    case T_DUR_S: case T_DUR_MS: case T_DUR_US: case T_DUR_NS: return (long double)((const int64_t*)d)[i];
    case T_UINT8: return ((const uint8_t*)d)[i];
    case T_BOOL8: return ((const uint8_t*)d)[i] ? 1 : 0;
    case T_UINT16: return ((const uint16_t*)d)[i];
    case T_UINT32: return ((const uint32_t*)d)[i];
    case T_UINT64: return (long double)((const uint64_t*)d)[i];
    case T_FLOAT32: return ((const float*)d)[i];
    case T_FLOAT64: return ((const double*)d)[i];
  }
  return 0;
}
static int64_t as_i64(const orc_col* c, int64_t i)
{
  const void* d = c->data;
  switch (c->type_id) {
    case T_INT8: return ((const int8_t*)d)[i];
    case T_INT16: return ((const int16_t*)d)[i];
    case T_INT32: case T_TS_D: case T_DUR_D: return ((const int32_t*)d)[i];
    case T_UINT8: return ((const uint8_t*)d)[i];
    case T_BOOL8: return ((const uint8_t*)d)[i] ? 1 : 0;
    case T_UINT16: return ((const uint16_t*)d)[i];
    case T_UINT32: return ((const uint32_t*)d)[i];
    case T_UINT64: return (int64_t)((const uint64_t*)d)[i];
    default: return ((const int64_t*)d)[i];
  }
}

/* Target type of (source type, kind): detail/aggregation/aggregation.hpp:1211-1331. */
int32_t orc_target_type(int32_t src, int32_t kind)
{
  switch (kind) {
    case A_MIN: case A_MAX: return src;
    case A_NTH_ELEMENT: return src == T_STRING ? -1 : src;
    case A_COUNT_VALID: case A_COUNT_ALL: return T_INT32;
    case A_ARGMIN: case A_ARGMAX: return T_INT32; /* size_type row index, aggregation.hpp:1346-1357 */
    case A_NUNIQUE: return T_INT32;               /* size_type count, group_nunique.cu:90-91 */
    case A_MEAN: case A_VARIANCE: case A_STD: return T_FLOAT64;
    case A_SUM: case A_SUM_OF_SQUARES: case A_PRODUCT:
      if (is_float(src)) return src;
      if (is_signed_int(src) || is_unsigned_int(src)) {
        /* durations sum to themselves (aggregation.hpp:1296-1301) */
        if (kind == A_SUM && src >= T_DUR_D && src <= T_DUR_NS) return src;
        return T_INT64;
      }
      return -1;
  }
  return -1;
}

/* Order of two valid values of a column for ARGMIN / ARGMAX: <0, 0, >0.
 * Integers by value, strings bytewise (then by length), floats by value with
 * NaN after every number (a NaN row never replaces a number). */
static int value_cmp(const orc_col* c, int64_t a, int64_t b)
{
  int t = c->type_id;
  if (t == T_STRING) {
    int64_t la = c->offsets[a + 1] - c->offsets[a], lb = c->offsets[b + 1] - c->offsets[b];
    const uint8_t* pa = (const uint8_t*)c->data + c->offsets[a];
    const uint8_t* pb = (const uint8_t*)c->data + c->offsets[b];
    int64_t m = la < lb ? la : lb;
    for (int64_t i = 0; i < m; i++)
      if (pa[i] != pb[i]) return pa[i] < pb[i] ? -1 : 1;
    return la < lb ? -1 : (la > lb ? 1 : 0);
  }
  if (is_float(t)) {
    long double x = as_ld(c, a), y = as_ld(c, b);
    int nx = isnan(x), ny = isnan(y);
    if (nx || ny) return nx == ny ? 0 : (nx ? 1 : -1);
    return x < y ? -1 : (x > y ? 1 : 0);
  }
  if (is_unsigned_int(t)) {
    uint64_t x = (uint64_t)as_i64(c, a), y = (uint64_t)as_i64(c, b);
    return x < y ? -1 : (x > y ? 1 : 0);
  }
  int64_t x = as_i64(c, a), y = as_i64(c, b);
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* NUNIQUE (cpp/src/groupby/sort/group_nunique.cu:44-64): the reference sorts
 * each group's values and counts a row when it is countable (valid, or any
 * row under null_policy INCLUDE) and it opens the group or differs from its
 * predecessor; nulls equal each other, NaN equals NaN and -0.0 equals 0.0
 * (the row comparator's defaults).  Restated the same way: the group's rows
 * sorted by (group, null first, value) and the heads of equal runs counted. */
static _Thread_local const orc_col* g_nu_col;  /* per thread: calls run concurrently */
static _Thread_local const int64_t* g_nu_group;
static int nunique_cmp(const void* pa, const void* pb)
{
  int64_t a = *(const int64_t*)pa, b = *(const int64_t*)pb;
  if (g_nu_group[a] != g_nu_group[b]) return g_nu_group[a] < g_nu_group[b] ? -1 : 1;
  int va = row_valid(g_nu_col, a), vb = row_valid(g_nu_col, b);
  if (va != vb) return va < vb ? -1 : 1;
  if (!va) return 0;
  return value_cmp(g_nu_col, a, b);
}

static void nunique_counts(const orc_col* v, const int64_t* group_of_row, int64_t n, int include_nulls,
                           int64_t G, int32_t* out)
{
  int64_t* rows = (int64_t*)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  int64_t m = 0;
  for (int64_t r = 0; r < n; r++)
    if (group_of_row[r] >= 0 && (include_nulls || row_valid(v, r))) rows[m++] = r;
  g_nu_col = v;
  g_nu_group = group_of_row;
  qsort(rows, (size_t)m, sizeof(int64_t), nunique_cmp);
  for (int64_t g = 0; g < G; g++) out[g] = 0;
  for (int64_t i = 0; i < m; i++)
    if (i == 0 || nunique_cmp(&rows[i - 1], &rows[i]) != 0) out[group_of_row[rows[i]]]++;
  free(rows);
}

/* ---- open-addressing map row -> group, first-occurrence group order ---- */
typedef struct { int64_t* slots; int64_t cap; } omap;

/* Group-by in the reference's semantics (hash/groupby.cu:548-622 with
 * groupby_kernels.cuh:98-105, aggregation.cuh:131-490 and
 * multi_pass_kernels.cuh:74-111):
 *   - rows whose key has any null are dropped under null_policy EXCLUDE (0)
 *     (hash/groupby.cu:479-485); under INCLUDE (1) nulls form groups;
 *   - SUM/SUM_OF_SQUARES/MIN/MAX/MEAN/VARIANCE/STD skip null values and are
 *     null when no value of the group is valid; COUNT_VALID counts valid
 *     values; COUNT_ALL counts rows;
 *   - integer SUM wraps in int64; float SUM is exact (long double) rounded to
 *     the target type once, i.e. the value an exact sum would give;
 *   - MIN/MAX keep the identity (+/-inf, type max/lowest) against NaN inputs,
 *     exactly like the CAS loop of device_atomics.cuh:107-153 with
 *     std::min/std::max (device_operators.cuh:43-58);
 *   - MEAN = double(sum) / count (hash/groupby.cu:287-306, binary_ops DIV in
 *     double);
 *   - VARIANCE = sum((x - mean)^2) / (n - ddof), null when n - ddof <= 0,
 *     two passes in long double (multi_pass_kernels.cuh:74-95);
 *   - PRODUCT multiplies the valid values (int64 for integers, wrapping;
 *     floats exact and rounded once), null when none is valid
 *     (aggregation.cuh update_target_element PRODUCT, DeviceProduct);
 *   - ARGMIN / ARGMAX = row index of the smallest / largest valid value,
 *     null when the group has none (aggregation.cuh:440-490).  The reference
 *     keeps whichever of equal values reaches the target first; this
 *     restatement takes the first row in input order, and for floats a NaN
 *     only when every value of the group is NaN (NaN compares false, so a
 *     number is never replaced by a NaN).
 * Outputs: `rep_rows[g]` = first row of group g; `outs[a]` = array of G
 * elements of the target type; `out_valid[a]` = G validity bytes.
 * Returns G, or -1 on an unsupported (type, kind). */
/* Groups of the rows in first-occurrence order: group_of_row[r] (-1 for a
 * dropped row), rep_rows[g] = first row of group g; returns G. */
static int64_t assign_groups(const orc_col* keys, int32_t nkeys, int32_t null_policy, int64_t* rep_rows,
                             int64_t* group_of_row)
{
  int64_t n = nkeys > 0 ? keys[0].size : 0;
  int64_t cap = 16;
  while (cap < 2 * n + 16) cap <<= 1;
  int64_t* slots = (int64_t*)malloc(sizeof(int64_t) * cap);
  int64_t* slot_group = (int64_t*)malloc(sizeof(int64_t) * cap);
  for (int64_t i = 0; i < cap; i++) slots[i] = -1;
  int64_t G = 0;
  for (int64_t r = 0; r < n; r++) {
    group_of_row[r] = -1;
    if (null_policy == 0) {
      int any_null = 0;
      for (int c = 0; c < nkeys; c++) any_null |= !row_valid(&keys[c], r);
      if (any_null) continue;
    }
    uint32_t h = orc_row_hash(keys, nkeys, r, 0, 0);
    int64_t s = h & (cap - 1);
    for (;;) {
      if (slots[s] < 0) {
        slots[s] = r; slot_group[s] = G; rep_rows[G] = r; group_of_row[r] = G; G++;
        break;
      }
      if (rows_equal(keys, nkeys, slots[s], r)) { group_of_row[r] = slot_group[s]; break; }
      s = (s + 1) & (cap - 1);
    }
  }
  free(slots);
  free(slot_group);
  return G;
}

int64_t orc_groupby(const orc_col* keys, int32_t nkeys, int32_t null_policy,
                    const orc_col* values, const int32_t* value_index,
                    const int32_t* kinds, const int32_t* ddofs, int32_t naggs,
                    int64_t* rep_rows, void** outs, uint8_t** out_valid)
{
  int64_t n = nkeys > 0 ? keys[0].size : 0;
  int64_t* group_of_row = (int64_t*)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  int64_t G = assign_groups(keys, nkeys, null_policy, rep_rows, group_of_row);

  long double* acc = (long double*)malloc(sizeof(long double) * (G > 0 ? G : 1));
  uint64_t* iacc = (uint64_t*)malloc(sizeof(uint64_t) * (G > 0 ? G : 1));
  int64_t* cnt = (int64_t*)malloc(sizeof(int64_t) * (G > 0 ? G : 1));
  int64_t* cnt_all = (int64_t*)malloc(sizeof(int64_t) * (G > 0 ? G : 1));
  int ok = 1;
  for (int a = 0; a < naggs && ok; a++) {
    const orc_col* v = &values[value_index[a]];
    int k = kinds[a];
    int src = v->type_id;
    int tgt = orc_target_type(src, k);
    if (tgt < 0 || (src == T_STRING && k != A_COUNT_VALID && k != A_COUNT_ALL && k != A_ARGMIN &&
                    k != A_ARGMAX && k != A_NUNIQUE)) { ok = 0; break; }
    if (k == A_NTH_ELEMENT) {
      /* group_nth_element.cu:50-140: the n-th row of the group (from the end
       * when n < 0) in row order, counting every row under INCLUDE and valid
       * rows under EXCLUDE; null when the group has too few such rows, and
       * the value's own null under INCLUDE.  ddofs[a] = 2 * n + include. */
      int include = ddofs[a] & 1;
      int64_t nth = (int64_t)(ddofs[a] - include) / 2;
      int64_t* total = (int64_t*)calloc((size_t)(G > 0 ? G : 1), sizeof(int64_t));
      int64_t* seen = (int64_t*)calloc((size_t)(G > 0 ? G : 1), sizeof(int64_t));
      int64_t* pick = (int64_t*)malloc(sizeof(int64_t) * (G > 0 ? G : 1));
      for (int64_t g = 0; g < G; g++) pick[g] = -1;
      for (int64_t r = 0; r < n; r++)
        if (group_of_row[r] >= 0 && (include || row_valid(v, r))) total[group_of_row[r]]++;
      for (int64_t r = 0; r < n; r++) {
        int64_t g = group_of_row[r];
        if (g < 0 || !(include || row_valid(v, r))) continue;
        int64_t want = nth >= 0 ? nth : total[g] + nth;
        if (seen[g]++ == want) pick[g] = r;
      }
      int w = type_width(src);
      for (int64_t g = 0; g < G; g++) {
        int ok_g = pick[g] >= 0 && row_valid(v, pick[g]);
        if (ok_g) memcpy((uint8_t*)outs[a] + g * w, (const uint8_t*)v->data + pick[g] * w, (size_t)w);
        else memset((uint8_t*)outs[a] + g * w, 0, (size_t)w);
        out_valid[a][g] = (uint8_t)ok_g;
      }
      free(total); free(seen); free(pick);
      continue;
    }
    if (k == A_NUNIQUE) {  /* ddofs[a] holds the null policy: 0 EXCLUDE, 1 INCLUDE */
      nunique_counts(v, group_of_row, n, ddofs[a] != 0, G, (int32_t*)outs[a]);
      for (int64_t g = 0; g < G; g++) out_valid[a][g] = 1;
      continue;
    }
    if (k == A_ARGMIN || k == A_ARGMAX) {
      int64_t* best = (int64_t*)malloc(sizeof(int64_t) * (G > 0 ? G : 1));
      for (int64_t g = 0; g < G; g++) best[g] = -1;
      for (int64_t r = 0; r < n; r++) {
        int64_t g = group_of_row[r];
        if (g < 0 || !row_valid(v, r)) continue;
        if (best[g] < 0) { best[g] = r; continue; }
        int c = value_cmp(v, r, best[g]);
        if (k == A_ARGMIN ? c < 0 : c > 0) {
          /* for ARGMAX a NaN orders after numbers but must not win: */
          if (k == A_ARGMAX && is_float(src) && isnan(as_ld(v, r))) continue;
          best[g] = r;
        } else if (k == A_ARGMAX && is_float(src) && isnan(as_ld(v, best[g])) && !isnan(as_ld(v, r))) {
          best[g] = r;
        }
      }
      for (int64_t g = 0; g < G; g++) {
        ((int32_t*)outs[a])[g] = best[g] < 0 ? 0 : (int32_t)best[g];
        out_valid[a][g] = (uint8_t)(best[g] >= 0);
      }
      free(best);
      continue;
    }
    uint8_t* ov = out_valid[a];
    for (int64_t g = 0; g < G; g++) { acc[g] = 0; iacc[g] = 0; cnt[g] = 0; cnt_all[g] = 0; }
    int fl = is_float(src);
    if (k == A_MIN || k == A_MAX) {
      for (int64_t g = 0; g < G; g++) acc[g] = (k == A_MIN) ? INFINITY : -INFINITY;
    }
    if (k == A_PRODUCT) {
      for (int64_t g = 0; g < G; g++) { acc[g] = 1; iacc[g] = 1; }
    }
    /* first pass */
    for (int64_t r = 0; r < n; r++) {
      int64_t g = group_of_row[r];
      if (g < 0) continue;
      cnt_all[g]++;
      if (!row_valid(v, r)) continue;
      cnt[g]++;
      long double x = as_ld(v, r);
      switch (k) {
        case A_SUM: case A_MEAN: case A_VARIANCE: case A_STD:
          if (fl) acc[g] += x; else iacc[g] += (uint64_t)as_i64(v, r);
          break;
        case A_PRODUCT:
          /* integers multiply in int64 with wrap-around; floats exactly
             (long double) and round once, as for SUM */
          if (fl) acc[g] *= x; else iacc[g] *= (uint64_t)as_i64(v, r);
          break;
        case A_SUM_OF_SQUARES:
          if (fl) {
            if (src == T_FLOAT32) { float f = (float)x; acc[g] += (long double)(f * f); }
            else { double d = (double)x; acc[g] += (long double)(d * d); }
          } else { uint64_t u = (uint64_t)as_i64(v, r); iacc[g] += u * u; }
          break;
        case A_MIN:
          if (fl) { if (x < acc[g]) acc[g] = x; }
          else if (cnt[g] == 1 || (is_unsigned_int(src) ? (uint64_t)as_i64(v, r) < iacc[g]
                                                         : as_i64(v, r) < (int64_t)iacc[g]))
            iacc[g] = (uint64_t)as_i64(v, r);
          break;
        case A_MAX:
          if (fl) { if (x > acc[g]) acc[g] = x; }
          else if (cnt[g] == 1 || (is_unsigned_int(src) ? (uint64_t)as_i64(v, r) > iacc[g]
                                                         : as_i64(v, r) > (int64_t)iacc[g]))
            iacc[g] = (uint64_t)as_i64(v, r);
          break;
        default: break;
      }
    }
    /* finalize */
    for (int64_t g = 0; g < G; g++) {
      int valid = cnt[g] > 0;
      switch (k) {
        case A_COUNT_VALID: ((int32_t*)outs[a])[g] = (int32_t)cnt[g]; valid = 1; break;
        case A_COUNT_ALL: ((int32_t*)outs[a])[g] = (int32_t)cnt_all[g]; valid = 1; break;
        case A_SUM: case A_SUM_OF_SQUARES: case A_PRODUCT:
          if (tgt == T_FLOAT32) ((float*)outs[a])[g] = (float)acc[g];
          else if (tgt == T_FLOAT64) ((double*)outs[a])[g] = (double)acc[g];
          else if (type_width(tgt) == 8) ((int64_t*)outs[a])[g] = (int64_t)iacc[g];
          else ((int32_t*)outs[a])[g] = (int32_t)iacc[g];
          break;
        case A_MIN: case A_MAX: {
          int w = type_width(src);
          if (fl) {
            if (src == T_FLOAT32) ((float*)outs[a])[g] = (float)acc[g];
            else ((double*)outs[a])[g] = (double)acc[g];
          } else if (w == 8) ((uint64_t*)outs[a])[g] = iacc[g];
          else if (w == 4) ((uint32_t*)outs[a])[g] = (uint32_t)iacc[g];
          else if (w == 2) ((uint16_t*)outs[a])[g] = (uint16_t)iacc[g];
          else ((uint8_t*)outs[a])[g] = (uint8_t)iacc[g];
          break;
        }
        case A_MEAN: {
          double s;
          if (fl) s = (src == T_FLOAT32) ? (double)(float)acc[g] : (double)acc[g];
          else s = (double)(int64_t)iacc[g];
          ((double*)outs[a])[g] = valid ? s / (double)cnt[g] : 0.0;
          break;
        }
        default: break;
      }
      ov[g] = (uint8_t)valid;
    }
    if (k == A_VARIANCE || k == A_STD) {
      /* second pass: sum of squared deviations from the (long double) mean */
      long double* m2 = (long double*)calloc((size_t)(G > 0 ? G : 1), sizeof(long double));
      long double* mean = (long double*)malloc(sizeof(long double) * (G > 0 ? G : 1));
      for (int64_t g = 0; g < G; g++) {
        long double s = fl ? acc[g] : (long double)(int64_t)iacc[g];
        mean[g] = cnt[g] > 0 ? s / cnt[g] : 0;
      }
      for (int64_t r = 0; r < n; r++) {
        int64_t g = group_of_row[r];
        if (g < 0 || !row_valid(v, r)) continue;
        long double d = as_ld(v, r) - mean[g];
        m2[g] += d * d;
      }
      for (int64_t g = 0; g < G; g++) {
        int64_t dn = cnt[g] - ddofs[a];
        int valid = cnt[g] > 0 && dn > 0;
        double var = valid ? (double)(m2[g] / dn) : 0.0;
        ((double*)outs[a])[g] = (k == A_STD) ? sqrt(var) : var;
        ov[g] = (uint8_t)valid;
      }
      free(m2);
      free(mean);
    }
  }
  free(acc); free(iacc); free(cnt); free(cnt_all); free(group_of_row);
  return ok ? G : -1;
}

/* ---- QUANTILE / MEDIAN --------------------------------------------------
 * sort/group_quantiles.cu:30-80 and quantiles/quantiles_util.hpp: each
 * group's valid values sorted ascending (NaN after every number); with
 * n = valid count, q clamped to [0, 1] and v = q * (n - 1):
 * lower = floor(v), higher = ceil(v), nearest = nearbyint(v) (ties to even),
 * fraction = v - lower;  LINEAR (1 - f) * x[lower] + f * x[higher] in
 * double;  MIDPOINT x[lower] / 2 + x[higher] / 2 in double, except for
 * int64 values: half = a / 2 + b / 2, rest = a % 2 + b % 2 (C division),
 * result = half + rest * 0.5;  LOWER / HIGHER / NEAREST the element.  Every
 * result is double; null when the group has no valid value.  MEDIAN is
 * QUANTILE {0.5} LINEAR (sort/aggregate.cpp:344-362).  Outputs are group-
 * major: out[g * nq + j] for quantile j of group g. */
enum { I_LINEAR = 0, I_LOWER, I_HIGHER, I_MIDPOINT, I_NEAREST };

/* the column q_cmp orders by: one per thread, since the sharded oracle runs
 * several calls at once */
static _Thread_local const orc_col* g_q_col;
static int q_cmp(const void* pa, const void* pb)
{
  int64_t a = *(const int64_t*)pa, b = *(const int64_t*)pb;
  const orc_col* c = g_q_col;
  if (is_float(c->type_id)) {
    long double x = as_ld(c, a), y = as_ld(c, b);
    int nx = isnan(x), ny = isnan(y);
    if (nx || ny) return nx - ny;
    return (x > y) - (x < y);
  }
  if (c->type_id == T_UINT64) {
    uint64_t x = (uint64_t)as_i64(c, a), y = (uint64_t)as_i64(c, b);
    return (x > y) - (x < y);
  }
  int64_t x = as_i64(c, a), y = as_i64(c, b);
  return (x > y) - (x < y);
}

static double q_elem(const orc_col* c, int64_t row)
{
  return (double)as_ld(c, row);
}

int64_t orc_group_quantiles(const orc_col* keys, int32_t nkeys, int32_t null_policy, const orc_col* values,
                            const double* q, int32_t nq, int32_t interp, int64_t* rep_rows, double* out,
                            uint8_t* out_valid)
{
  int t = values->type_id;
  if (!(is_float(t) || is_signed_int(t) || (is_unsigned_int(t) && t != T_BOOL8))) return -1;
  if (interp < I_LINEAR || interp > I_NEAREST) return -1;
  int64_t n = nkeys > 0 ? keys[0].size : 0;
  int64_t* group_of_row = (int64_t*)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  int64_t G = assign_groups(keys, nkeys, null_policy, rep_rows, group_of_row);
  /* rows of each group's valid values, grouped by a counting sort */
  int64_t* start = (int64_t*)calloc((size_t)G + 1, sizeof(int64_t));
  for (int64_t r = 0; r < n; r++)
    if (group_of_row[r] >= 0 && row_valid(values, r)) start[group_of_row[r] + 1]++;
  for (int64_t g = 0; g < G; g++) start[g + 1] += start[g];
  int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * ((size_t)G + 1));
  memcpy(fill, start, sizeof(int64_t) * ((size_t)G + 1));
  int64_t* rows = (int64_t*)malloc(sizeof(int64_t) * (size_t)(start[G] > 0 ? start[G] : 1));
  for (int64_t r = 0; r < n; r++)
    if (group_of_row[r] >= 0 && row_valid(values, r)) rows[fill[group_of_row[r]]++] = r;
  g_q_col = values;
  for (int64_t g = 0; g < G; g++) {
    int64_t m = start[g + 1] - start[g];
    int64_t* x = rows + start[g];
    qsort(x, (size_t)m, sizeof(int64_t), q_cmp);
    for (int32_t j = 0; j < nq; j++) {
      double r = 0.0;
      if (m > 0) {
        double qq = q[j] < 0.0 ? 0.0 : (q[j] > 1.0 ? 1.0 : q[j]);
        double v = qq * (double)(m - 1);
        int64_t lo = (int64_t)floor(v), hi = (int64_t)ceil(v), ne = (int64_t)nearbyint(v);
        double f = v - (double)lo;
        switch (interp) {
          case I_LINEAR: r = (1.0 - f) * q_elem(values, x[lo]) + f * q_elem(values, x[hi]); break;
          case I_LOWER: r = q_elem(values, x[lo]); break;
          case I_HIGHER: r = q_elem(values, x[hi]); break;
          case I_NEAREST: r = q_elem(values, x[ne]); break;
          case I_MIDPOINT:
            if (t == T_INT64) {
              int64_t a = as_i64(values, x[lo]), b = as_i64(values, x[hi]);
              int64_t half = a / 2 + b / 2, rest = a % 2 + b % 2;
              r = (double)half + (double)rest * 0.5;
            } else {
              r = q_elem(values, x[lo]) / 2 + q_elem(values, x[hi]) / 2;
            }
            break;
        }
      }
      out[g * nq + j] = r;
      out_valid[g * nq + j] = (uint8_t)(m > 0);
    }
  }
  free(start); free(fill); free(rows); free(group_of_row);
  return G;
}
