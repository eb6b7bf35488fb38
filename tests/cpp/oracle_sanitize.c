/*
 * Drives every entry point of the CPU oracle (oracle/groupby_oracle.c) over
 * seeded inputs -- empty, single-row and ragged sizes, every key and value
 * storage kind the oracle takes, nulls, NaN and -0.0, strings of length 0,
 * every aggregation kind and both null policies -- for a build with
 * AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_oracle_sanitize.py).
 * The memory-checking counterpart of the reference's compute-sanitizer test
 * runs (ci/run_cudf_memcheck_ctests.sh:12-20), for the test oracle.
 *
 * Besides running clean under the sanitizers, each call is checked for
 * invariants that hold whatever the data: group count <= kept rows,
 * representative rows in range, COUNT_ALL summing to the kept rows, and
 * COUNT_VALID <= COUNT_ALL per group.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct orc_col {
  int32_t type_id;
  int32_t size;
  const void* data;
  const uint8_t* valid;
  const int64_t* offsets;
} orc_col;

uint32_t orc_murmur3_x86_32(const uint8_t* data, int64_t len, uint32_t seed);
void orc_hash_rows(const orc_col* cols, int32_t ncols, uint32_t seed, uint32_t elem_seed, uint32_t* out);
int64_t orc_groupby(const orc_col* keys, int32_t nkeys, int32_t null_policy, const orc_col* values,
                    const int32_t* value_index, const int32_t* kinds, const int32_t* ddofs, int32_t naggs,
                    int64_t* rep_rows, void** outs, uint8_t** out_valid);
int64_t orc_group_quantiles(const orc_col* keys, int32_t nkeys, int32_t null_policy, const orc_col* values,
                            const double* q, int32_t nq, int32_t interp, int64_t* rep_rows, double* out,
                            uint8_t* out_valid);

enum { T_INT8 = 1, T_INT16, T_INT32, T_INT64, T_UINT8, T_UINT16, T_UINT32, T_UINT64, T_FLOAT32, T_FLOAT64,
       T_BOOL8, T_STRING = 23 };
enum { A_SUM = 0, A_PRODUCT = 1, A_MIN = 2, A_MAX = 3, A_COUNT_VALID = 4, A_COUNT_ALL = 5,
       A_SUM_OF_SQUARES = 8, A_MEAN = 9, A_VARIANCE = 11, A_STD = 12, A_ARGMAX = 15, A_ARGMIN = 16,
       A_NUNIQUE = 17, A_NTH_ELEMENT = 18 };

static uint64_t rng_state;
static uint64_t next_u64(void)
{
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return rng_state;
}

static int failures = 0;
#define CHECK(cond, ...)                      \
  do {                                        \
    if (!(cond)) {                            \
      fprintf(stderr, "check failed: " __VA_ARGS__); \
      fprintf(stderr, "\n");                  \
      failures++;                             \
    }                                         \
  } while (0)

typedef struct {
  orc_col col;
  void* data;
  uint8_t* valid;
  int64_t* offsets;
} owned;

static int width_of(int t)
{
  switch (t) {
    case T_INT8: case T_UINT8: case T_BOOL8: return 1;
    case T_INT16: case T_UINT16: return 2;
    case T_INT32: case T_UINT32: case T_FLOAT32: return 4;
    default: return 8;
  }
}

/* A column of n rows drawn over `card` distinct values (strings: lengths 0..12). */
static owned make_col(int t, int32_t n, int64_t card, double null_frac)
{
  owned o;
  memset(&o, 0, sizeof(o));
  o.col.type_id = t;
  o.col.size    = n;
  if (null_frac > 0) {
    o.valid = (uint8_t*)malloc((size_t)(n > 0 ? n : 1));
    for (int32_t i = 0; i < n; i++) o.valid[i] = (double)(next_u64() % 10000) / 10000.0 >= null_frac;
  }
  if (t == T_STRING) {
    o.offsets = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    char* chars = (char*)malloc((size_t)n * 12 + 1);
    o.offsets[0] = 0;
    for (int32_t i = 0; i < n; i++) {
      uint64_t v  = next_u64() % (uint64_t)card;
      int len     = (int)(v % 13);
      for (int j = 0; j < len; j++) chars[o.offsets[i] + j] = (char)('a' + (v >> (j % 8)) % 26);
      o.offsets[i + 1] = o.offsets[i] + len;
    }
    o.data = chars;
  } else {
    int w = width_of(t);
    uint8_t* d = (uint8_t*)malloc((size_t)(n > 0 ? n : 1) * (size_t)w);
    for (int32_t i = 0; i < n; i++) {
      uint64_t v = next_u64() % (uint64_t)card;
      if (t == T_FLOAT32) {
        float f = (v % 17 == 0) ? NAN : ((v % 13 == 0) ? -0.0f : (float)v - (float)card / 2);
        memcpy(d + (size_t)i * 4, &f, 4);
      } else if (t == T_FLOAT64) {
        double f = (v % 17 == 0) ? NAN : ((v % 13 == 0) ? -0.0 : (double)v * 0.25 - (double)card / 8);
        memcpy(d + (size_t)i * 8, &f, 8);
      } else if (t == T_BOOL8) {
        d[i] = (uint8_t)(v & 1);
      } else {
        int64_t s = (int64_t)v - (int64_t)(card / 2);
        memcpy(d + (size_t)i * (size_t)w, &s, (size_t)w);  /* little-endian low bytes */
      }
    }
    o.data = d;
  }
  o.col.data    = o.data;
  o.col.valid   = o.valid;
  o.col.offsets = o.offsets;
  return o;
}

static void free_col(owned* o)
{
  free(o->data);
  free(o->valid);
  free(o->offsets);
}

static void run_case(int32_t n, int nkeys, const int* key_types, int64_t card, double key_nulls, int policy)
{
  owned keys[2];
  orc_col kc[2];
  for (int k = 0; k < nkeys; k++) {
    keys[k] = make_col(key_types[k], n, card, key_nulls);
    kc[k]   = keys[k].col;
  }
  static const int vtypes[] = {T_INT8, T_INT16, T_INT32, T_INT64, T_UINT8, T_UINT16, T_UINT32, T_UINT64,
                               T_FLOAT32, T_FLOAT64, T_BOOL8, T_STRING};
  enum { NV = sizeof(vtypes) / sizeof(vtypes[0]) };
  owned vals[NV];
  orc_col vc[NV];
  for (int v = 0; v < NV; v++) {
    vals[v] = make_col(vtypes[v], n, 1000, v % 2 ? 0.2 : 0.0);
    vc[v]   = vals[v].col;
  }
  static const int kinds_all[] = {A_SUM, A_PRODUCT, A_MIN, A_MAX, A_COUNT_VALID, A_COUNT_ALL, A_SUM_OF_SQUARES,
                                  A_MEAN, A_VARIANCE, A_STD, A_ARGMAX, A_ARGMIN, A_NUNIQUE, A_NUNIQUE,
                                  A_NTH_ELEMENT, A_NTH_ELEMENT, A_NTH_ELEMENT};
  static const int ddof_all[] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 2 * 0 + 1, 2 * 2 + 0, 2 * -1 + 0};
  enum { NK = sizeof(kinds_all) / sizeof(kinds_all[0]) };
  int64_t* rep   = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  for (int v = 0; v < NV; v++) {
    int kinds[NK], ddofs[NK], vidx[NK], na = 0;
    for (int a = 0; a < NK; a++) {
      int k = kinds_all[a];
      if (vtypes[v] == T_STRING && !(k == A_COUNT_VALID || k == A_COUNT_ALL || k == A_ARGMIN || k == A_ARGMAX ||
                                     k == A_NUNIQUE))
        continue;
      kinds[na] = k;
      ddofs[na] = ddof_all[a];
      vidx[na]  = v;
      na++;
    }
    void* outs[NK];
    uint8_t* ov[NK];
    for (int a = 0; a < na; a++) {
      outs[a] = calloc((size_t)(n > 0 ? n : 1), 8);
      ov[a]   = (uint8_t*)calloc((size_t)(n > 0 ? n : 1), 1);
    }
    int64_t G = orc_groupby(kc, nkeys, policy, vc, vidx, kinds, ddofs, na, rep, outs, ov);
    CHECK(G >= 0, "orc_groupby rejected value type %d", vtypes[v]);
    CHECK(G <= n, "groups %lld > rows %d", (long long)G, n);
    int64_t kept = 0;
    for (int32_t r = 0; r < n; r++) {
      int any_null = 0;
      for (int k = 0; k < nkeys; k++) any_null |= kc[k].valid != NULL && !kc[k].valid[r];
      kept += !(policy == 0 && any_null);
    }
    for (int64_t g = 0; g < G; g++) CHECK(rep[g] >= 0 && rep[g] < n, "representative row out of range");
    int a_all = -1, a_valid = -1;
    for (int a = 0; a < na; a++) {
      if (kinds[a] == A_COUNT_ALL) a_all = a;
      if (kinds[a] == A_COUNT_VALID) a_valid = a;
    }
    if (G >= 0 && a_all >= 0) {
      int64_t total = 0;
      for (int64_t g = 0; g < G; g++) {
        int32_t c_all = ((int32_t*)outs[a_all])[g];
        total += c_all;
        if (a_valid >= 0) CHECK(((int32_t*)outs[a_valid])[g] <= c_all, "COUNT_VALID > COUNT_ALL");
      }
      CHECK(total == kept, "COUNT_ALL sums to %lld, kept rows %lld", (long long)total, (long long)kept);
    }
    for (int a = 0; a < na; a++) {
      free(outs[a]);
      free(ov[a]);
    }
    if (vtypes[v] != T_STRING && vtypes[v] != T_BOOL8) {
      static const double q[] = {0.0, 0.25, 0.5, 1.0};
      double* qo   = (double*)malloc(sizeof(double) * 4 * (size_t)(n > 0 ? n : 1));
      uint8_t* qv  = (uint8_t*)malloc(4 * (size_t)(n > 0 ? n : 1));
      for (int interp = 0; interp < 5; interp++) {
        int64_t Gq = orc_group_quantiles(kc, nkeys, policy, &vc[v], q, 4, interp, rep, qo, qv);
        CHECK(Gq == G, "quantile groups %lld vs %lld", (long long)Gq, (long long)G);
      }
      free(qo);
      free(qv);
    }
  }
  uint32_t* h = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  orc_hash_rows(kc, nkeys, 0, 0, h);
  free(h);
  free(rep);
  for (int v = 0; v < NV; v++) free_col(&vals[v]);
  for (int k = 0; k < nkeys; k++) free_col(&keys[k]);
}

int main(void)
{
  rng_state = 0x9E3779B97F4A7C15ull;
  uint8_t bytes[64];
  for (int i = 0; i < 64; i++) bytes[i] = (uint8_t)next_u64();
  for (int len = 0; len <= 64; len++) (void)orc_murmur3_x86_32(bytes + (64 - len), len, (uint32_t)len);
  static const int32_t sizes[] = {0, 1, 2, 7, 63, 64, 65, 1000, 5003};
  static const int key_sets[][2] = {{T_INT32, 0}, {T_INT64, 0}, {T_FLOAT64, 0}, {T_STRING, 0},
                                    {T_INT32, T_STRING}, {T_UINT8, T_FLOAT32}};
  static const int nkeys_of[] = {1, 1, 1, 1, 2, 2};
  int cases = 0;
  for (size_t s = 0; s < sizeof(sizes) / sizeof(sizes[0]); s++)
    for (size_t k = 0; k < sizeof(nkeys_of) / sizeof(nkeys_of[0]); k++)
      for (int policy = 0; policy < 2; policy++) {
        run_case(sizes[s], nkeys_of[k], key_sets[k], sizes[s] > 100 ? 97 : 5, (k % 2) ? 0.1 : 0.0, policy);
        cases++;
      }
  printf("oracle sanitizer run: %d cases, %d failed checks\n", cases, failures);
  return failures == 0 ? 0 : 1;
}
