/*
 * CPU oracle of the seeded input generator — TEST INFRASTRUCTURE ONLY.
 *
 * Restates, from the C++ standard's published algorithms, what the
 * reference's DataGenerator (src/data_generator.cuh:12-211) computes with
 * libstdc++ on x86-64:
 *   mt19937                 (C++ [rand.eng.mers], 32-bit Mersenne Twister)
 *   seed_seq::generate      (C++ [rand.util.seedseq])
 *   mt19937(seed_seq&)      (state from 624 generated words, all-zero guard)
 *   generate_canonical<double,53>   two 32-bit draws: (g0 + g1*2^32) / 2^64
 *   uniform_real_distribution<double>  a + u * (b - a)
 *   normal_distribution<double>     Marsaglia polar, returns y*mult and keeps
 *                                   x*mult for the next call
 *   SeedRandomGenerator     1248 words (8-byte result_type * 624 / 4), each
 *                           the seed rotated left once more (:193-205)
 *   128 threads, libgomp static schedule (:210, :42-51 etc.)
 * plus the added distributions of the MI355X benchmark configs
 * (normal-clamped, zipf, zipf-1.2, nearly-sorted-swaps, float-specials).
 *
 * nearly-sorted uses snapshot semantics (every gap from the sorted array);
 * the reference races on each thread's last element (src/data_generator.cuh
 * :140-159), so fixtures mask those positions.
 *
 * Pinned by tests/test_datagen.py against fixtures written by the reference's
 * own data_generator.cuh compiled in oracle/_ref (tests/golden/).
 */
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NTHREADS 128
#define NBUCKETS 10

/* ---------------------------------------------------------------- mt19937 */
typedef struct {
  uint32_t x[624];
  int p;
} mt_t;

static uint32_t mt_next(mt_t* m) {
  if (m->p >= 624) {
    for (int i = 0; i < 624; ++i) {
      const uint32_t y = (m->x[i] & 0x80000000u) | (m->x[(i + 1) % 624] & 0x7FFFFFFFu);
      m->x[i] = m->x[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908B0DFu : 0u);
    }
    m->p = 0;
  }
  uint32_t z = m->x[m->p++];
  z ^= z >> 11;
  z ^= (z << 7) & 0x9D2C5680u;
  z ^= (z << 15) & 0xEFC60000u;
  z ^= z >> 18;
  return z;
}

static uint32_t seq_T(uint32_t v) { return v ^ (v >> 27); }

/* seed_seq::generate over `out` (n words) from v (s words). */
static void seed_seq_generate(const uint32_t* v, size_t s, uint32_t* out, size_t n) {
  for (size_t i = 0; i < n; ++i) out[i] = 0x8b8b8b8bu;
  const size_t t = (n >= 623) ? 11 : (n >= 68) ? 7 : (n >= 39) ? 5 : (n >= 7) ? 3 : (n - 1) / 2;
  const size_t p = (n - t) / 2;
  const size_t q = p + t;
  const size_t m = (s + 1 > n) ? s + 1 : n;
  for (size_t k = 0; k < m; ++k) {
    const uint32_t r1 = 1664525u * seq_T(out[k % n] ^ out[(k + p) % n] ^ out[(k + n - 1) % n]);
    uint32_t r2 = r1;
    if (k == 0)
      r2 += (uint32_t)s;
    else if (k <= s)
      r2 += (uint32_t)(k % n) + v[k - 1];
    else
      r2 += (uint32_t)(k % n);
    out[(k + p) % n] += r1;
    out[(k + q) % n] += r2;
    out[k % n] = r2;
  }
  for (size_t k = m; k < m + n; ++k) {
    const uint32_t r3 = 1566083941u * seq_T(out[k % n] + out[(k + p) % n] + out[(k + n - 1) % n]);
    const uint32_t r4 = r3 - (uint32_t)(k % n);
    out[(k + p) % n] ^= r3;
    out[(k + q) % n] ^= r4;
    out[k % n] = r4;
  }
}

/* SeedRandomGenerator(seed) of the reference. */
static void seed_mt(mt_t* m, uint32_t seed) {
  enum { S = 1248 };
  uint32_t v[S];
  for (int i = 0; i < S; ++i) {
    seed = (seed << 1) | (seed >> 31);
    v[i] = seed;
  }
  seed_seq_generate(v, S, m->x, 624);
  int zero = (m->x[0] & 0x80000000u) == 0;
  for (int i = 1; i < 624 && zero; ++i)
    if (m->x[i] != 0) zero = 0;
  if (zero) m->x[0] = 0x80000000u;
  m->p = 624;
}

static double canonical(mt_t* m) {
  double sum = 0.0, tmp = 1.0;
  for (int k = 0; k < 2; ++k) {
    sum += (double)mt_next(m) * tmp;
    tmp *= 4294967296.0;
  }
  double r = sum / tmp;
  if (r >= 1.0) r = nextafter(1.0, 0.0);
  return r;
}

static double uniform_real(mt_t* m, double a, double b) { return canonical(m) * (b - a) + a; }

typedef struct {
  double saved;
  int has;
} normal_t;

static double normal_draw(normal_t* st, mt_t* m, double mean, double sd) {
  double r;
  if (st->has) {
    st->has = 0;
    r = st->saved;
  } else {
    double x, y, r2;
    do {
      x = 2.0 * canonical(m) - 1.0;
      y = 2.0 * canonical(m) - 1.0;
      r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0.0);
    const double mult = sqrt(-2.0 * log(r2) / r2);
    st->saved = x * mult;
    st->has = 1;
    r = y * mult;
  }
  return r * sd + mean;
}

/* libgomp static schedule: thread t of nt over n iterations. */
static void static_range(size_t n, int nt, int t, size_t* b, size_t* e) {
  size_t q = n / (size_t)nt, r = n % (size_t)nt;
  if ((size_t)t < r) {
    ++q;
    r = 0;
  }
  *b = q * (size_t)t + r;
  *e = *b + q;
}

/* ------------------------------------------------------------ splitmix64 */
static uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
static uint64_t stream_base(uint32_t seed, uint32_t stream) { return splitmix64(((uint64_t)stream << 32) | seed); }
static double unit53(uint64_t r) { return (double)(r >> 11) * 0x1.0p-53; }

/* -------------------------------------------------- conversions per type */
static int32_t cvt_i32(double x) {
  if (!(x >= -2147483648.0 && x < 2147483648.0)) return INT32_MIN;
  return (int32_t)x;
}
static int64_t cvt_i64(double x) {
  if (!(x >= -9223372036854775808.0 && x < 9223372036854775808.0)) return INT64_MIN;
  return (int64_t)x;
}
static uint32_t cvt_u32(double x) {
  if (!(x > -1.0)) return 0;
  if (x >= 4294967296.0) return UINT32_MAX;
  return (uint32_t)x;
}
static uint64_t cvt_u64(double x) {
  if (!(x > -1.0)) return 0;
  if (x >= 18446744073709551616.0) return UINT64_MAX;
  return (uint64_t)x;
}
static float cvt_f32(double x) { return (float)x; }
static double cvt_f64(double x) { return x; }

static int cmp_i32(const void* a, const void* b) { int32_t x = *(const int32_t*)a, y = *(const int32_t*)b; return (x > y) - (x < y); }
static int cmp_i64(const void* a, const void* b) { int64_t x = *(const int64_t*)a, y = *(const int64_t*)b; return (x > y) - (x < y); }
static int cmp_u32(const void* a, const void* b) { uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b; return (x > y) - (x < y); }
static int cmp_u64(const void* a, const void* b) { uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b; return (x > y) - (x < y); }
static int cmp_f32(const void* a, const void* b) { float x = *(const float*)a, y = *(const float*)b; return (x > y) - (x < y); }
static int cmp_f64(const void* a, const void* b) { double x = *(const double*)a, y = *(const double*)b; return (x > y) - (x < y); }

/* Wrapping subtraction in T's arithmetic (signed overflow wraps as on x86). */
#define WRAPSUB_INT(T, U) static T wrapsub_##T(T a, T b) { return (T)(U)((U)a - (U)b); }
WRAPSUB_INT(int32_t, uint32_t)
WRAPSUB_INT(int64_t, uint64_t)
WRAPSUB_INT(uint32_t, uint32_t)
WRAPSUB_INT(uint64_t, uint64_t)
static float wrapsub_float(float a, float b) { return a - b; }
static double wrapsub_double(double a, double b) { return a - b; }

/* One generator per type.  NAME: suffix; T: C type; CVT: double->T;
 * CMP: qsort comparator; MAXV: numeric_limits<T>::max(); MINV: min(). */
#define DEFINE_GEN(NAME, T, CVT, CMP, MAXV, MINV, WSUB, ISFLOAT, ISSIGNED, WIDTH)                          \
  static void uniform_##NAME(T* b, size_t n, uint32_t seed) {                                             \
    for (int t = 0; t < NTHREADS; ++t) {                                                                  \
      size_t lo, hi;                                                                                      \
      static_range(n, NTHREADS, t, &lo, &hi);                                                             \
      if (lo == hi) continue;                                                                             \
      mt_t m;                                                                                             \
      seed_mt(&m, seed + (uint32_t)t);                                                                    \
      for (size_t i = lo; i < hi; ++i) b[i] = CVT(uniform_real(&m, 0.0, (double)(MAXV)));                 \
    }                                                                                                     \
  }                                                                                                       \
  static void normal_##NAME(T* b, size_t n, uint32_t seed) {                                              \
    const double mean = (double)(MAXV) / 2.0, sd = mean / 3.0;                                            \
    for (int t = 0; t < NTHREADS; ++t) {                                                                  \
      size_t lo, hi;                                                                                      \
      static_range(n, NTHREADS, t, &lo, &hi);                                                             \
      if (lo == hi) continue;                                                                             \
      mt_t m;                                                                                             \
      seed_mt(&m, seed + (uint32_t)t);                                                                    \
      normal_t st = {0.0, 0};                                                                             \
      for (size_t i = lo; i < hi; ++i) b[i] = CVT(fabs(normal_draw(&st, &m, mean, sd)));                  \
    }                                                                                                     \
  }                                                                                                       \
  static void stag_bucket_##NAME(mt_t* m, double* lo, double* hi) {                                       \
    const T upper = CVT(uniform_real(m, 0.0, (double)(MAXV)));                                            \
    const T lower = ((double)NBUCKETS > (double)upper) ? (T)0 : WSUB(upper, (T)NBUCKETS);                 \
    *lo = (double)lower;                                                                                  \
    *hi = (double)upper;                                                                                  \
  }                                                                                                       \
  static void staggered_##NAME(T* b, size_t n, uint32_t seed) {                                           \
    if (n < NBUCKETS) {                                                                                   \
      uniform_##NAME(b, n, seed);                                                                         \
      return;                                                                                             \
    }                                                                                                     \
    const size_t bs = n / NBUCKETS, iters = n / bs;                                                       \
    for (int t = 0; t < NTHREADS; ++t) {                                                                  \
      size_t lo, hi;                                                                                      \
      static_range(iters, NTHREADS, t, &lo, &hi);                                                         \
      if (lo == hi) continue;                                                                             \
      mt_t m;                                                                                             \
      seed_mt(&m, seed + (uint32_t)t);                                                                    \
      for (size_t i = lo; i < hi; ++i) {                                                                  \
        double a, c;                                                                                      \
        stag_bucket_##NAME(&m, &a, &c);                                                                   \
        for (size_t j = i * bs; j < (i + 1) * bs; ++j) b[j] = CVT(uniform_real(&m, a, c));                \
      }                                                                                                   \
    }                                                                                                     \
    mt_t m;                                                                                               \
    seed_mt(&m, seed + (uint32_t)NTHREADS);                                                               \
    double a, c;                                                                                          \
    stag_bucket_##NAME(&m, &a, &c);                                                                       \
    for (size_t i = n - (n % bs); i < n; ++i) b[i] = CVT(uniform_real(&m, a, c));                         \
  }                                                                                                       \
  static void sorted_##NAME(T* b, size_t n, uint32_t seed) {                                              \
    uniform_##NAME(b, n, seed);                                                                           \
    qsort(b, n, sizeof(T), CMP);                                                                          \
  }                                                                                                       \
  static void reverse_##NAME(T* b, size_t n, uint32_t seed) {                                             \
    sorted_##NAME(b, n, seed);                                                                            \
    for (size_t i = 0; i < n / 2; ++i) {                                                                  \
      T x = b[i];                                                                                         \
      b[i] = b[n - 1 - i];                                                                                \
      b[n - 1 - i] = x;                                                                                   \
    }                                                                                                     \
  }                                                                                                       \
  static void nearly_##NAME(T* b, size_t n, uint32_t seed) {                                              \
    sorted_##NAME(b, n, seed);                                                                            \
    if (n < 2) return;                                                                                    \
    const size_t iters = n - 1;                                                                           \
    T* snap = (T*)malloc(n * sizeof(T));                                                                  \
    memcpy(snap, b, n * sizeof(T));                                                                       \
    for (int t = 0; t < NTHREADS; ++t) {                                                                  \
      size_t lo, hi;                                                                                      \
      static_range(iters, NTHREADS, t, &lo, &hi);                                                         \
      if (lo == hi) continue;                                                                             \
      mt_t m;                                                                                             \
      seed_mt(&m, seed + (uint32_t)t);                                                                    \
      for (size_t i = lo; i < hi; ++i) {                                                                  \
        double sd = (double)WSUB(snap[i + 1], snap[i]);                                                   \
        if (sd < DBL_MAX / 2.0) sd *= 2.0;                                                                \
        normal_t st = {0.0, 0};                                                                           \
        const T diff = CVT(fabs(normal_draw(&st, &m, 0.0, sd)));                                          \
        if (b[i] > WSUB((T)(MAXV), diff))                                                                 \
          b[i] = (T)(MAXV);                                                                               \
        else                                                                                              \
          b[i] = (T)(b[i] + diff);                                                                        \
      }                                                                                                   \
    }                                                                                                     \
    free(snap);                                                                                           \
  }                                                                                                       \
  static void bucket_##NAME(T* b, size_t n, uint32_t seed) {                                              \
    const size_t bs = n / NBUCKETS;                                                                       \
    const size_t tail = bs ? n - n % bs : 0;                                                              \
    for (int t = 0; t < NTHREADS; ++t) {                                                                  \
      size_t blo, bhi, tlo, thi;                                                                          \
      static_range(NBUCKETS, NTHREADS, t, &blo, &bhi);                                                    \
      static_range(n - tail, NTHREADS, t, &tlo, &thi);                                                    \
      if (blo == bhi && tlo == thi) continue;                                                             \
      mt_t m;                                                                                             \
      seed_mt(&m, seed + (uint32_t)t);                                                                    \
      for (size_t i = blo; i < bhi; ++i)                                                                  \
        for (size_t j = i * bs; j < (i + 1) * bs; ++j) b[j] = CVT(uniform_real(&m, 0.0, (double)(MAXV))); \
      for (size_t i = tail + tlo; i < tail + thi; ++i) b[i] = CVT(uniform_real(&m, 0.0, (double)(MAXV))); \
    }                                                                                                     \
    for (size_t i = NBUCKETS * bs; i < tail; ++i) b[i] = 0;                                               \
    if (bs == 0) {                                                                                        \
      qsort(b, n, sizeof(T), CMP);                                                                        \
      return;                                                                                             \
    }                                                                                                     \
    for (size_t i = 0; i <= NBUCKETS; ++i) {                                                              \
      const size_t lo = i * bs, hi = ((i + 1) * bs < n) ? (i + 1) * bs : n;                               \
      if (lo < hi) qsort(b + lo, hi - lo, sizeof(T), CMP);                                                \
    }                                                                                                     \
  }                                                                                                       \
  static void nclamp_##NAME(T* b, size_t n, uint32_t seed) {                                              \
    const uint64_t base = stream_base(seed, 1);                                                           \
    double mean, sd, lo, hi;                                                                              \
    if (ISFLOAT) {                                                                                        \
      mean = 0x1.0p31; sd = 0x1.0p28; lo = -INFINITY; hi = INFINITY;                                      \
    } else if (ISSIGNED) {                                                                                \
      mean = 0.0; sd = ldexp(1.0, WIDTH - 4); lo = (double)(MINV); hi = (double)(MAXV);                   \
    } else {                                                                                              \
      mean = ldexp(1.0, WIDTH - 1); sd = ldexp(1.0, WIDTH - 4); lo = 0.0; hi = (double)(MAXV);            \
    }                                                                                                     \
    for (size_t i = 0; i < n; ++i) {                                                                      \
      const double u1 = 1.0 - unit53(splitmix64(base + 2 * i));                                           \
      const double u2 = unit53(splitmix64(base + 2 * i + 1));                                             \
      const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);                                \
      const double x = mean + sd * z;                                                                     \
      if (ISFLOAT) b[i] = (T)x;                                                                           \
      else if (x <= lo) b[i] = (T)(MINV);                                                                 \
      else if (x >= hi) b[i] = (T)(MAXV);                                                                 \
      else b[i] = (T)x;                                                                                   \
    }                                                                                                     \
  }                                                                                                       \
  static void zipf_##NAME(T* b, size_t n, uint32_t seed, double s) {                                      \
    const size_t K = (size_t)1 << 20;                                                                     \
    double* cdf = (double*)malloc(K * sizeof(double));                                                    \
    double acc = 0.0;                                                                                     \
    for (size_t k = 0; k < K; ++k) {                                                                      \
      acc += pow((double)(k + 1), -s);                                                                    \
      cdf[k] = acc;                                                                                       \
    }                                                                                                     \
    const uint64_t base = stream_base(seed, 2), vbase = stream_base(seed, 3);                             \
    for (size_t i = 0; i < n; ++i) {                                                                      \
      const double u = unit53(splitmix64(base + i)) * acc;                                                \
      size_t lo = 0, hi = K; /* first k with cdf[k] > u */                                                \
      while (lo < hi) {                                                                                   \
        const size_t mid = (lo + hi) / 2;                                                                 \
        if (cdf[mid] > u) hi = mid; else lo = mid + 1;                                                    \
      }                                                                                                   \
      const uint64_t h = splitmix64(vbase + (lo < K ? lo : K - 1));                                       \
      if (ISFLOAT) b[i] = (T)(uint32_t)h;                                                                 \
      else b[i] = (T)h;                                                                                   \
    }                                                                                                     \
    free(cdf);                                                                                            \
  }                                                                                                       \
  static void swaps_##NAME(T* b, size_t n, uint32_t seed) {                                               \
    sorted_##NAME(b, n, seed);                                                                            \
    if (n < 2) return;                                                                                    \
    const uint64_t base = stream_base(seed, 4);                                                           \
    for (size_t j = 0; j < n / 100; ++j) {                                                                \
      const size_t x = (size_t)(splitmix64(base + 2 * j) % n), y = (size_t)(splitmix64(base + 2 * j + 1) % n); \
      T v = b[x];                                                                                         \
      b[x] = b[y];                                                                                        \
      b[y] = v;                                                                                           \
    }                                                                                                     \
  }                                                                                                       \
  static int gen_##NAME(const char* dist, uint32_t seed, T* b, size_t n) {                                \
    if (!strcmp(dist, "uniform")) uniform_##NAME(b, n, seed);                                             \
    else if (!strcmp(dist, "normal")) normal_##NAME(b, n, seed);                                          \
    else if (!strcmp(dist, "zero")) memset(b, 0, n * sizeof(T));                                          \
    else if (!strcmp(dist, "staggered")) staggered_##NAME(b, n, seed);                                    \
    else if (!strcmp(dist, "sorted")) sorted_##NAME(b, n, seed);                                          \
    else if (!strcmp(dist, "reverse-sorted")) reverse_##NAME(b, n, seed);                                 \
    else if (!strcmp(dist, "nearly-sorted")) nearly_##NAME(b, n, seed);                                   \
    else if (!strcmp(dist, "bucket-sorted")) bucket_##NAME(b, n, seed);                                   \
    else if (!strcmp(dist, "normal-clamped")) nclamp_##NAME(b, n, seed);                                  \
    else if (!strcmp(dist, "zipf")) zipf_##NAME(b, n, seed, 1.0);                                         \
    else if (!strcmp(dist, "zipf-1.2")) zipf_##NAME(b, n, seed, 1.2);                                     \
    else if (!strcmp(dist, "nearly-sorted-swaps")) swaps_##NAME(b, n, seed);                              \
    else return -1;                                                                                       \
    return 0;                                                                                             \
  }

DEFINE_GEN(i32, int32_t, cvt_i32, cmp_i32, INT32_MAX, INT32_MIN, wrapsub_int32_t, 0, 1, 32)
DEFINE_GEN(i64, int64_t, cvt_i64, cmp_i64, INT64_MAX, INT64_MIN, wrapsub_int64_t, 0, 1, 64)
DEFINE_GEN(u32, uint32_t, cvt_u32, cmp_u32, UINT32_MAX, 0, wrapsub_uint32_t, 0, 0, 32)
DEFINE_GEN(u64, uint64_t, cvt_u64, cmp_u64, UINT64_MAX, 0, wrapsub_uint64_t, 0, 0, 64)
DEFINE_GEN(f32, float, cvt_f32, cmp_f32, FLT_MAX, -FLT_MAX, wrapsub_float, 1, 1, 32)
DEFINE_GEN(f64, double, cvt_f64, cmp_f64, DBL_MAX, -DBL_MAX, wrapsub_double, 1, 1, 64)

/* float-specials: uniform, then 1% (at least 16) seeded positions replaced
 * by +-0, +-inf, NaNs with payloads, denormals, smallest normals. */
static void specials32(uint32_t* b, size_t n, uint32_t seed) {
  const uint64_t base = stream_base(seed, 5);
  size_t count = n / 100 > 16 ? n / 100 : 16;
  if (count > n) count = n;
  for (size_t j = 0; j < count; ++j) {
    const size_t pos = (size_t)(splitmix64(base + 2 * j) % n);
    const uint64_t r = splitmix64(base + 2 * j + 1);
    const uint32_t sign = (r & 1) ? 0x80000000u : 0u, payload = (uint32_t)(r >> 8) & 0x7FFFFFu;
    uint32_t v;
    switch ((r >> 1) % 6) {
      case 0: v = sign; break;
      case 1: v = sign | 0x7F800000u; break;
      case 2: v = sign | 0x7F800000u | 0x400000u | (payload >> 1); break;
      case 3: v = sign | 0x7F800000u | (payload ? payload : 1u); break;
      case 4: v = sign | (payload ? payload : 1u); break;
      default: v = sign | 0x800000u | payload; break;
    }
    b[pos] = v;
  }
}
static void specials64(uint64_t* b, size_t n, uint32_t seed) {
  const uint64_t base = stream_base(seed, 5);
  size_t count = n / 100 > 16 ? n / 100 : 16;
  if (count > n) count = n;
  const uint64_t E = 0x7FF0000000000000ull, M = (1ull << 52) - 1, S = 1ull << 63;
  for (size_t j = 0; j < count; ++j) {
    const size_t pos = (size_t)(splitmix64(base + 2 * j) % n);
    const uint64_t r = splitmix64(base + 2 * j + 1);
    const uint64_t sign = (r & 1) ? S : 0, payload = (r >> 8) & M;
    uint64_t v;
    switch ((r >> 1) % 6) {
      case 0: v = sign; break;
      case 1: v = sign | E; break;
      case 2: v = sign | E | (1ull << 51) | (payload >> 1); break;
      case 3: v = sign | E | (payload ? payload : 1); break;
      case 4: v = sign | (payload ? payload : 1); break;
      default: v = sign | (1ull << 52) | payload; break;
    }
    b[pos] = v;
  }
}

/* dtype: "int"/"i32", "long"/"i64", "float"/"f32", "double"/"f64", "u32", "u64".
 * Returns 0, or -1 for an unknown dtype/distribution. */
int oracle_generate(const char* dtype, const char* dist, uint32_t seed, void* out, size_t n, int reserved) {
  (void)reserved;
  const int specials = !strcmp(dist, "float-specials");
  const char* d = specials ? "uniform" : dist;
  if (!strcmp(dtype, "int") || !strcmp(dtype, "i32")) return specials ? -1 : gen_i32(d, seed, (int32_t*)out, n);
  if (!strcmp(dtype, "long") || !strcmp(dtype, "i64")) return specials ? -1 : gen_i64(d, seed, (int64_t*)out, n);
  if (!strcmp(dtype, "u32")) return specials ? -1 : gen_u32(d, seed, (uint32_t*)out, n);
  if (!strcmp(dtype, "u64")) return specials ? -1 : gen_u64(d, seed, (uint64_t*)out, n);
  if (!strcmp(dtype, "float") || !strcmp(dtype, "f32")) {
    const int r = gen_f32(d, seed, (float*)out, n);
    if (r == 0 && specials) specials32((uint32_t*)out, n, seed);
    return r;
  }
  if (!strcmp(dtype, "double") || !strcmp(dtype, "f64")) {
    const int r = gen_f64(d, seed, (double*)out, n);
    if (r == 0 && specials) specials64((uint64_t*)out, n, seed);
    return r;
  }
  return -1;
}

/* Positions whose nearly-sorted value the reference computes racily: the
 * last iteration of each of the 128 static blocks of n-1 iterations.
 * Writes up to 128 indices to out, returns how many. */
size_t oracle_nearly_sorted_racy_positions(size_t n, size_t* out) {
  size_t c = 0;
  if (n < 2) return 0;
  for (int t = 0; t < NTHREADS; ++t) {
    size_t lo, hi;
    static_range(n - 1, NTHREADS, t, &lo, &hi);
    if (hi > lo && hi < n - 1) out[c++] = hi - 1;
  }
  return c;
}

/* Known-answer hook: the k-th output (1-based) of mt19937 seeded with the
 * standard integer seeding (C++ [rand.eng.mers]: x0 = s,
 * xi = 1812433253 * (x(i-1) ^ (x(i-1) >> 30)) + i).  The standard fixes the
 * 10000th output of seed 5489 at 4123659995 ([rand.predef]). */
uint32_t oracle_mt19937_kat(uint32_t seed, int k) {
  mt_t m;
  m.x[0] = seed;
  for (int i = 1; i < 624; ++i) m.x[i] = 1812433253u * (m.x[i - 1] ^ (m.x[i - 1] >> 30)) + (uint32_t)i;
  m.p = 624;
  uint32_t v = 0;
  for (int i = 0; i < k; ++i) v = mt_next(&m);
  return v;
}

/* Known-answer hook for seed_seq::generate: fills out[0, n) from v[0, s). */
void oracle_seed_seq_generate(const uint32_t* v, size_t s, uint32_t* out, size_t n) { seed_seq_generate(v, s, out, n); }
