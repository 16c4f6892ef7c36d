/*
 * zkref.c - C restatement of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * The same algorithms as oracle/zkref (which restates the reference's
 * src/zkgpu/ntt.py, domain.py, curve.py and msm.py), on 4x64-bit Montgomery
 * arithmetic so that full-size inputs (2^20-point MSMs, 2^22-point NTTs)
 * can be checked in seconds. Every exported function takes and returns
 * plain (non-Montgomery) canonical values as little-endian uint64 limbs.
 * It is checked against the Python oracle by tests/test_oracle_c.py.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

typedef unsigned __int128 u128;
typedef struct { uint64_t v[4]; } fe;
typedef struct { uint64_t p[4]; uint64_t ninv; fe r2; fe one; } field;

static const uint64_t R_LIMBS[4] = {0x43e1f593f0000001ULL, 0x2833e84879b97091ULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL};
static const uint64_t P_LIMBS[4] = {0x3c208c16d87cfd47ULL, 0x97816a916871ca8dULL, 0xb85045b68181585dULL, 0x30644e72e131a029ULL};

static int geq(const uint64_t a[4], const uint64_t b[4]) {
  for (int i = 3; i >= 0; --i) { if (a[i] != b[i]) return a[i] > b[i]; }
  return 1;
}
static void sub4(uint64_t r[4], const uint64_t a[4], const uint64_t b[4]) {
  u128 br = 0;
  for (int i = 0; i < 4; ++i) { u128 d = (u128)a[i] - b[i] - br; r[i] = (uint64_t)d; br = (d >> 64) ? 1 : 0; }
}
static fe f_add(const field* F, fe a, fe b) {
  fe r; u128 c = 0;
  for (int i = 0; i < 4; ++i) { c += (u128)a.v[i] + b.v[i]; r.v[i] = (uint64_t)c; c >>= 64; }
  if (c || geq(r.v, F->p)) sub4(r.v, r.v, F->p);
  return r;
}
static fe f_sub(const field* F, fe a, fe b) {
  fe r; u128 br = 0;
  for (int i = 0; i < 4; ++i) { u128 d = (u128)a.v[i] - b.v[i] - br; r.v[i] = (uint64_t)d; br = (d >> 64) ? 1 : 0; }
  if (br) { u128 c = 0; for (int i = 0; i < 4; ++i) { c += (u128)r.v[i] + F->p[i]; r.v[i] = (uint64_t)c; c >>= 64; } }
  return r;
}
/* CIOS Montgomery product a*b*2^-256 mod p (p < 2^255). */
static fe f_mul(const field* F, fe a, fe b) {
  uint64_t t[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 4; ++i) {
    u128 c = 0;
    for (int j = 0; j < 4; ++j) { c += (u128)a.v[j] * b.v[i] + t[j]; t[j] = (uint64_t)c; c >>= 64; }
    c += t[4]; t[4] = (uint64_t)c; t[5] = (uint64_t)(c >> 64);
    uint64_t m = t[0] * F->ninv;
    c = (u128)m * F->p[0] + t[0]; c >>= 64;
    for (int j = 1; j < 4; ++j) { c += (u128)m * F->p[j] + t[j]; t[j - 1] = (uint64_t)c; c >>= 64; }
    c += t[4]; t[3] = (uint64_t)c; t[4] = t[5] + (uint64_t)(c >> 64);
  }
  fe r; memcpy(r.v, t, 32);
  if (t[4] || geq(r.v, F->p)) sub4(r.v, r.v, F->p);
  return r;
}
static int f_is_zero(fe a) { return (a.v[0] | a.v[1] | a.v[2] | a.v[3]) == 0; }
static fe f_from(const uint64_t* limbs) { fe r; memcpy(r.v, limbs, 32); return r; }
static fe f_to_mont(const field* F, fe a) { return f_mul(F, a, F->r2); }
static fe f_from_mont(const field* F, fe a) { fe one = {{1, 0, 0, 0}}; return f_mul(F, a, one); }
static fe f_pow(const field* F, fe base_m, const uint64_t e[4]) {
  fe acc = F->one;
  for (int i = 255; i >= 0; --i) {
    acc = f_mul(F, acc, acc);
    if ((e[i >> 6] >> (i & 63)) & 1) acc = f_mul(F, acc, base_m);
  }
  return acc;
}
static fe f_inv(const field* F, fe a_m) {
  uint64_t e[4]; uint64_t two[4] = {2, 0, 0, 0}; sub4(e, F->p, two);
  return f_pow(F, a_m, e);
}
static fe f_small(const field* F, uint64_t x) { fe r = {{x, 0, 0, 0}}; return f_to_mont(F, r); }

static void field_init(field* F, const uint64_t p[4]) {
  memcpy(F->p, p, 32);
  uint64_t inv = 1;  /* Newton iteration for p^-1 mod 2^64 */
  for (int i = 0; i < 7; ++i) inv *= 2 - p[0] * inv;
  F->ninv = (uint64_t)0 - inv;
  /* R mod p by doubling 1 two hundred fifty-six times, then R^2 = (R mod p) doubled 256 more times */
  fe x = {{1, 0, 0, 0}};
  for (int i = 0; i < 512; ++i) {
    if (i == 256) F->one = x;
    x = f_add(F, x, x);
  }
  F->r2 = x;
}
static field FR, FQ;
static pthread_once_t once = PTHREAD_ONCE_INIT;
static void init_fields(void) { field_init(&FR, R_LIMBS); field_init(&FQ, P_LIMBS); }
static void ensure(void) { pthread_once(&once, init_fields); }

/* ---------------- NTT (ntt.py:149-203) ---------------- */
static int reverse_bits(uint64_t i, int bits) { uint64_t r = 0; for (int b = 0; b < bits; ++b) { r = (r << 1) | (i & 1); i >>= 1; } return (int)r; }

static void ntt_mont(const field* F, fe* x, int log_n, fe omega_m) {
  long n = 1L << log_n;
  for (long i = 0; i < n; ++i) { long j = reverse_bits(i, log_n); if (j > i) { fe t = x[i]; x[i] = x[j]; x[j] = t; } }
  for (int s = 1; s <= log_n; ++s) {
    long len = 1L << s, half = len >> 1;
    uint64_t e[4] = {(uint64_t)(n / len), 0, 0, 0};
    fe wl = f_pow(F, omega_m, e);
    for (long st = 0; st < n; st += len) {
      fe w = F->one;
      for (long j = 0; j < half; ++j) {
        fe u = x[st + j], v = f_mul(F, x[st + j + half], w);
        x[st + j] = f_add(F, u, v); x[st + j + half] = f_sub(F, u, v);
        w = f_mul(F, w, wl);
      }
    }
  }
}
/* omega_plain: the primitive 2^log_n-th root (plain form); inverse transforms pass omega^-1 and n_inv. */
int zr_ntt(const uint64_t* in, uint64_t* out, int log_n, const uint64_t* omega_plain, const uint64_t* scale_plain) {
  ensure();
  long n = 1L << log_n;
  fe* x = (fe*)malloc(sizeof(fe) * n);
  if (!x) return -1;
  for (long i = 0; i < n; ++i) x[i] = f_to_mont(&FR, f_from(in + 4 * i));
  if (log_n > 0) ntt_mont(&FR, x, log_n, f_to_mont(&FR, f_from(omega_plain)));
  fe sc = scale_plain ? f_to_mont(&FR, f_from(scale_plain)) : FR.one;
  for (long i = 0; i < n; ++i) { fe v = f_from_mont(&FR, f_mul(&FR, x[i], sc)); memcpy(out + 4 * i, v.v, 32); }
  free(x);
  return 0;
}
/* out[j] = in[j] * g^j (domain.py scaling); g plain. */
int zr_scale_by_powers(const uint64_t* in, uint64_t* out, long n, const uint64_t* g_plain) {
  ensure();
  fe g = f_to_mont(&FR, f_from(g_plain)), p = FR.one;
  for (long i = 0; i < n; ++i) {
    fe v = f_from_mont(&FR, f_mul(&FR, f_to_mont(&FR, f_from(in + 4 * i)), p));
    memcpy(out + 4 * i, v.v, 32); p = f_mul(&FR, p, g);
  }
  return 0;
}
/* out[j] = (a[j]*b[j] - c[j]) * z (plain). */
int zr_pointwise_quotient(const uint64_t* a, const uint64_t* b, const uint64_t* c, const uint64_t* z_plain, uint64_t* out, long n) {
  ensure();
  fe z = f_to_mont(&FR, f_from(z_plain));
  for (long i = 0; i < n; ++i) {
    fe x = f_to_mont(&FR, f_from(a + 4 * i)), y = f_to_mont(&FR, f_from(b + 4 * i)), w = f_to_mont(&FR, f_from(c + 4 * i));
    fe v = f_from_mont(&FR, f_mul(&FR, f_sub(&FR, f_mul(&FR, x, y), w), z));
    memcpy(out + 4 * i, v.v, 32);
  }
  return 0;
}

/* ---------------- G1 (curve.py:236-324), Montgomery Fq coordinates ---------------- */
typedef struct { fe x, y, z; } jac;
static jac jac_identity(void) { jac r; r.x = FQ.one; r.y = FQ.one; memset(&r.z, 0, 32); return r; }
static jac jdbl(jac a) {
  const field* F = &FQ;
  if (f_is_zero(a.z) || f_is_zero(a.y)) return jac_identity();
  fe A = f_mul(F, a.x, a.x), B = f_mul(F, a.y, a.y), C = f_mul(F, B, B);
  fe t = f_add(F, a.x, B); t = f_mul(F, t, t); t = f_sub(F, f_sub(F, t, A), C);
  fe D = f_add(F, t, t), E = f_add(F, f_add(F, A, A), A), Fv = f_mul(F, E, E);
  jac r;
  r.x = f_sub(F, Fv, f_add(F, D, D));
  fe c8 = f_add(F, C, C); c8 = f_add(F, c8, c8); c8 = f_add(F, c8, c8);
  r.y = f_sub(F, f_mul(F, E, f_sub(F, D, r.x)), c8);
  fe yz = f_mul(F, a.y, a.z); r.z = f_add(F, yz, yz);
  return r;
}
static jac jadd(jac a, jac b) {
  const field* F = &FQ;
  if (f_is_zero(a.z)) return b;
  if (f_is_zero(b.z)) return a;
  fe z1z1 = f_mul(F, a.z, a.z), z2z2 = f_mul(F, b.z, b.z);
  fe u1 = f_mul(F, a.x, z2z2), u2 = f_mul(F, b.x, z1z1);
  fe s1 = f_mul(F, f_mul(F, a.y, b.z), z2z2), s2 = f_mul(F, f_mul(F, b.y, a.z), z1z1);
  fe h = f_sub(F, u2, u1), rr = f_sub(F, s2, s1); rr = f_add(F, rr, rr);
  if (f_is_zero(h)) return f_is_zero(rr) ? jdbl(a) : jac_identity();
  fe h2 = f_add(F, h, h), I = f_mul(F, h2, h2), J = f_mul(F, h, I), V = f_mul(F, u1, I);
  jac r;
  r.x = f_sub(F, f_sub(F, f_mul(F, rr, rr), J), f_add(F, V, V));
  fe s1j = f_mul(F, s1, J);
  r.y = f_sub(F, f_mul(F, rr, f_sub(F, V, r.x)), f_add(F, s1j, s1j));
  fe zz = f_add(F, a.z, b.z); zz = f_mul(F, zz, zz);
  r.z = f_mul(F, f_sub(F, f_sub(F, zz, z1z1), z2z2), h);
  return r;
}
static jac jmadd(jac a, fe x2, fe y2) {
  const field* F = &FQ;
  if (f_is_zero(a.z)) { jac r; r.x = x2; r.y = y2; r.z = F->one; return r; }
  fe z1z1 = f_mul(F, a.z, a.z), u2 = f_mul(F, x2, z1z1), s2 = f_mul(F, f_mul(F, y2, a.z), z1z1);
  fe h = f_sub(F, u2, a.x), rr = f_sub(F, s2, a.y); rr = f_add(F, rr, rr);
  if (f_is_zero(h)) return f_is_zero(rr) ? jdbl(a) : jac_identity();
  fe hh = f_mul(F, h, h), I = f_add(F, hh, hh); I = f_add(F, I, I);
  fe J = f_mul(F, h, I), V = f_mul(F, a.x, I);
  jac r;
  r.x = f_sub(F, f_sub(F, f_mul(F, rr, rr), J), f_add(F, V, V));
  fe y1j = f_mul(F, a.y, J);
  r.y = f_sub(F, f_mul(F, rr, f_sub(F, V, r.x)), f_add(F, y1j, y1j));
  fe zh = f_add(F, a.z, h); zh = f_mul(F, zh, zh);
  r.z = f_sub(F, f_sub(F, zh, z1z1), hh);
  return r;
}
static void jac_out(jac a, uint64_t* out) {
  fe x = f_from_mont(&FQ, a.x), y = f_from_mont(&FQ, a.y), z = f_from_mont(&FQ, a.z);
  memcpy(out, x.v, 32); memcpy(out + 4, y.v, 32); memcpy(out + 8, z.v, 32);
}
static jac jac_in(const uint64_t* in) {
  jac r; r.x = f_to_mont(&FQ, f_from(in)); r.y = f_to_mont(&FQ, f_from(in + 4)); r.z = f_to_mont(&FQ, f_from(in + 8)); return r;
}
/* Element-wise Jacobian ops on plain (n,3,4) rows; b_step 0 broadcasts. op: 0 add, 1 madd (b affine (n,2,4)), 2 double. */
int zr_g1_op(int op, const uint64_t* a, long a_step, const uint64_t* b, long b_step, uint64_t* out, long n) {
  ensure();
  for (long i = 0; i < n; ++i) {
    jac x = jac_in(a + i * a_step), r;
    if (op == 0) r = jadd(x, jac_in(b + i * b_step));
    else if (op == 1) r = jmadd(x, f_to_mont(&FQ, f_from(b + i * b_step)), f_to_mont(&FQ, f_from(b + i * b_step + 4)));
    else r = jdbl(x);
    jac_out(r, out + 12 * i);
  }
  return 0;
}
/* Affine (x, y) of a Jacobian point in Montgomery form; returns 0 for the identity. */
static int to_affine(jac a, fe* x, fe* y) {
  if (f_is_zero(a.z)) return 0;
  fe zi = f_inv(&FQ, a.z), zi2 = f_mul(&FQ, zi, zi);
  *x = f_mul(&FQ, a.x, zi2); *y = f_mul(&FQ, f_mul(&FQ, a.y, zi2), zi);
  return 1;
}
static void affine_out(jac a, uint64_t* out8, int* finite) {
  fe x, y;
  *finite = to_affine(a, &x, &y);
  if (*finite) { x = f_from_mont(&FQ, x); y = f_from_mont(&FQ, y); memcpy(out8, x.v, 32); memcpy(out8 + 4, y.v, 32); }
  else memset(out8, 0, 64);
}

/* ---------------- Pippenger (msm.py:132-207) ---------------- */
static unsigned digit_of(const uint64_t* s, int bit, int w) {
  unsigned d = 0;
  for (int k = 0; k < w; ++k) { int b = bit + k; if (b < 256 && ((s[b >> 6] >> (b & 63)) & 1)) d |= 1u << k; }
  return d;
}
/* points: (n,2,4) plain affine (all finite); scalars (n,4) plain, already reduced mod r. */
int zr_msm_pippenger(const uint64_t* points, const uint64_t* scalars, long n, int w, uint64_t* out8, int* finite) {
  ensure();
  int windows = (254 + w - 1) / w;
  long nb = 1L << w;
  jac* buckets = (jac*)malloc(sizeof(jac) * nb);
  fe* xs = (fe*)malloc(sizeof(fe) * n); fe* ys = (fe*)malloc(sizeof(fe) * n);
  jac* sums = (jac*)malloc(sizeof(jac) * windows);
  if (!buckets || !xs || !ys || !sums) return -1;
  for (long i = 0; i < n; ++i) { xs[i] = f_to_mont(&FQ, f_from(points + 8 * i)); ys[i] = f_to_mont(&FQ, f_from(points + 8 * i + 4)); }
  for (int win = 0; win < windows; ++win) {
    for (long b = 0; b < nb; ++b) buckets[b] = jac_identity();
    for (long i = 0; i < n; ++i) { unsigned d = digit_of(scalars + 4 * i, win * w, w); if (d) buckets[d] = jmadd(buckets[d], xs[i], ys[i]); }
    jac running = jac_identity(), total = jac_identity();
    for (long b = nb - 1; b > 0; --b) { running = jadd(running, buckets[b]); total = jadd(total, running); }
    sums[win] = total;
  }
  jac acc = sums[windows - 1];
  for (int win = windows - 2; win >= 0; --win) { for (int k = 0; k < w; ++k) acc = jdbl(acc); acc = jadd(acc, sums[win]); }
  affine_out(acc, out8, finite);
  free(buckets); free(xs); free(ys); free(sums);
  return 0;
}

/* ---------------- Fixed-base multiples k*G (curve.py:306-324 semantics; result affine) ---------------- */
typedef struct { const uint64_t* k; uint64_t* out; long lo, hi; const fe* tx; const fe* ty; const int* tfin; } fb_job;
/* table[j][d] = d * 2^(8j) * G in affine Montgomery form, j < 32, d < 256 */
static fe* TX; static fe* TY; static int* TFIN;
static pthread_once_t table_once = PTHREAD_ONCE_INIT;
static void build_table(void) {
  ensure();
  TX = (fe*)malloc(sizeof(fe) * 32 * 256); TY = (fe*)malloc(sizeof(fe) * 32 * 256); TFIN = (int*)malloc(sizeof(int) * 32 * 256);
  jac base; base.x = f_small(&FQ, 1); base.y = f_small(&FQ, 2); base.z = FQ.one;
  for (int j = 0; j < 32; ++j) {
    jac acc = jac_identity();
    for (int d = 0; d < 256; ++d) {
      if (d) acc = jadd(acc, base);
      fe x, y; TFIN[j * 256 + d] = to_affine(acc, &x, &y); TX[j * 256 + d] = x; TY[j * 256 + d] = y;
    }
    for (int k = 0; k < 8; ++k) base = jdbl(base);
  }
}
static void* fb_worker(void* arg) {
  fb_job* jb = (fb_job*)arg;
  for (long i = jb->lo; i < jb->hi; ++i) {
    const uint64_t* k = jb->k + 4 * i;
    jac acc = jac_identity();
    for (int j = 0; j < 32; ++j) {
      int d = (int)((k[j >> 3] >> (8 * (j & 7))) & 0xff);
      if (d && TFIN[j * 256 + d]) acc = jmadd(acc, TX[j * 256 + d], TY[j * 256 + d]);
    }
    int fin; affine_out(acc, jb->out + 8 * i, &fin);
  }
  return NULL;
}
/* k: (n,4) plain scalars < 2^256; out: (n,2,4) plain affine (zeros for the identity). */
int zr_fixed_base_g1(const uint64_t* k, uint64_t* out, long n, int threads) {
  pthread_once(&table_once, build_table);
  if (threads < 1) threads = 1;
  if (threads > 64) threads = 64;
  pthread_t th[64]; fb_job jobs[64];
  for (int t = 0; t < threads; ++t) {
    jobs[t].k = k; jobs[t].out = out; jobs[t].lo = n * t / threads; jobs[t].hi = n * (t + 1) / threads;
    pthread_create(&th[t], NULL, fb_worker, &jobs[t]);
  }
  for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
  return 0;
}

/* ---------------- Closed form of an MSM over a basis P_i = k_i G ---------------- */
/* For P_i = k_i G, sum_i s_i P_i = (sum_i k_i s_i mod r) G (msm.py:112-207 returns that unique
 * group element). out = sum_{i<n} k_i s_i mod r (plain), for plain k_i, s_i < 2^256 (reduced mod r
 * first, as msm_gpu.py:433 reduces scalars). Montgomery products carry one factor 2^-256 per
 * term, so the sum is brought back by one product with 2^512 mod r. */
static fe reduce_r(const uint64_t* limbs) {
  fe v = f_from(limbs);
  while (geq(v.v, FR.p)) sub4(v.v, v.v, FR.p);
  return v;
}
int zr_dot_mod_r(const uint64_t* k, const uint64_t* s, long n, uint64_t* out) {
  ensure();
  fe acc = {{0, 0, 0, 0}};
  for (long i = 0; i < n; ++i) acc = f_add(&FR, acc, f_mul(&FR, reduce_r(k + 4 * i), reduce_r(s + 4 * i)));
  acc = f_mul(&FR, acc, FR.r2);
  memcpy(out, acc.v, 32);
  return 0;
}
