// Chinese remainder step of the residue split GEMMs (crt_bf16.hip with BF16
// residues, crt_i8.hip with INT8 residues), shared by the device kernels and
// a host build of the same code (tests/native), so the integer
// reconstruction is checked on the CPU against exact integers, including the
// ends of the balanced range (-P/2, P/2).
//
// For N moduli p_1..p_N (P = p_1...p_N, M_l = P / p_l, y_l = M_l^-1 mod p_l)
// and residues r_l of an integer C' with |C'| < P/2 and |C'| < 2^126:
//   C' = sum_l t_l M_l - q P,  t_l = r_l y_l mod p_l,  q = rint(sum_l t_l / p_l).
// M_l and P are held as six 24-bit limbs in fp64: with |t_l| < 2^9 every
// limb sum S_k = sum_l t_l mu_lk (< 2^37) and T_k = S_k - q pi_k is an exact
// fp64 integer; C' = sum_k T_k 2^(24k) is then assembled modulo 2^128
// (wrapping), exact because C' fits.
// sum_l t_l / p_l is within 10^-10 of C'/P + q_true, so q is exact unless
// |C'| is within 10^-10 P of P/2. That needs P < 2 k 2^(2 beta) (1 + 10^-10),
// which only an uncapped beta can give. When P < 2^127 (BF16 moduli: N <= 14;
// INT8 moduli: every N), C' +- P still fits and one compare against
// floor(P/2) fixes it. Otherwise (BF16, N >= 15) beta is capped, |C'| <=
// 2^125 < 2^-8 P and q is always exact.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define MOMMS_HD __host__ __device__
#else
#define MOMMS_HD
#endif

namespace momms {

constexpr int kCrtMaxModuli = 16;
// BF16 residues: the 16 largest pairwise-coprime odd integers <= 511
// (balanced residues |r| <= 255, integers bf16 holds exactly).
constexpr int kCrtModuli[kCrtMaxModuli] = {511, 509, 507, 505, 503, 499, 493, 491,
                                           487, 479, 473, 467, 463, 461, 457, 449};
// INT8 residues: the 16 largest pairwise-coprime odd integers <= 255
// (balanced residues |r| <= 127, int8 values).
constexpr int kCrtI8Moduli[kCrtMaxModuli] = {255, 253, 251, 247, 241, 239, 233, 229,
                                             227, 223, 217, 211, 199, 197, 193, 191};

constexpr int mod_inverse(int a, int m) {  // extended Euclid, gcd(a, m) = 1
  int t = 0, nt = 1, r = m, nr = a % m;
  while (nr != 0) {
    const int q = r / nr;
    const int t2 = t - q * nt;
    t = nt;
    nt = t2;
    const int r2 = r - q * nr;
    r = nr;
    nr = r2;
  }
  return t < 0 ? t + m : t;
}

constexpr int kCrtLimbs = 6;  // 24-bit limbs: 144 bits hold M_l and P for N <= 16

struct CrtRecon {
  double mu[kCrtMaxModuli][kCrtLimbs];                 // 24-bit limbs of M_l
  double pi[kCrtLimbs];                                // 24-bit limbs of P
  uint64_t p_lo, p_hi, h_lo, h_hi;                     // P and floor(P/2) mod 2^128
  float y[kCrtMaxModuli];                              // M_l^-1 mod p_l
  float pf[kCrtMaxModuli], invpf[kCrtMaxModuli];      // p_l, fp32(1/p_l)
  double invpd[kCrtMaxModuli];                         // fp64(1/p_l)
  int fix;                                             // P < 2^127: the +-P fix
  int limbs;                                           // 24-bit limbs that P (and so every M_l) needs
};

// Small unsigned big integers (32-bit limbs) for the constants.
struct CrtBig {
  uint32_t w[6] = {0, 0, 0, 0, 0, 0};
};
constexpr void crt_big_mul(CrtBig& a, uint32_t b) {
  uint64_t carry = 0;
  for (int i = 0; i < 6; ++i) {
    const uint64_t t = (uint64_t)a.w[i] * b + carry;
    a.w[i] = (uint32_t)t;
    carry = t >> 32;
  }
}
constexpr uint32_t crt_big_div(const CrtBig& a, uint32_t b, CrtBig& q) {  // returns a mod b
  uint64_t rem = 0;
  for (int i = 5; i >= 0; --i) {
    const uint64_t cur = (rem << 32) | a.w[i];
    q.w[i] = (uint32_t)(cur / b);
    rem = cur % b;
  }
  return (uint32_t)rem;
}

constexpr double crt_big_limb24(const CrtBig& a, int k) {  // bits 24k .. 24k+23
  const int bit = 24 * k, w = bit / 32, sh = bit % 32;
  uint64_t v = a.w[w] >> sh;
  if (sh > 8 && w + 1 < 6) v |= (uint64_t)a.w[w + 1] << (32 - sh);
  return (double)(v & 0xFFFFFF);
}

constexpr uint64_t crt_big_lo(const CrtBig& a) { return (uint64_t)a.w[0] | ((uint64_t)a.w[1] << 32); }
constexpr uint64_t crt_big_hi(const CrtBig& a) { return (uint64_t)a.w[2] | ((uint64_t)a.w[3] << 32); }

constexpr int crt_big_bitlen(const CrtBig& a) {
  for (int i = 5; i >= 0; --i)
    if (a.w[i]) {
      int b = 0;
      while (b < 32 && (a.w[i] >> b)) ++b;
      return 32 * i + b;
    }
  return 0;
}

// The constants for the first N moduli of a table (kCrtModuli or
// kCrtI8Moduli); constexpr, so the kernel for a given N holds them as
// instruction literals (a 24-bit limb as a double has a zero low word: one
// 32-bit literal).
constexpr CrtRecon crt_recon_constants(int N, const int* mods = kCrtModuli) {
  CrtRecon K{};
  CrtBig P;
  P.w[0] = 1;
  for (int l = 0; l < N; ++l) crt_big_mul(P, (uint32_t)mods[l]);
  for (int l = 0; l < N; ++l) {
    CrtBig M, unused;
    (void)crt_big_div(P, (uint32_t)mods[l], M);
    const uint32_t mr = crt_big_div(M, (uint32_t)mods[l], unused);
    for (int k = 0; k < kCrtLimbs; ++k) K.mu[l][k] = crt_big_limb24(M, k);
    K.y[l] = (float)mod_inverse((int)mr, mods[l]);
    K.pf[l] = (float)mods[l];
    K.invpf[l] = 1.0f / (float)mods[l];
    K.invpd[l] = 1.0 / (double)mods[l];
  }
  CrtBig H;
  (void)crt_big_div(P, 2, H);
  for (int k = 0; k < kCrtLimbs; ++k) K.pi[k] = crt_big_limb24(P, k);
  K.p_lo = crt_big_lo(P);
  K.p_hi = crt_big_hi(P);
  K.h_lo = crt_big_lo(H);
  K.h_hi = crt_big_hi(H);
  // C' +- P must fit in 128 bits for the compare against floor(P/2): P < 2^127
  // (BF16 moduli: N <= 14; above, beta is capped and q is always exact;
  // INT8 moduli: every N <= 16)
  K.fix = crt_big_bitlen(P) <= 127 ? 1 : 0;
  K.limbs = kCrtLimbs;  // the higher limbs of P are 0: fewer suffice, and M_l < P
  while (K.limbs > 1 && K.pi[K.limbs - 1] == 0.0) --K.limbs;
  return K;
}

MOMMS_HD inline unsigned __int128 crt_u128(uint64_t lo, uint64_t hi) {
  return ((unsigned __int128)hi << 64) | lo;
}

// C' from its balanced residues r[0..N-1] (integers in fp32).
// Running sums of one entry: f = sum_l t_l / p_l and the limb sums S_k.
struct CrtSums {
  double f, S[kCrtLimbs];
};

MOMMS_HD inline void crt_sums_init(CrtSums& a) {
  a.f = 0.0;
#pragma unroll
  for (int k = 0; k < kCrtLimbs; ++k) a.S[k] = 0.0;
}

// Adds the balanced residue r (an integer in fp32) modulo p_l.
MOMMS_HD inline void crt_sums_add(CrtSums& a, float r, int l, const CrtRecon& K) {
  // t = r y mod p: |r y| < 2^17 is exact in fp32; |t| < 0.51 p < 2^9
  float t = r * K.y[l];
  t = __builtin_fmaf(-__builtin_rintf(t * K.invpf[l]), K.pf[l], t);
  const double td = (double)t;
  a.f = __builtin_fma(td, K.invpd[l], a.f);
#pragma unroll
  for (int k = 0; k < kCrtLimbs; ++k)
    if (k < K.limbs) a.S[k] = __builtin_fma(td, K.mu[l][k], a.S[k]);
}

// C' from the sums over all N residues.
MOMMS_HD inline __int128 crt_sums_value(const CrtSums& a, const CrtRecon& K) {
  const double q = __builtin_rint(a.f);  // |q| <= 9
  unsigned __int128 x = 0;
#pragma unroll
  for (int k = 0; k < kCrtLimbs; ++k) {
    if (k >= K.limbs) break;  // limbs of P and of every M_l above are 0
    const double T = __builtin_fma(-q, K.pi[k], a.S[k]);  // exact integer, |T| < 2^38
    x += (unsigned __int128)(__int128)(int64_t)T << (24 * k);
  }
  __int128 c = (__int128)x;
  if (K.fix) {
    const __int128 P = (__int128)crt_u128(K.p_lo, K.p_hi), H = (__int128)crt_u128(K.h_lo, K.h_hi);
    if (c > H) c -= P;
    else if (c < -H) c += P;
  }
  return c;
}

// C' from its balanced residues r[0..N-1] (integers in fp32) modulo the
// moduli K was built for.
inline __int128 crt_combine(int N, const float* r, const CrtRecon& K) {
  CrtSums a;
  crt_sums_init(a);
  for (int l = 0; l < N; ++l) crt_sums_add(a, r[l], l, K);
  return crt_sums_value(a, K);
}

// Correctly rounded (to nearest, ties to even) conversion of a signed 128-bit
// integer: the top 64 bits with the shifted-out bits folded into a sticky
// bit, converted by the once-rounding u64 -> f64 conversion.
MOMMS_HD inline double crt_i128_to_f64(__int128 y) {
  const bool neg = y < 0;
  const unsigned __int128 u = neg ? (unsigned __int128)0 - (unsigned __int128)y : (unsigned __int128)y;
  const uint64_t hi = (uint64_t)(u >> 64), lo = (uint64_t)u;
  double d;
  if (hi == 0) {
    d = (double)lo;
  } else {
    const int s = 64 - __builtin_clzll(hi);  // 1..64 bits above the low word
    const uint64_t sticky_mask = s == 64 ? ~0ull : ((1ull << s) - 1);
    uint64_t top = (uint64_t)(u >> s);
    top |= (lo & sticky_mask) != 0 ? 1ull : 0ull;
    d = __builtin_ldexp((double)top, s);
  }
  return neg ? -d : d;
}

}  // namespace momms
