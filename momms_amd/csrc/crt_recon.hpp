// Chinese remainder step of the residue split GEMMs (crt_bf16.hip with BF16
// residues, crt_i8.hip with INT8 residues), shared by the device kernels and
// a host build of the same code (tests/native), so the integer
// reconstruction is checked on the CPU against exact integers, including the
// ends of the balanced range (-P/2, P/2).
//
// For N moduli p_1..p_N (P = p_1...p_N, M_l = P / p_l, y_l = M_l^-1 mod p_l)
// and balanced residues r_l (|r_l| <= 255) of an integer C' with |C'| < P/2
// and |C'| < 2^126:
//   X = sum_l r_l y_l M_l = C' + q P,  q = rint(X / P) = rint(sum_l r_l y_l / p_l),
// so C' = X - q P, and it is computed modulo 2^128 (exact, because C' fits in
// 128 bits signed): u_l = y_l M_l mod 2^128 and P mod 2^128 are held as four
// 32-bit limbs c_lk, pi_k in fp64. Every limb sum S_k = sum_l r_l c_lk
// (|r c| < 2^40, N <= 16: |S_k| < 2^44) and T_k = S_k - q pi_k (|q| <= 16 *
// 255 + 1 < 2^12: |T_k| < 2^45) is an exact fp64 integer, and C' = sum_k T_k
// 2^(32k) modulo 2^128. Per residue: one conversion, one fma for the
// quotient sum and four for the limbs (r y is never reduced modulo p).
// The quotient sum f = sum_l r_l fl(y_l / p_l) is within 10^-11 of X / P =
// q + C'/P, so q is exact unless |C'| is within 10^-11 P of P/2. That needs
// P < 2 k 2^(2 beta) (1 + 10^-11), which only an uncapped beta can give.
// When P < 2^127 (BF16 moduli: N <= 14; INT8 moduli: every N), C' +- P then
// has magnitude about P/2 and still fits, and one compare against floor(P/2)
// fixes it. Otherwise (BF16, N >= 15) beta is capped, |C'| <= 2^125 <
// 2^-8 P and q is always exact.
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

constexpr int kCrtLimbs = 4;  // 32-bit limbs: everything is computed modulo 2^128

struct CrtRecon {
  double c[kCrtMaxModuli][kCrtLimbs];  // 32-bit limbs of y_l M_l mod 2^128
  double pi[kCrtLimbs];                // 32-bit limbs of P mod 2^128
  double yp[kCrtMaxModuli];            // fp64(y_l / p_l)
  uint64_t p_lo, p_hi, h_lo, h_hi;     // P and floor(P/2) mod 2^128
  int fix;                             // P < 2^127: the +-P fix
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

constexpr double crt_big_limb32(const CrtBig& a, int k) { return (double)a.w[k]; }

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
// instruction literals.
constexpr CrtRecon crt_recon_constants(int N, const int* mods = kCrtModuli) {
  CrtRecon K{};
  CrtBig P;
  P.w[0] = 1;
  for (int l = 0; l < N; ++l) crt_big_mul(P, (uint32_t)mods[l]);
  for (int l = 0; l < N; ++l) {
    CrtBig M, unused;
    (void)crt_big_div(P, (uint32_t)mods[l], M);
    const uint32_t mr = crt_big_div(M, (uint32_t)mods[l], unused);
    const int y = mod_inverse((int)mr, mods[l]);
    crt_big_mul(M, (uint32_t)y);  // y_l M_l < 2^9 P < 2^153: within the 192 bits
    for (int k = 0; k < kCrtLimbs; ++k) K.c[l][k] = crt_big_limb32(M, k);
    K.yp[l] = (double)y / (double)mods[l];
  }
  CrtBig H;
  (void)crt_big_div(P, 2, H);
  for (int k = 0; k < kCrtLimbs; ++k) K.pi[k] = crt_big_limb32(P, k);
  K.p_lo = crt_big_lo(P);
  K.p_hi = crt_big_hi(P);
  K.h_lo = crt_big_lo(H);
  K.h_hi = crt_big_hi(H);
  // C' +- P must fit in 128 bits for the compare against floor(P/2): P < 2^127
  // (BF16 moduli: N <= 14; above, beta is capped and q is always exact;
  // INT8 moduli: every N <= 16)
  K.fix = crt_big_bitlen(P) <= 127 ? 1 : 0;
  return K;
}

MOMMS_HD inline unsigned __int128 crt_u128(uint64_t lo, uint64_t hi) {
  return ((unsigned __int128)hi << 64) | lo;
}

// Running sums of one entry: f = sum_l r_l y_l / p_l and the limb sums S_k.
struct CrtSums {
  double f, S[kCrtLimbs];
};

MOMMS_HD inline void crt_sums_init(CrtSums& a) {
  a.f = 0.0;
#pragma unroll
  for (int k = 0; k < kCrtLimbs; ++k) a.S[k] = 0.0;
}

// Adds the balanced residue r (an integer, |r| <= 255, as int or as an
// integer-valued float) modulo p_l.
template <typename R>
MOMMS_HD inline void crt_sums_add(CrtSums& a, R r, int l, const CrtRecon& K) {
  const double rd = (double)r;
  a.f = __builtin_fma(rd, K.yp[l], a.f);
#pragma unroll
  for (int k = 0; k < kCrtLimbs; ++k) a.S[k] = __builtin_fma(rd, K.c[l][k], a.S[k]);
}

// C' from the sums over all N residues.
MOMMS_HD inline __int128 crt_sums_value(const CrtSums& a, const CrtRecon& K) {
  const double q = __builtin_rint(a.f);  // |q| < 2^12
  unsigned __int128 x = 0;
#pragma unroll
  for (int k = 0; k < kCrtLimbs; ++k) {
    const double T = __builtin_fma(-q, K.pi[k], a.S[k]);  // exact integer, |T| < 2^45
    x += (unsigned __int128)(__int128)(int64_t)T << (32 * k);
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
