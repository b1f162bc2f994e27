// Host build of the residue split GEMM's Chinese remainder step
// (momms_amd/csrc/crt_recon.hpp, the code the device kernel runs).
// Reads lines "N r_1 ... r_N" from stdin and prints, per line, the hex of
// the correctly rounded double of the reconstructed integer and the integer
// itself in decimal, for comparison with exact Python integers. The residues
// are taken modulo the BF16 form's moduli, or the INT8 form's when the first
// argument is "i8".
#include <cstdio>

#include "crt_recon.hpp"

using namespace momms;

static void print_i128(__int128 v) {
  char buf[64];
  int n = 0;
  const bool neg = v < 0;
  unsigned __int128 u = neg ? (unsigned __int128)0 - (unsigned __int128)v : (unsigned __int128)v;
  do {
    buf[n++] = (char)('0' + (int)(u % 10));
    u /= 10;
  } while (u);
  if (neg) putchar('-');
  while (n) putchar(buf[--n]);
}

int main(int argc, char** argv) {
  const bool i8 = argc > 1 && argv[1][0] == 'i';
  CrtRecon K[kCrtMaxModuli + 1];
  for (int N = 1; N <= kCrtMaxModuli; ++N)
    K[N] = crt_recon_constants(N, i8 ? kCrtI8Moduli : kCrtModuli);
  int N;
  while (scanf("%d", &N) == 1) {
    if (N < 1 || N > kCrtMaxModuli) return 2;
    float r[kCrtMaxModuli];
    for (int l = 0; l < N; ++l) {
      int x;
      if (scanf("%d", &x) != 1) return 2;
      r[l] = (float)x;
    }
    const __int128 c = crt_combine(N, r, K[N]);
    printf("%a ", crt_i128_to_f64(c));
    print_i128(c);
    putchar('\n');
  }
  return 0;
}
