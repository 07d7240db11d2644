// Host-side check of the engine's multi-word string helpers (sbd_amd/csrc/common.h): reads cases
// from stdin, one per line,
//   x_lo x_hi y_lo y_hi i a I J A B
// (two strings of two 64-bit words, orbital indices) and prints, per line,
//   popc first second count_range(x, i, a) sign_single(x, i, a) parity_ref(x, i, a)
//   sign_double(x, I, J, A, B) det_sign(x, y) exc_i exc_a exc_neg exc_pair
// with i < a.  When both strings fit one word the Str<1> results follow on the same line, so
// the test compares the two widths.  No GPU call is made.  Test program for
// tests/test_str_words.py.
#include "common.h"

#include <cstdio>

using namespace sbdx;

template <int W>
static void emit(const Str<W>& x, const Str<W>& y, int i, int a, int I, int J, int A, int B) {
  std::printf(" %d %d %d %d %d %d %d %d", popc(x), first(x), first(drop_first(x)), count_range(x, i, a),
              sign_single(x, i, a), parity_ref(x, i, a), sign_double(x, I, J, A, B), det_sign(x, y));
}

int main() {
  unsigned long long xl, xh, yl, yh;
  int i, a, I, J, A, B;
  while (std::scanf("%llu %llu %llu %llu %d %d %d %d %d %d", &xl, &xh, &yl, &yh, &i, &a, &I, &J, &A, &B) == 10) {
    const Str<2> x{{xl, xh}}, y{{yl, yh}};
    emit(x, y, i, a, I, J, A, B);
    const uint32_t p = exc_pack(i, a, (i + a) & 1);
    std::printf(" %d %d %d %u", exc_i(p), exc_a(p), exc_neg(p) ? 1 : 0, exc_pair(p, 128));
    if (xh == 0 && yh == 0 && a <= 64 && B < 64 && I < 64 && J < 64 && A < 64) {
      const Str<1> x1{{xl}}, y1{{yl}};
      emit(x1, y1, i, a, I, J, A, B);
    }
    std::printf("\n");
  }
  return 0;
}
