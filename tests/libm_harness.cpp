// CPU check of trinity_amd/csrc/trn_libm.hpp against the host C library.
// Built by tests/test_libm.py with g++ (no floating-point contraction):
// the same source the device compiles, evaluated on the host.
//
//   libm_harness eval <op> <file_in> <file_out>   op: exp log pow log1p_small expm1 cube
//       evaluates trn::libm on the raw doubles of file_in (pairs for pow)
//   libm_harness sweep <op> <n> <seed>
//       draws n inputs per range and prints "range mismatches n" lines,
//       comparing bit patterns with exp/log/pow of the C library
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../trinity_amd/csrc/trn_libm.hpp"

static uint64_t rng_state;
static double urand() {   // splitmix64 -> [0, 1)
  uint64_t z = (rng_state += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1p-53;
}
static bool same(double a, double b) {
  if (isnan(a) && isnan(b)) return true;
  uint64_t ua, ub;
  memcpy(&ua, &a, 8);
  memcpy(&ub, &b, 8);
  return ua == ub;
}

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const char* mode = argv[1];
  const char* op = argv[2];
  if (!strcmp(mode, "eval")) {
    FILE* fi = fopen(argv[3], "rb");
    if (!fi) return 3;
    std::vector<double> in;
    double v;
    while (fread(&v, 8, 1, fi) == 1) in.push_back(v);
    fclose(fi);
    std::vector<double> out;
    if (!strcmp(op, "pow")) {
      for (size_t i = 0; i + 1 < in.size(); i += 2) out.push_back(trn::libm::pow(in[i], in[i + 1]));
    } else {
      for (double x : in) {
        if (!strcmp(op, "exp")) out.push_back(trn::libm::exp(x));
        else if (!strcmp(op, "log")) out.push_back(trn::libm::log(x));
        else if (!strcmp(op, "log1p_small")) out.push_back(trn::libm::log1p_small(x));
        else if (!strcmp(op, "cube")) out.push_back(trn::libm::cube(x));
        else out.push_back(trn::libm::expm1(x));
      }
    }
    FILE* fo = fopen(argv[4], "wb");
    if (!fo) return 3;
    fwrite(out.data(), 8, out.size(), fo);
    fclose(fo);
    return 0;
  }
  const long n = atol(argv[3]);
  rng_state = strtoull(argv[4], nullptr, 10);
  if (!strcmp(op, "exp")) {
    // the path's range (logs of metric eigenvalues), the normal range, the
    // scaled ranges near overflow and into the subnormals, tiny arguments
    const double lo[5] = {-50, -708, 512, -745.2, -1e-9};
    const double hi[5] = {50, 708, 709.8, -512, 1e-9};
    for (int r = 0; r < 5; ++r) {
      long bad = 0;
      for (long i = 0; i < n; ++i) {
        const double x = lo[r] + (hi[r] - lo[r]) * urand();
        volatile double g = exp(x);
        if (!same(g, trn::libm::exp(x))) ++bad;
      }
      printf("exp%d %ld %ld\n", r, bad, n);
    }
  } else if (!strcmp(op, "pow")) {
    // L^p exponents -1/(2p+2) for p = 1, 2, 4, the p = 0 exponent 1/2,
    // the cube of the quality, general exponents, subnormal bases
    const double ys[5] = {-0.25, -1.0 / 6, -0.1, 0.5, 3.0};
    for (int r = 0; r < 7; ++r) {
      long bad = 0;
      for (long i = 0; i < n; ++i) {
        double x, y;
        if (r < 4) { x = exp((urand() * 2 - 1) * 100); y = ys[r]; }
        else if (r == 4) { x = urand() * 1.2; y = ys[4]; }
        else if (r == 5) { x = exp((urand() * 2 - 1) * 20); y = (urand() * 2 - 1) * 30; }
        else { x = urand() * 0x1p-1022; y = (urand() * 2 - 1) * 0.9; }
        volatile double g = pow(x, y);
        if (!same(g, trn::libm::pow(x, y))) ++bad;
      }
      printf("pow%d %ld %ld\n", r, bad, n);
    }
  } else if (!strcmp(op, "cube")) {
    // the quality's m (2 - m) in [0, 1], general magnitudes of both signs,
    // arguments whose cubes are subnormal or overflow; also counts how
    // often the fast form defers to pow (cube_pow lines)
    for (int r = 0; r < 4; ++r) {
      long bad = 0;
      for (long i = 0; i < n; ++i) {
        double x;
        if (r == 0) x = urand();
        else if (r == 1) x = (urand() * 2 - 1) * exp((urand() * 2 - 1) * 200);
        else if (r == 2) x = exp(-urand() * 260);
        else x = exp(230 + urand() * 20) * (urand() < 0.5 ? -1 : 1);
        volatile double g = pow(x, 3.0);
        if (!same(g, trn::libm::cube(x))) ++bad;
      }
      printf("cube%d %ld %ld\n", r, bad, n);
    }
  } else {
    // eigenvalues of metrics (wide), the Riemann-distance ratio (near 1),
    // very close to 1, subnormals
    for (int r = 0; r < 4; ++r) {
      long bad = 0;
      for (long i = 0; i < n; ++i) {
        double x;
        if (r == 0) x = exp((urand() * 2 - 1) * 700);
        else if (r == 1) x = 1 + urand() * 3;
        else if (r == 2) x = 1 + (urand() * 2 - 1) * 1e-3;
        else x = urand() * 0x1p-1022;
        volatile double g = log(x);
        if (!same(g, trn::libm::log(x))) ++bad;
      }
      printf("log%d %ld %ld\n", r, bad, n);
    }
  }
  return 0;
}
