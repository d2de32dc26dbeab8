// Op-counting build of the oracle (test/measurement infrastructure).
//
// mam4_oracle.c is compiled as C++ with `double` replaced by CD, a scalar
// that counts the FP64 operations performed on values that depend on the
// cell's inputs.  Operations whose operands are all literal constants are not
// counted (the HIP kernel evaluates those once per handle), nor are
// comparisons, min/max selects, fabs or casts.  Weights, in FP64 flops:
//   + - * /  and sqrt : 1 each (IEEE basic operations)
//   exp 32, log 90, pow 196, cbrt 28, erfc 121
// The transcendental weights are the flops (FMA = 2) of the gfx950 device
// libm sequences, read from the ISA of one-call kernels (hipcc -O3, gfx950).
//
// Usage: opcount <scene.bin>  -- reads ncol, nlev, dt and the scene planes,
// runs the composed step cell by cell and prints, per cell, the op count and
// the SOA sub-step count.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

static thread_local unsigned long long g_ops = 0;

struct CD {
  double v;
  bool dyn;  // depends on a per-cell input
  CD() = default;
  CD(double x) : v(x), dyn(false) {}
  static CD mk(double x, bool d) {
    CD r(x);
    r.dyn = d;
    return r;
  }
  explicit operator int() const { return (int)v; }
  explicit operator unsigned() const { return (unsigned)v; }
  explicit operator long() const { return (long)v; }
};
static inline CD op(double r, bool d, unsigned long long w = 1) {
  if (d) g_ops += w;
  return CD::mk(r, d);
}
static inline CD operator+(CD a, CD b) { return op(a.v + b.v, a.dyn || b.dyn); }
static inline CD operator-(CD a, CD b) { return op(a.v - b.v, a.dyn || b.dyn); }
static inline CD operator*(CD a, CD b) { return op(a.v * b.v, a.dyn || b.dyn); }
static inline CD operator/(CD a, CD b) { return op(a.v / b.v, a.dyn || b.dyn); }
static inline CD operator-(CD a) { return CD::mk(-a.v, a.dyn); }
static inline CD &operator+=(CD &a, CD b) { return a = a + b; }
static inline CD &operator-=(CD &a, CD b) { return a = a - b; }
static inline CD &operator*=(CD &a, CD b) { return a = a * b; }
static inline CD &operator/=(CD &a, CD b) { return a = a / b; }
static inline bool operator<(CD a, CD b) { return a.v < b.v; }
static inline bool operator>(CD a, CD b) { return a.v > b.v; }
static inline bool operator<=(CD a, CD b) { return a.v <= b.v; }
static inline bool operator>=(CD a, CD b) { return a.v >= b.v; }
static inline bool operator==(CD a, CD b) { return a.v == b.v; }
static inline bool operator!=(CD a, CD b) { return a.v != b.v; }
static inline CD exp(CD a) { return op(std::exp(a.v), a.dyn, 32); }
static inline CD log(CD a) { return op(std::log(a.v), a.dyn, 90); }
static inline CD pow(CD a, CD b) { return op(std::pow(a.v, b.v), a.dyn || b.dyn, 196); }
static inline CD cbrt(CD a) { return op(std::cbrt(a.v), a.dyn, 28); }
static inline CD erfc(CD a) { return op(std::erfc(a.v), a.dyn, 121); }
static inline CD sqrt(CD a) { return op(std::sqrt(a.v), a.dyn, 1); }
static inline CD fabs(CD a) { return CD::mk(std::fabs(a.v), a.dyn); }
static inline CD round(CD a) { return CD::mk(std::round(a.v), a.dyn); }
static inline bool isfinite(CD a) { return std::isfinite(a.v); }
static inline CD nextafter(CD a, double b) { return CD::mk(std::nextafter(a.v, b), a.dyn); }
// functions of the oracle's other entry points (water uptake, chemistry),
// which the chain's step never calls; counted at weight 1
static inline CD operator+(CD a) { return a; }
static inline CD hypot(CD a, CD b) { return op(std::hypot(a.v, b.v), a.dyn || b.dyn); }
static inline CD atan2(CD a, CD b) { return op(std::atan2(a.v, b.v), a.dyn || b.dyn); }
static inline CD cos(CD a) { return op(std::cos(a.v), a.dyn); }
static inline CD sin(CD a) { return op(std::sin(a.v), a.dyn); }
static inline CD log10(CD a) { return op(std::log10(a.v), a.dyn); }
static inline CD acos(CD a) { return op(std::acos(a.v), a.dyn); }
static inline CD tanh(CD a) { return op(std::tanh(a.v), a.dyn); }
static inline bool isnan(CD a) { return std::isnan(a.v); }

#define double CD
#define _Thread_local thread_local
extern "C" {
#include "mam4_oracle.c"
}
#undef double

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s scene.bin\n", argv[0]);
    return 2;
  }
  FILE *f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int32_t hdr[2];
  double dt;
  if (std::fread(hdr, 4, 2, f) != 2 || std::fread(&dt, 8, 1, f) != 1) return 2;
  const int ncol = hdr[0], nlev = hdr[1];
  const long n = (long)ncol * nlev;
  // planes in the order of orc_chain_step: state[52], T, p, z, qv, cld, pblh[ncol],
  // dgncur_a[4], dgncur_awet[4], wetdens[4]
  const long total = 52 * n + 5 * n + ncol + 12 * n;
  std::vector<double> raw(total);
  if ((long)std::fread(raw.data(), 8, total, f) != total) return 2;
  std::fclose(f);
  std::vector<CD> buf(total);
  for (long i = 0; i < total; ++i) buf[i] = CD::mk(raw[i], true);
  CD *st = buf.data(), *T = st + 52 * n, *p = T + n, *z = p + n, *qv = z + n, *cld = qv + n,
     *pblh = cld + n, *dga = pblh + ncol, *dgw = dga + 4 * n, *wd = dgw + 4 * n;
  orc_config cfg = {1, 1, 1, 1, 2, 2, 8, 1, 1, 1, 1};
  std::vector<CD> one(52 + 5 + 1 + 12), tend(52), di(4), dc(4);
  for (int k = 0; k < nlev; ++k)
    for (int i = 0; i < ncol; ++i) {
      const long c = (long)k * ncol + i;
      CD *o = one.data();
      for (int q = 0; q < 52; ++q) o[q] = st[q * n + c];
      o[52] = T[c];
      o[53] = p[c];
      o[54] = z[c];
      o[55] = qv[c];
      o[56] = cld[c];
      o[57] = pblh[i];
      for (int m = 0; m < 4; ++m) {
        o[58 + m] = dga[m * n + c];
        o[62 + m] = dgw[m * n + c];
        o[66 + m] = wd[m * n + c];
      }
      g_ops = 0;
      const long it = orc_chain_step(&cfg, 1, 1, dt, o, o + 52, o + 53, o + 54, o + 55, o + 56,
                                     o + 57, o + 58, o + 62, o + 66, nullptr, nullptr,
                                     tend.data(), di.data(), dc.data(), 1);
      std::printf("%llu %ld\n", g_ops, it);
    }
  return 0;
}
