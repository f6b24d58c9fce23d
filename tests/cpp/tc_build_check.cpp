// Host-only check of the einsum-string front end (tc_build.cpp): build a
// descriptor for each contraction of the CTMRG step and evaluate it with a
// naive strided loop; compare with a direct index-by-index einsum.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>
#include <map>
#include <functional>

#define TNAD_TC_DESC_DEFINED
#define TNAD_ERR_ARG (-1)
#define TNAD_ERR_SHAPE (-2)
namespace tnad {
struct Error {
  int code;
  std::string msg;
  Error(int c, std::string m) : code(c), msg(std::move(m)) {}
};
struct TcDesc {
  int M0 = 1, M1 = 1, N0 = 1, N1 = 1, K0 = 1, K1 = 1;
  long long am0 = 0, am1 = 0, ak0 = 0, ak1 = 0;
  long long bk0 = 0, bk1 = 0, bn0 = 0, bn1 = 0;
  long long cm0 = 0, cm1 = 0, cn0 = 0, cn1 = 0;
  int batch = 1;
  long long abat = 0, bbat = 0, cbat = 0;
  double alpha = 1.0, beta = 0.0;
};
}  // namespace tnad
#include "tc_build.h"
#include "tc_build.cpp"

using namespace tnad;

static void run_desc(const TcDesc& d, const double* A, const double* B, double* C) {
  for (int z = 0; z < d.batch; ++z)
    for (int m = 0; m < d.M0 * d.M1; ++m)
      for (int n = 0; n < d.N0 * d.N1; ++n) {
        double acc = 0;
        for (int k = 0; k < d.K0 * d.K1; ++k) {
          long long a = z * d.abat + (m % d.M0) * d.am0 + (m / d.M0) * d.am1 + (k % d.K0) * d.ak0 + (k / d.K0) * d.ak1;
          long long b = z * d.bbat + (k % d.K0) * d.bk0 + (k / d.K0) * d.bk1 + (n % d.N0) * d.bn0 + (n / d.N0) * d.bn1;
          acc += A[a] * B[b];
        }
        long long c = z * d.cbat + (m % d.M0) * d.cm0 + (m / d.M0) * d.cm1 + (n % d.N0) * d.cn0 + (n / d.N0) * d.cn1;
        C[c] = d.alpha * acc + (d.beta != 0 ? d.beta * C[c] : 0.0);
      }
}

// direct einsum over named indices (views carry strides)
static void naive(const TView& A, const double* a, const TView& B, const double* b, const TView& Cv, double* c) {
  std::map<char, long long> ext;
  for (const TView* v : {&A, &B, &Cv})
    for (size_t i = 0; i < v->idx.size(); ++i) ext[v->idx[i]] = v->dim[i];
  std::vector<char> names;
  for (auto& kv : ext) names.push_back(kv.first);
  std::vector<long long> cur(names.size(), 0);
  auto off = [&](const TView& v) {
    long long o = 0;
    for (size_t i = 0; i < v.idx.size(); ++i)
      for (size_t j = 0; j < names.size(); ++j)
        if (names[j] == v.idx[i]) o += cur[j] * v.str[i];
    return o;
  };
  long long total = 1;
  for (auto& kv : ext) total *= kv.second;
  // zero output
  std::function<void(size_t)> zero;
  for (long long t = 0; t < total; ++t) {
    long long r = t;
    for (size_t j = 0; j < names.size(); ++j) { cur[j] = r % ext[names[j]]; r /= ext[names[j]]; }
    c[off(Cv)] = 0.0;
  }
  for (long long t = 0; t < total; ++t) {
    long long r = t;
    for (size_t j = 0; j < names.size(); ++j) { cur[j] = r % ext[names[j]]; r /= ext[names[j]]; }
    c[off(Cv)] += a[off(A)] * b[off(B)];
  }
}

static size_t span(const TView& v) {
  long long s = 1;
  for (size_t i = 0; i < v.idx.size(); ++i) s += (v.dim[i] - 1) * v.str[i];
  return (size_t)s;
}

int main() {
  std::mt19937_64 rng(7);
  std::normal_distribution<double> nd;
  const long long c = 5, D = 3, s = 2, n = c * D;
  struct Case { TView A, B, C; };
  std::vector<Case> cases = {
    {cm("iba", {c, D, c}), cm("ad", {c, c}), cm("ibd", {c, D, c})},
    {cm("ibd", {c, D, c}), cm("dcl", {c, D, c}), cm("ibcl", {c, D, D, c})},
    {cm("ibcl", {c, D, D, c}), cm("jkcb", {D, D, D, D}), cm("ijlk", {c, D, c, D})},
    {cm("aed", {c, D, c}), cm("abx", {c, D, c}), cm("edbx", {D, c, D, c})},
    {cm("edbx", {D, c, D, c}), cm("bjce", {D, D, D, D}), cm("dcxj", {c, D, c, D})},
    {cm("dcxj", {c, D, c, D}), cm("dcy", {c, D, c}), cm("xjy", {c, D, c})},
    {cm("abx", {c, D, c}), cm("xjy", {c, D, c}), cm("abjy", {c, D, D, c})},
    {cm("abjy", {c, D, D, c}), cm("bjce", {D, D, D, D}), cm("aecy", {c, D, D, c})},
    {cm("aecy", {c, D, D, c}), cm("dcy", {c, D, c}), cm("aed", {c, D, c})},
    {cm("abjy", {c, D, D, c}), cm("aecy", {c, D, D, c}), cm("bjce", {D, D, D, D})},
    {cm("ijlk", {c, D, c, D}), cm("jkcb", {D, D, D, D}), cm("ibcl", {c, D, D, c})},
    {cm("ijlk", {c, D, c, D}), cm("ibcl", {c, D, D, c}), cm("jkcb", {D, D, D, D})},
    {cm("ibcl", {c, D, D, c}), cm("dcl", {c, D, c}), cm("ibd", {c, D, c})},
    {cm("ibd", {c, D, c}), cm("ibcl", {c, D, D, c}), cm("dcl", {c, D, c})},
    {cm("iba", {c, D, c}), cm("ibd", {c, D, c}), cm("ad", {c, c})},
    {cm("icde", {c, D, D, c}), cm("cjfd", {D, D, D, D}), with_strides("ijef", {c, D, c, D}, {1, c, n, n * c})},
    {with_strides("ijeflm", {c, D, c, D, s, s}, {1, c, n, n * c, n * n, n * n * s}), cm("efk", {c, D, c}),
     cm("ijklm", {c, D, c, s, s})},
    {cm("ijklm", {c, D, c, s, s}), cm("efk", {c, D, c}),
     with_strides("ijeflm", {c, D, c, D, s, s}, {1, c, n, n * c, n * n, n * n * s})},
    {with_strides("ijeflm", {c, D, c, D, s, s}, {1, c, n, n * c, n * n, n * n * s}), cm("ijklm", {c, D, c, s, s}),
     cm("efk", {c, D, c})},
    {cm("abcij", {c, D, c, s, s}), cm("ijkl", {s, s, s, s}), cm("abckl", {c, D, c, s, s})},
  };
  int bad = 0;
  for (size_t ci = 0; ci < cases.size(); ++ci) {
    auto& cs = cases[ci];
    std::vector<double> a(span(cs.A)), b(span(cs.B)), o1(span(cs.C), 0.0), o2(span(cs.C), 0.0);
    for (auto& x : a) x = nd(rng);
    for (auto& x : b) x = nd(rng);
    try {
      TcDesc d = contract_desc(cs.A, cs.B, cs.C);
      run_desc(d, a.data(), b.data(), o1.data());
    } catch (const Error& e) {
      std::printf("case %zu: builder error %s\n", ci, e.msg.c_str());
      ++bad;
      continue;
    }
    naive(cs.A, a.data(), cs.B, b.data(), cs.C, o2.data());
    double err = 0;
    for (size_t i = 0; i < o1.size(); ++i) err = std::max(err, std::abs(o1[i] - o2[i]));
    if (err > 1e-10) { std::printf("case %zu: max err %g\n", ci, err); ++bad; }
  }
  std::printf("%zu cases, %d bad\n", cases.size(), bad);
  return bad ? 1 : 0;
}
