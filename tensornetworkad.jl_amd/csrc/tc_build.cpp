#include "tc_build.h"

#include <algorithm>
#include <stdexcept>

namespace tnad {

TView cm(const std::string& idx, const std::vector<long long>& dims) {
  if (idx.size() != dims.size()) throw Error(TNAD_ERR_ARG, "tc view: index/extent count mismatch");
  TView v{idx, dims, {}};
  long long s = 1;
  for (long long d : dims) { v.str.push_back(s); s *= d; }
  return v;
}

TView with_strides(const std::string& idx, const std::vector<long long>& dims,
                   const std::vector<long long>& strides) {
  if (idx.size() != dims.size() || idx.size() != strides.size())
    throw Error(TNAD_ERR_ARG, "tc view: index/extent/stride count mismatch");
  return TView{idx, dims, strides};
}

namespace {

struct Sub {
  long long dim = 1;
  long long sa = 0, sb = 0, sc = 0;   // stride in A, B, C (0 if absent)
};

int pos(const TView& v, char c) {
  auto p = v.idx.find(c);
  return p == std::string::npos ? -1 : (int)p;
}

// try to merge a group (ordered list of subs) down to at most `cap` sub-indices
bool merge_group(std::vector<Sub> g, int cap, std::vector<Sub>& out) {
  auto lex = [](const Sub& x, const Sub& y) {
    if (x.sa != y.sa) return x.sa < y.sa;
    if (x.sb != y.sb) return x.sb < y.sb;
    if (x.sc != y.sc) return x.sc < y.sc;
    return x.dim < y.dim;
  };
  std::sort(g.begin(), g.end(), lex);
  // try every ordering (groups have at most a handful of members) and keep the
  // first that merges down to `cap` sub-indices
  do {
    std::vector<Sub> cur;
    for (const Sub& s : g) {
      if (!cur.empty()) {
        Sub& last = cur.back();
        auto ok = [&](long long s0, long long s1) { return s0 == 0 ? s1 == 0 : s1 == s0 * last.dim; };
        if (ok(last.sa, s.sa) && ok(last.sb, s.sb) && ok(last.sc, s.sc)) {
          last.dim *= s.dim;
          continue;
        }
      }
      cur.push_back(s);
    }
    if ((int)cur.size() <= cap) {
      out = cur;
      return true;
    }
  } while (std::next_permutation(g.begin(), g.end(), lex));
  return false;
}

}  // namespace

TcDesc contract_desc(const TView& A, const TView& B, const TView& C, double alpha, double beta) {
  std::vector<Sub> gb, gm, gn, gk;
  std::string seen;
  auto dim_of = [&](char c) -> long long {
    long long d = -1;
    for (const TView* v : {&A, &B, &C}) {
      int p = pos(*v, c);
      if (p >= 0) {
        if (d >= 0 && d != v->dim[p])
          throw Error(TNAD_ERR_SHAPE, std::string("tc: extent mismatch for index ") + c);
        d = v->dim[p];
      }
    }
    return d;
  };
  auto sub_of = [&](char c) {
    Sub s;
    s.dim = dim_of(c);
    int pa = pos(A, c), pb = pos(B, c), pc = pos(C, c);
    s.sa = pa >= 0 ? A.str[pa] : 0;
    s.sb = pb >= 0 ? B.str[pb] : 0;
    s.sc = pc >= 0 ? C.str[pc] : 0;
    return s;
  };
  for (char c : C.idx) {
    bool ia = pos(A, c) >= 0, ib = pos(B, c) >= 0;
    if (ia && ib) gb.push_back(sub_of(c));
    else if (ia) gm.push_back(sub_of(c));
    else if (ib) gn.push_back(sub_of(c));
    else throw Error(TNAD_ERR_ARG, std::string("tc: output index only in C: ") + c);
  }
  for (char c : A.idx) {
    bool ib = pos(B, c) >= 0, ic = pos(C, c) >= 0;
    if (ib && !ic) gk.push_back(sub_of(c));
    if (!ib && !ic) throw Error(TNAD_ERR_ARG, std::string("tc: index summed inside A only: ") + c);
  }
  for (char c : B.idx) {
    if (pos(A, c) < 0 && pos(C, c) < 0)
      throw Error(TNAD_ERR_ARG, std::string("tc: index summed inside B only: ") + c);
  }
  std::vector<Sub> mb, mm, mn, mk;
  if (!merge_group(gb, 1, mb) || !merge_group(gm, 2, mm) || !merge_group(gn, 2, mn) ||
      !merge_group(gk, 2, mk))
    throw Error(TNAD_ERR_ARG, "tc: index group does not reduce to two strided sub-indices");
  if (mk.empty()) throw Error(TNAD_ERR_ARG, "tc: no contracted index");
  TcDesc d;
  auto fill2 = [](const std::vector<Sub>& g, int& x0, int& x1, long long& sa0, long long& sa1,
                  long long Sub::*f1, long long& sb0, long long& sb1, long long Sub::*f2) {
    x0 = 1; x1 = 1; sa0 = sa1 = sb0 = sb1 = 0;
    if (g.size() > 0) { x0 = (int)g[0].dim; sa0 = g[0].*f1; sb0 = g[0].*f2; }
    if (g.size() > 1) { x1 = (int)g[1].dim; sa1 = g[1].*f1; sb1 = g[1].*f2; }
  };
  fill2(mm, d.M0, d.M1, d.am0, d.am1, &Sub::sa, d.cm0, d.cm1, &Sub::sc);
  fill2(mn, d.N0, d.N1, d.bn0, d.bn1, &Sub::sb, d.cn0, d.cn1, &Sub::sc);
  fill2(mk, d.K0, d.K1, d.ak0, d.ak1, &Sub::sa, d.bk0, d.bk1, &Sub::sb);
  if (!mb.empty()) {
    d.batch = (int)mb[0].dim;
    d.abat = mb[0].sa; d.bbat = mb[0].sb; d.cbat = mb[0].sc;
  }
  d.alpha = alpha;
  d.beta = beta;
  return d;
}

}  // namespace tnad
