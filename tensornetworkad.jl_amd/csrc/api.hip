// extern "C" entry points of libtnad (include/tnad.h): context and memory
// management, the CTMRG fixed point with the reference's StopFunction
// semantics, checkpoint tapes for the reverse sweep, and the fused
// energy + gradient call.
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>

#include "api_internal.h"
#include "ctmrg.h"
#include "trg.h"
#include "zc.h"

namespace tnad {

// ---------------------------------------------------------------------------
// memory: a size-keyed cache of device blocks per context
// ---------------------------------------------------------------------------
void* Ctx::alloc(size_t bytes) {
  if (bytes == 0) bytes = 8;
  bytes = (bytes + 255) & ~size_t(255);
  auto it = free_blocks.find(bytes);
  if (it != free_blocks.end()) {
    void* p = it->second;
    free_blocks.erase(it);
    cached_bytes -= bytes;
    live[(uintptr_t)p] = bytes;
    return p;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    trim();
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error(TNAD_ERR_ALLOC, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
  }
  live[(uintptr_t)p] = bytes;
  return p;
}
void Ctx::check_range(const void* p, size_t bytes, const char* what) const {
  const uintptr_t a = (uintptr_t)p;
  auto it = live.upper_bound(a);
  if (it == live.begin()) throw Error(TNAD_ERR_ARG, std::string(what) + ": operand is not a live device block");
  --it;
  if (a + bytes > it->first + it->second)
    throw Error(TNAD_ERR_ARG, std::string(what) + ": operand extent " + std::to_string(bytes) +
                                  " bytes exceeds its device block (" + std::to_string(it->first + it->second - a) +
                                  " bytes available)");
}
void Ctx::release(void* p, size_t bytes) {
  if (!p) return;
  if (bytes == 0) bytes = 8;
  bytes = (bytes + 255) & ~size_t(255);
  live.erase((uintptr_t)p);
  free_blocks.emplace(bytes, p);
  cached_bytes += bytes;
  if (cached_bytes > (size_t(32) << 30)) trim();
}
void Ctx::trim() {
  // blocks may still be referenced by queued work
  (void)hipStreamSynchronize(stream);
  for (auto& kv : free_blocks) (void)hipFree(kv.second);
  free_blocks.clear();
  cached_bytes = 0;
}
hipEvent_t Ctx::mark(int slot) {
  if (!ev_[slot]) TNAD_HIP_TRY(hipEventCreate(&ev_[slot]));
  TNAD_HIP_TRY(hipEventRecord(ev_[slot], stream));
  return ev_[slot];
}
float Ctx::elapsed(int a, int b) {
  float ms = 0.0f;
  if (ev_[a] && ev_[b] && hipEventElapsedTime(&ms, ev_[a], ev_[b]) != hipSuccess) {
    (void)hipGetLastError();
    ms = 0.0f;
  }
  return ms;
}
Ctx::~Ctx() {
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto& e : ev_)
    if (e) (void)hipEventDestroy(e);
  for (auto& kv : free_blocks) (void)hipFree(kv.second);
  if (pinned_) (void)hipHostFree(pinned_);
  if (stream) (void)hipStreamDestroy(stream);
}

// ---------------------------------------------------------------------------
// pinned staging of small host <-> device transfers
// ---------------------------------------------------------------------------
constexpr size_t kPinnedBytes = size_t(1) << 20;
char* Ctx::stage(size_t bytes, size_t& off) {
  if (bytes > kPinnedBytes / 2) return nullptr;
  if (!pinned_) {
    void* p = nullptr;
    TNAD_HIP_TRY(hipHostMalloc(&p, kPinnedBytes, hipHostMallocDefault));
    pinned_ = static_cast<char*>(p);
  }
  size_t need = (bytes + 15) & ~size_t(15);
  if (pinned_used_ + need > kPinnedBytes) sync();   // drains the queued transfers, frees the area
  off = pinned_used_;
  pinned_used_ += need;
  return pinned_ + off;
}
void Ctx::read_async(void* dst_host, const void* src_dev, size_t bytes) {
  if (bytes == 0) return;
  size_t off = 0;
  char* st = stage(bytes, off);
  if (!st) {   // large: straight into the destination (complete at the next sync as well)
    TNAD_HIP_TRY(hipMemcpyAsync(dst_host, src_dev, bytes, hipMemcpyDeviceToHost, stream));
    return;
  }
  TNAD_HIP_TRY(hipMemcpyAsync(st, src_dev, bytes, hipMemcpyDeviceToHost, stream));
  pending_.push_back({dst_host, off, bytes});
}
void Ctx::write_async(void* dst_dev, const void* src_host, size_t bytes) {
  if (bytes == 0) return;
  size_t off = 0;
  char* st = stage(bytes, off);
  if (!st) {
    TNAD_HIP_TRY(hipMemcpyAsync(dst_dev, src_host, bytes, hipMemcpyHostToDevice, stream));
    return;
  }
  std::memcpy(st, src_host, bytes);
  TNAD_HIP_TRY(hipMemcpyAsync(dst_dev, st, bytes, hipMemcpyHostToDevice, stream));
}
void Ctx::sync() {
  TNAD_HIP_TRY(hipStreamSynchronize(stream));
  for (const auto& r : pending_) std::memcpy(r.dst, pinned_ + r.off, r.bytes);
  pending_.clear();
  pinned_used_ = 0;
}

DevBuf::DevBuf(Ctx* c, size_t b) : ptr(c->alloc(b)), bytes(b), owner(c) {}
DevBuf& DevBuf::operator=(DevBuf&& o) noexcept {
  if (this != &o) {
    if (ptr && owner) owner->release(ptr, bytes);
    ptr = o.ptr; bytes = o.bytes; owner = o.owner;
    o.ptr = nullptr; o.bytes = 0; o.owner = nullptr;
  }
  return *this;
}
DevBuf::~DevBuf() {
  // blocks return to the cache; work queued on the context's single stream
  // that still reads them is ordered before any later reuse.
  if (ptr && owner) owner->release(ptr, bytes);
}

void copy_in(Ctx* c, void* dst_dev, const void* src_any, size_t bytes) {
  if (!src_any) throw Error(TNAD_ERR_ARG, "null input pointer");
  TNAD_HIP_TRY(hipMemcpyAsync(dst_dev, src_any, bytes, hipMemcpyDefault, c->stream));
}
void copy_out(Ctx* c, void* dst_any, const void* src_dev, size_t bytes) {
  if (!dst_any) throw Error(TNAD_ERR_ARG, "null output pointer");
  TNAD_HIP_TRY(hipMemcpyAsync(dst_any, src_dev, bytes, hipMemcpyDefault, c->stream));
}

}  // namespace tnad

using namespace tnad;


struct tnad_tape {
  tnad_ctx* ctx = nullptr;
  StepDims d;
  std::vector<std::unique_ptr<StepTape>> steps;
  DevBuf a64;
  PBuf ap[3];          // a at each contraction precision (lazily filled)
  bool have[3] = {false, false, false};
  tnad_opts opts{};
};

namespace {

Prec to_prec_enum(int p) {
  if (p == TNAD_PREC_FP64) return Prec::F64;
  if (p == TNAD_PREC_FP32) return Prec::F32;
  if (p == TNAD_PREC_BF16) return Prec::BF16;
  throw Error(TNAD_ERR_ARG, "unknown precision");
}

void check_dims(int D, int chi) { enforce(ctmrg_dims_limit(D, chi)); }

const PBuf& ap_at(Ctx* c, tnad_tape& t, Prec p) {
  int i = (int)p;
  if (!t.have[i]) {
    size_t nA = (size_t)t.d.D * t.d.D * t.d.D * t.d.D;
    t.ap[i] = PBuf(c, p, nA);
    to_prec(c, t.a64.as<double>(), t.ap[i]);
    t.have[i] = true;
  }
  return t.ap[i];
}

// The fixed point (src/ctmrg.jl:110-117, src/fixedpoint.jl:11-41) on device buffers.
// `t` always carries a and its precision copies; its `steps` are filled only when keep_steps.
void run_ctmrg(tnad_ctx* ctx, tnad_tape& t, const double* C0, const double* T0, double tol, int maxit,
               bool keep_steps, double* C_out, double* T_out, std::vector<double>& vals_out, int& iters) {
  Ctx* c = &ctx->c;
  const tnad_opts& o = t.opts;
  const StepDims d = t.d;
  const int chi = d.chi, D = d.D, n = d.n;
  const size_t nC = (size_t)chi * chi, nT = (size_t)chi * D * chi;
  DevBuf Ca(c, nC * 8), Ta(c, nT * 8), Cb(c, nC * 8), Tb(c, nT * 8);
  k_copy(C0, Ca.as<double>(), nC, c->stream);
  k_copy(T0, Ta.as<double>(), nT, c->stream);
  const Prec low = to_prec_enum(o.precision);
  const bool lowp = low != Prec::F64;
  bool switched = !lowp;
  StepOpts so;
  so.eig_block = o.eig_block;
  so.eig_max_iters = o.eig_max_iters > 0 ? o.eig_max_iters : 200;
  so.eig_warm_iters = o.eig_warm_iters > 0 ? o.eig_warm_iters : 200;
  so.eig_tol = o.eig_tol > 0 ? o.eig_tol : 1e-13;
  const double base_eig_tol = so.eig_tol;
  // a low-precision step's contractions carry ~1e-7 relative error: its eigenvectors need not
  // be resolved beyond what the switch test can see
  so.eig_tol_low = std::max(so.eig_tol, 1e-3 * o.switch_tol);
  const int k = step_k(d, so);
  DevBuf Xprev;            // warm start block when no tape is kept
  const double* X0 = nullptr;
  std::vector<double> oldvals(n, std::numeric_limits<double>::infinity());
  std::vector<double> vals = oldvals;
  // n > 192: the spectrum beyond the chi kept values enters through its norm (StepState::rest)
  const bool partial = n > JACOBI_MAX;
  const int nk = partial ? chi : n;
  double rest = std::numeric_limits<double>::infinity(), rest_err = 0.0;
  double oldrest = rest, oldrest_err = 0.0;
  // StopFunction(oldvals, -1, tol, maxit) (src/ctmrg.jl:114): the counter starts at -1 and the test runs on
  // the initial state too, so at most maxit + 1 steps are taken and maxit = 0 still takes one
  int counter = -1;
  iters = 0;
  tnad_stats& st = ctx->stats;
  st = tnad_stats{};
  st.eig_k = k;
  c->mark(0);
  while (true) {
    // StopFunction (src/fixedpoint.jl:31-41), evaluated on every state including the first
    counter += 1;
    if (counter > maxit) break;
    // n <= 192: |vals - oldvals|_2 over the whole spectrum, as the reference.  n > 192: the change of the
    // chi kept values plus | |rest| - |oldrest| | (less both error bounds) for the discarded ones, a lower
    // bound on the reference's norm (triangle inequality): a run never stops after the reference, and
    // before it only when the discarded values move by more than their norm shows
    double diff2 = 0.0;
    for (int i = 0; i < nk; ++i) {
      double dv = vals[i] - oldvals[i];
      diff2 += dv * dv;
    }
    if (partial) {
      double lb = std::fabs(rest - oldrest) - rest_err - oldrest_err;   // Inf while oldrest is Inf
      if (lb > 0.0) diff2 += lb * lb;
    }
    double diff = std::sqrt(diff2);   // NaN for the first state (Inf - Inf), Inf for the second
    if (diff <= tol) break;
    if (!switched && diff <= o.switch_tol) switched = true;
    oldvals = vals;
    oldrest = rest;
    oldrest_err = rest_err;
    so.prec = switched ? Prec::F64 : low;
    // in a mixed run an FP64 step before convergence resolves its eigenvectors only as far as the
    // step-to-step change can see: 1e-3 of |vals - oldvals|, between eig_tol and 1e-10 (the FP64 steps'
    // adjoint assumes accurate eigenpairs: with tolerances near the low-precision steps' ones, e.g. 1e-5 for
    // a BF16 switch at 1e-2, the reverse sweep's resolvent solves lose the gradient); the plain path keeps
    // eig_tol on every step
    if (lowp)
      so.eig_tol = std::isfinite(diff) ? std::max(base_eig_tol, std::min(1e-10, 1e-3 * diff)) : base_eig_tol;
    const PBuf& ap = ap_at(c, t, so.prec);
    StepTape* tp = nullptr;
    if (keep_steps) {
      t.steps.emplace_back(new StepTape(c, d, so.prec, k));
      tp = t.steps.back().get();
    }
    StepState ss;
    double* X_keep = nullptr;
    const bool warm = o.eig_warm && n > JACOBI_MAX;
    if (warm && !tp) {
      if (!Xprev.ptr) Xprev = DevBuf(c, (size_t)n * k * 8);
      X_keep = Xprev.as<double>();
    }
    ctmrg_step(c, d, so, t.a64.as<double>(), ap, Ca.as<double>(), Ta.as<double>(), warm ? X0 : nullptr, X_keep,
               Cb.as<double>(), Tb.as<double>(), vals, tp, ss);
    if (warm) X0 = tp ? tp->X.as<double>() : Xprev.as<double>();
    rest = ss.rest;
    rest_err = ss.rest_err;
    std::swap(Ca, Cb);
    std::swap(Ta, Tb);
    iters += 1;
    st.iters = iters;
    if (so.prec != Prec::F64) st.low_prec_iters += 1;
    st.eig_iters_total += ss.eig_iters;
    st.eig_resid_max = std::max(st.eig_resid_max, ss.eig_resid);
    st.eig_ms += ss.eig_ms;
  }
  c->mark(1);
  c->sync();
  st.fwd_ms = c->elapsed(0, 1);
  k_copy(Ca.as<double>(), C_out, nC, c->stream);
  k_copy(Ta.as<double>(), T_out, nT, c->stream);
  vals_out = vals;
}

}  // namespace

// ---------------------------------------------------------------------------
// Device scratch of the entry point that allocates nothing (tnad_symeig_noalloc): one file-scope
// array per GPU, sized for tnad_symeig's full-spectrum solver (divide and conquer) at n = 192
// (reflectors, two eigenvector buffers and the merge matrices, 4 x 192 x 192 doubles, plus the
// tridiagonal matrix, tau, the eigenvalues, the reflector block products and the top merge's vectors:
// 150,008 doubles, 1.20 MB).  Every call on a GPU shares
// that GPU's copy, and every word a call reads there is written earlier in the same call.  Host code
// never passes `scratch` itself to a kernel: its address on each device comes from hipGetSymbolAddress
// at tnad_create.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kScratchDoubles = 4 * 192 * 192 + 13 * 192 + 56;   // syev_dc_work_doubles(192)
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};   // device address of `scratch` on each device
}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
// with `device` current: look up this device's copy of `scratch` once
hipError_t scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipSuccess;   // tnad_symeig_noalloc reports it
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return hipSuccess;
  double* p = nullptr;
  hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch[device] = p;
  return hipSuccess;
}
double* scratch_for(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices || !g_scratch[device])
    throw Error(TNAD_ERR_UNSUPPORTED, "tnad_symeig_noalloc: no scratch on device " + std::to_string(device));
  return g_scratch[device];
}
// throw unless p is device memory of the context's GPU (no allocation: an attribute query)
void require_device(const Ctx* c, const void* p, const char* what) {
  if (!p) throw Error(TNAD_ERR_ARG, std::string("tnad_symeig_noalloc: null ") + what);
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();
    throw Error(TNAD_ERR_ARG, std::string("tnad_symeig_noalloc: ") + what + " is not device memory");
  }
  if (at.type != hipMemoryTypeDevice || at.device != c->device)
    throw Error(TNAD_ERR_ARG, std::string("tnad_symeig_noalloc: ") + what +
                                  " must be device memory on the context's GPU");
}
}  // namespace

extern "C" {

int tnad_version(void) { return TNAD_ABI_VERSION; }

void tnad_opts_default(tnad_opts* o) {
  if (!o) return;
  o->precision = TNAD_PREC_FP64;
  o->switch_tol = 0.0;
  o->bwd_precision = TNAD_PREC_FP64;
  o->eig_block = 0;
  o->eig_max_iters = 200;
  o->eig_warm = 0;
  o->eig_warm_iters = 200;
  o->eig_tol = 1e-13;
  o->bwd_last = 0;
}

int tnad_create(int device, tnad_ctx** out) {
  if (!out) return TNAD_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return TNAD_ERR_HIP;
  if (hipSetDevice(device) != hipSuccess) return TNAD_ERR_HIP;
  tnad_ctx* ctx = new tnad_ctx();
  ctx->c.device = device;
  if (hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return TNAD_ERR_HIP;
  }
  if (scratch_setup(device) != hipSuccess) {
    (void)hipGetLastError();
    delete ctx;
    return TNAD_ERR_HIP;
  }
  *out = ctx;
  return TNAD_OK;
}

int tnad_destroy(tnad_ctx* ctx) {
  if (!ctx) return TNAD_ERR_ARG;
  (void)hipSetDevice(ctx->c.device);
  delete ctx;
  return TNAD_OK;
}

const char* tnad_last_error(const tnad_ctx* ctx) { return ctx ? ctx->c.last_error.c_str() : "null context"; }

int tnad_get_stats(const tnad_ctx* ctx, tnad_stats* out) {
  if (!ctx || !out) return TNAD_ERR_ARG;
  *out = ctx->stats;
  return TNAD_OK;
}

int tnad_indexperm_symmetrize(tnad_ctx* ctx, const double* bulk, int d, int s, double* out) {
  return guarded(ctx, [&] {
    if (d < 1 || s < 1) throw Error(TNAD_ERR_SHAPE, "size of tensor error, should be (d, d, d, d, s)");
    Ctx* c = &ctx->c;
    size_t N = (size_t)d * d * d * d * s;
    DevBuf x(c, N * 8), y(c, N * 8);
    copy_in(c, x.ptr, bulk, N * 8);
    k_symmetrize_bulk(x.as<double>(), d, s, y.as<double>(), c->stream);
    copy_out(c, out, y.ptr, N * 8);
  });
}

int tnad_double_layer(tnad_ctx* ctx, const double* bulk, int d, int s, double* ap, double* a) {
  return guarded(ctx, [&] {
    if (d < 1 || s < 1) throw Error(TNAD_ERR_SHAPE, "size of tensor error, should be (d, d, d, d, s)");
    Ctx* c = &ctx->c;
    size_t N = (size_t)d * d * d * d * s, D4 = (size_t)d * d * d * d * d * d * d * d;
    DevBuf x(c, N * 8), bap(c, D4 * s * s * 8), ba(c, D4 * 8);
    copy_in(c, x.ptr, bulk, N * 8);
    k_double_layer(x.as<double>(), d, s, ap ? bap.as<double>() : nullptr, ba.as<double>(), c->stream);
    if (ap) copy_out(c, ap, bap.ptr, D4 * s * s * 8);
    if (a) copy_out(c, a, ba.ptr, D4 * 8);
  });
}

int tnad_init_raw(tnad_ctx* ctx, const double* a, int D, int chi, double* C, double* T) {
  return guarded(ctx, [&] {
    enforce(env_dims_limit(D, chi));
    Ctx* c = &ctx->c;
    size_t nA = (size_t)D * D * D * D, nC = (size_t)chi * chi, nT = (size_t)chi * D * chi;
    DevBuf ba(c, nA * 8), bC(c, nC * 8), bT(c, nT * 8);
    copy_in(c, ba.ptr, a, nA * 8);
    k_init_raw(ba.as<double>(), D, chi, bC.as<double>(), bT.as<double>(), c->stream);
    copy_out(c, C, bC.ptr, nC * 8);
    copy_out(c, T, bT.ptr, nT * 8);
  });
}

int tnad_symeig(tnad_ctx* ctx, const double* A, int n, int want, int k, double* evals, double* evecs, int* k_out) {
  return guarded(ctx, [&] {
    enforce(symeig_limit(n, want));
    Ctx* c = &ctx->c;
    const int kk = eig_block_width(n, want, k);
    DevBuf bA(c, (size_t)n * n * 8), X(c, (size_t)n * kk * 8), AX(c, (size_t)n * kk * 8), th(c, (size_t)kk * 8);
    copy_in(c, bA.ptr, A, (size_t)n * n * 8);
    int kr = topk_eig(c, bA.as<double>(), n, want, kk, nullptr, 200, 1e-13, X.as<double>(), AX.as<double>(),
                      th.as<double>(), nullptr);
    if (evals) copy_out(c, evals, th.ptr, (size_t)kr * 8);
    if (evecs) copy_out(c, evecs, X.ptr, (size_t)n * kr * 8);
    if (k_out) *k_out = kr;
  });
}

int tnad_symeig_noalloc(tnad_ctx* ctx, const double* A, int n, double* evals, double* evecs) {
  return guarded(ctx, [&] {
    if (n < 1 || n > JACOBI_MAX)
      throw Error(TNAD_ERR_UNSUPPORTED, "tnad_symeig_noalloc: n must be 1..192, got " + std::to_string(n));
    Ctx* c = &ctx->c;
    require_device(c, A, "A");
    require_device(c, evals, "evals");
    require_device(c, evecs, "evecs");
    double* ws = scratch_for(c->device);
    if (syev_dc_work_doubles(n) > kScratchDoubles)
      throw Error(TNAD_ERR_UNSUPPORTED, "tnad_symeig_noalloc: work exceeds the scratch array");
    // the kernels of tnad_symeig's full-spectrum path, on the scratch array instead of a pool block
    syev_dc(A, n, n, evals, evecs, ws, c->stream);
  });
}

int tnad_energy_z(tnad_ctx* ctx, const double* h, const double* bulk, int d, int s, int chi, double tol, int maxit,
                  double* e, double* grad, int* iters) {
  return guarded(ctx, [&] {
    if (d < 1 || s < 1) throw Error(TNAD_ERR_SHAPE, "size of tensor error, should be (d, d, d, d, s)");
    enforce(zctmrg_dims_limit((long long)d * d, chi));
    if (!h || !bulk) throw Error(TNAD_ERR_ARG, "null input");
    Ctx* c = &ctx->c;
    const size_t nb5 = (size_t)d * d * d * d * s, nh = (size_t)s * s * s * s;
    DevBuf hb(c, 2 * nh * 8), bb(c, 2 * nb5 * 8), gb;
    z_from_interleaved(c, h, nh, hb.as<double>());
    z_from_interleaved(c, bulk, nb5, bb.as<double>());
    if (grad) gb = DevBuf(c, 2 * nb5 * 8);
    double eh[2] = {0.0, 0.0};
    int it = 0;
    zenergy_run(c, hb.as<double>(), bb.as<double>(), d, s, chi, tol, maxit, eh, grad ? gb.as<double>() : nullptr, it);
    if (grad) z_to_interleaved(c, gb.as<double>(), nb5, grad);
    if (e) { e[0] = eh[0]; e[1] = eh[1]; }
    if (iters) *iters = it;
    ctx->stats = tnad_stats{};
    ctx->stats.iters = it;
    ctx->stats.eig_k = chi * d * d;
  });
}

int tnad_trg_z(tnad_ctx* ctx, const double* a, int q, int chi, int niter, double tol, double* lnz, double* abar) {
  return guarded(ctx, [&] {
    enforce(ztrg_limit(q, chi, niter));
    if (!a) throw Error(TNAD_ERR_ARG, "null input");
    Ctx* c = &ctx->c;
    const size_t n = (size_t)q * q * q * q;
    DevBuf ab(c, 2 * n * 8), gb;
    z_from_interleaved(c, a, n, ab.as<double>());
    if (abar) gb = DevBuf(c, 2 * n * 8);
    double l2[2] = {0.0, 0.0};
    ztrg_run(c, ab.as<double>(), q, chi, niter, tol, l2, abar ? gb.as<double>() : nullptr);
    if (abar) z_to_interleaved(c, gb.as<double>(), n, abar);
    if (lnz) { lnz[0] = l2[0]; lnz[1] = l2[1]; }
  });
}

int tnad_svd_z(tnad_ctx* ctx, const double* A, int m, int p, double* U, double* S, double* V) {
  return guarded(ctx, [&] {
    enforce(zsvd_limit(m, p));
    if (!A) throw Error(TNAD_ERR_ARG, "null input");
    Ctx* c = &ctx->c;
    const int K = std::min(m, p);
    DevBuf Ab(c, 2 * (size_t)m * p * 8), Ub(c, 2 * (size_t)m * K * 8), Vb(c, 2 * (size_t)p * K * 8), Sb(c, (size_t)K * 8);
    z_from_interleaved(c, A, (size_t)m * p, Ab.as<double>());
    zsvd_small(c, Ab.as<double>(), m, p, Ub.as<double>(), Sb.as<double>(), Vb.as<double>());
    if (U) z_to_interleaved(c, Ub.as<double>(), (size_t)m * K, U);
    if (V) z_to_interleaved(c, Vb.as<double>(), (size_t)p * K, V);
    if (S) copy_out(c, S, Sb.ptr, (size_t)K * 8);
  });
}

int tnad_svd_back_z(tnad_ctx* ctx, const double* A, int m, int p, const double* dU, const double* dS, const double* dV,
                    double* dA) {
  return guarded(ctx, [&] {
    enforce(zsvd_limit(m, p));
    if (!A || !dA) throw Error(TNAD_ERR_ARG, "null input");
    Ctx* c = &ctx->c;
    const int K = std::min(m, p);
    DevBuf Ab(c, 2 * (size_t)m * p * 8), dAb(c, 2 * (size_t)m * p * 8), dUb, dVb, dSb;
    z_from_interleaved(c, A, (size_t)m * p, Ab.as<double>());
    if (dU) { dUb = DevBuf(c, 2 * (size_t)m * K * 8); z_from_interleaved(c, dU, (size_t)m * K, dUb.as<double>()); }
    if (dV) { dVb = DevBuf(c, 2 * (size_t)p * K * 8); z_from_interleaved(c, dV, (size_t)p * K, dVb.as<double>()); }
    if (dS) { dSb = DevBuf(c, (size_t)K * 8); copy_in(c, dSb.ptr, dS, (size_t)K * 8); }
    if (!dU && !dS && !dV) {   // svd_back returns nothing (src/trg.jl:73): a zero cotangent
      k_fill(dAb.as<double>(), 2 * (size_t)m * p, 0.0, c->stream);
    } else {
      zsvd_back_small(c, Ab.as<double>(), m, p, dU ? dUb.as<double>() : nullptr, dS ? dSb.as<double>() : nullptr,
                      dV ? dVb.as<double>() : nullptr, dAb.as<double>());
    }
    z_to_interleaved(c, dAb.as<double>(), (size_t)m * p, dA);
  });
}

int tnad_svd_back_sym(tnad_ctx* ctx, const double* A, int n, int chi, int k, const double* G, double* Abar) {
  return guarded(ctx, [&] {
    enforce(svd_back_sym_limit(n, chi));
    Ctx* c = &ctx->c;
    StepDims d;
    d.chi = chi; d.D = n / chi; d.n = n;
    const int kk = eig_block_width(n, chi, k);
    StepTape tp(c, d, Prec::F64, kk);
    copy_in(c, tp.M.ptr, A, (size_t)n * n * 8);
    tp.k = topk_eig(c, tp.M.as<double>(), n, chi, kk, nullptr, 200, 1e-13, tp.X.as<double>(), tp.AX.as<double>(),
                    tp.theta.as<double>(), nullptr);
    DevBuf bG(c, (size_t)n * chi * 8), zero(c, (size_t)chi * chi * 8), out(c, (size_t)n * n * 8);
    copy_in(c, bG.ptr, G, (size_t)n * chi * 8);
    k_fill(zero.as<double>(), (size_t)chi * chi, 0.0, c->stream);
    eig_adjoint(c, d, tp, bG.as<double>(), zero.as<double>(), out.as<double>(), nullptr);
    copy_out(c, Abar, out.ptr, (size_t)n * n * 8);
  });
}

int tnad_ctmrgstep(tnad_ctx* ctx, const double* a, int D, int chi, const double* C, const double* T,
                   const tnad_opts* opts, double* Cout, double* Tout, double* vals) {
  return guarded(ctx, [&] {
    check_dims(D, chi);
    Ctx* c = &ctx->c;
    tnad_opts o;
    tnad_opts_default(&o);
    if (opts) o = *opts;
    StepDims d;
    d.chi = chi; d.D = D; d.n = chi * D;
    size_t nA = (size_t)D * D * D * D, nC = (size_t)chi * chi, nT = (size_t)chi * D * chi;
    DevBuf ba(c, nA * 8), bC(c, nC * 8), bT(c, nT * 8), oC(c, nC * 8), oT(c, nT * 8);
    copy_in(c, ba.ptr, a, nA * 8);
    copy_in(c, bC.ptr, C, nC * 8);
    copy_in(c, bT.ptr, T, nT * 8);
    StepOpts so;
    so.prec = to_prec_enum(o.precision);
    so.eig_block = o.eig_block;
    so.eig_max_iters = o.eig_max_iters > 0 ? o.eig_max_iters : 200;
    so.eig_tol = o.eig_tol > 0 ? o.eig_tol : 1e-13;
    so.eig_tol_low = so.eig_tol;
    PBuf ap(c, so.prec, nA);
    to_prec(c, ba.as<double>(), ap);
    std::vector<double> v;
    StepState ss;
    ctmrg_step(c, d, so, ba.as<double>(), ap, bC.as<double>(), bT.as<double>(), nullptr, nullptr,
               oC.as<double>(), oT.as<double>(), v, nullptr, ss);
    copy_out(c, Cout, oC.ptr, nC * 8);
    copy_out(c, Tout, oT.ptr, nT * 8);
    if (vals) std::memcpy(vals, v.data(), v.size() * 8);
  });
}

int tnad_ctmrg(tnad_ctx* ctx, const double* a, int D, int chi, const double* C0, const double* T0, double tol,
               int maxit, const tnad_opts* opts, double* C, double* T, double* vals, int* iters,
               tnad_tape** tape_out) {
  std::unique_ptr<tnad_tape> t(new tnad_tape());
  int rc = guarded(ctx, [&] {
    check_dims(D, chi);
    if (maxit < 0) throw Error(TNAD_ERR_ARG, "maxit must be >= 0");
    Ctx* c = &ctx->c;
    t->ctx = ctx;
    tnad_opts_default(&t->opts);
    if (opts) t->opts = *opts;
    t->d.chi = chi; t->d.D = D; t->d.n = chi * D;
    size_t nA = (size_t)D * D * D * D, nC = (size_t)chi * chi, nT = (size_t)chi * D * chi;
    t->a64 = DevBuf(c, nA * 8);
    copy_in(c, t->a64.ptr, a, nA * 8);
    DevBuf bC0(c, nC * 8), bT0(c, nT * 8), oC(c, nC * 8), oT(c, nT * 8);
    copy_in(c, bC0.ptr, C0, nC * 8);
    copy_in(c, bT0.ptr, T0, nT * 8);
    std::vector<double> v;
    int it = 0;
    run_ctmrg(ctx, *t, bC0.as<double>(), bT0.as<double>(), tol, maxit, tape_out != nullptr, oC.as<double>(),
              oT.as<double>(), v, it);
    copy_out(c, C, oC.ptr, nC * 8);
    copy_out(c, T, oT.ptr, nT * 8);
    c->sync();
    if (vals) std::memcpy(vals, v.data(), v.size() * 8);
    if (iters) *iters = it;
  });
  if (rc == TNAD_OK && tape_out) *tape_out = t.release();
  else if (tape_out) *tape_out = nullptr;
  return rc;
}

static void ctmrg_back_impl(tnad_ctx* ctx, tnad_tape* t, const double* Cbar_dev, const double* Tbar_dev,
                            double* abar_dev) {
  Ctx* c = &ctx->c;
  const StepDims d = t->d;
  const int chi = d.chi, D = d.D;
  size_t nA = (size_t)D * D * D * D, nC = (size_t)chi * chi, nT = (size_t)chi * D * chi;
  DevBuf Cb(c, nC * 8), Tb(c, nT * 8), Cb2(c, nC * 8), Tb2(c, nT * 8);
  k_copy(Cbar_dev, Cb.as<double>(), nC, c->stream);
  k_copy(Tbar_dev, Tb.as<double>(), nT, c->stream);
  k_fill(abar_dev, nA, 0.0, c->stream);
  const int ns = (int)t->steps.size();
  const int first = (t->opts.bwd_last > 0 && t->opts.bwd_last < ns) ? ns - t->opts.bwd_last : 0;
  const Prec bp = to_prec_enum(t->opts.bwd_precision);
  double sweep_ref = 0.0;   // largest resolvent right-hand side of the sweep so far
  for (int i = ns - 1; i >= first; --i) {
    const PBuf& ap = ap_at(c, *t, bp);
    ctmrg_step_back(c, d, bp, *t->steps[i], t->a64.as<double>(), ap, Cb.as<double>(), Tb.as<double>(),
                    Cb2.as<double>(), Tb2.as<double>(), abar_dev, &sweep_ref);
    std::swap(Cb, Cb2);
    std::swap(Tb, Tb2);
  }
  ctx->stats.bwd_steps = ns - first;
}

int tnad_ctmrg_back(tnad_ctx* ctx, tnad_tape* t, const double* Cbar, const double* Tbar, double* abar) {
  return guarded(ctx, [&] {
    if (!t || t->ctx != ctx) throw Error(TNAD_ERR_ARG, "tape does not belong to this context");
    Ctx* c = &ctx->c;
    const StepDims d = t->d;
    size_t nA = (size_t)d.D * d.D * d.D * d.D, nC = (size_t)d.chi * d.chi, nT = (size_t)d.chi * d.D * d.chi;
    DevBuf bC(c, nC * 8), bT(c, nT * 8), out(c, nA * 8);
    copy_in(c, bC.ptr, Cbar, nC * 8);
    copy_in(c, bT.ptr, Tbar, nT * 8);
    ctmrg_back_impl(ctx, t, bC.as<double>(), bT.as<double>(), out.as<double>());
    copy_out(c, abar, out.ptr, nA * 8);
  });
}

int tnad_tape_free(tnad_tape* t) {
  if (!t) return TNAD_ERR_ARG;
  if (t->ctx) (void)hipSetDevice(t->ctx->c.device);
  delete t;
  return TNAD_OK;
}

int tnad_expectationvalue(tnad_ctx* ctx, const double* h, int s, const double* ap, int D, int chi, const double* C,
                          const double* T, double* e) {
  return guarded(ctx, [&] {
    enforce(env_dims_limit(D, chi));
    if (s < 1) throw Error(TNAD_ERR_SHAPE, "physical dimension must be positive");
    Ctx* c = &ctx->c;
    size_t nap = (size_t)D * D * D * D * s * s, nC = (size_t)chi * chi, nT = (size_t)chi * D * chi;
    DevBuf bh(c, (size_t)s * s * s * s * 8), bap(c, nap * 8), bC(c, nC * 8), bT(c, nT * 8);
    copy_in(c, bh.ptr, h, (size_t)s * s * s * s * 8);
    copy_in(c, bap.ptr, ap, nap * 8);
    copy_in(c, bC.ptr, C, nC * 8);
    copy_in(c, bT.ptr, T, nT * 8);
    EvCache ev(c, chi, D, s);
    expectation_fwd(c, chi, D, s, bh.as<double>(), bap.as<double>(), bC.as<double>(), bT.as<double>(), ev, e);
  });
}

int tnad_energy(tnad_ctx* ctx, const double* h, const double* bulk, int d, int s, int chi, double tol, int maxit,
                const tnad_opts* opts, double* e, double* grad, int* iters) {
  return guarded(ctx, [&] {
    if (d < 1 || s < 1) throw Error(TNAD_ERR_SHAPE, "size of tensor error, should be (d, d, d, d, s)");
    const int D = d * d;
    check_dims(D, chi);
    if (maxit < 0) throw Error(TNAD_ERR_ARG, "maxit must be >= 0");
    Ctx* c = &ctx->c;
    hipStream_t st = c->stream;
    const size_t N = (size_t)d * d * d * d * s, nA = (size_t)D * D * D * D, nap = nA * s * s;
    const size_t nC = (size_t)chi * chi, nT = (size_t)chi * D * chi;
    DevBuf bh(c, (size_t)s * s * s * s * 8), bx(c, N * 8), bs(c, N * 8), bap(c, nap * 8);
    copy_in(c, bh.ptr, h, (size_t)s * s * s * s * 8);
    copy_in(c, bx.ptr, bulk, N * 8);
    // ipeps = indexperm_symmetrize(ipeps); ap, a = double layer    (src/variationalipeps.jl:29-34)
    k_symmetrize_bulk(bx.as<double>(), d, s, bs.as<double>(), st);
    tnad_tape t;
    t.ctx = ctx;
    tnad_opts_default(&t.opts);
    if (opts) t.opts = *opts;
    t.d.chi = chi; t.d.D = D; t.d.n = chi * D;
    t.a64 = DevBuf(c, nA * 8);
    k_double_layer(bs.as<double>(), d, s, bap.as<double>(), t.a64.as<double>(), st);
    // rt = SquareCTMRGRuntime(a, Val(:raw), chi); rt = ctmrg(rt; tol, maxit)   (src/variationalipeps.jl:36-37)
    DevBuf C0(c, nC * 8), T0(c, nT * 8), C(c, nC * 8), T(c, nT * 8);
    k_init_raw(t.a64.as<double>(), D, chi, C0.as<double>(), T0.as<double>(), st);
    std::vector<double> v;
    int it = 0;
    run_ctmrg(ctx, t, C0.as<double>(), T0.as<double>(), tol, maxit, grad != nullptr, C.as<double>(), T.as<double>(),
              v, it);
    // e = expectationvalue(h, ap, rt)                                       (src/variationalipeps.jl:38)
    EvCache ev(c, chi, D, s);
    double eh = 0.0;
    expectation_fwd(c, chi, D, s, bh.as<double>(), bap.as<double>(), C.as<double>(), T.as<double>(), ev, &eh);
    if (e) *e = eh;
    if (iters) *iters = it;
    if (grad) {
      c->mark(2);
      DevBuf apbar(c, nap * 8), Cbar(c, nC * 8), Tbar(c, nT * 8), abar(c, nA * 8), bsbar(c, N * 8), g(c, N * 8);
      k_fill(Cbar.as<double>(), nC, 0.0, st);
      k_fill(Tbar.as<double>(), nT, 0.0, st);
      expectation_back(c, chi, D, s, bh.as<double>(), bap.as<double>(), C.as<double>(), T.as<double>(), ev,
                       apbar.as<double>(), Cbar.as<double>(), Tbar.as<double>());
      ctmrg_back_impl(ctx, &t, Cbar.as<double>(), Tbar.as<double>(), abar.as<double>());
      k_double_layer_back(bs.as<double>(), d, s, apbar.as<double>(), abar.as<double>(), bsbar.as<double>(), st);
      k_symmetrize_bulk_back2(bsbar.as<double>(), bx.as<double>(), d, s, g.as<double>(), st);
      c->mark(3);
      copy_out(c, grad, g.ptr, N * 8);
      c->sync();
      ctx->stats.bwd_ms = c->elapsed(2, 3);
    }
  });
}

int tnad_trg(tnad_ctx* ctx, const double* a, int q, int chi, int niter, double tol, double* lnz, double* abar) {
  return guarded(ctx, [&] { trg_run(&ctx->c, a, q, chi, niter, tol, lnz, abar); });
}

int tnad_trg_svd(tnad_ctx* ctx, const double* t, int d1, int d2, int d3, int d4, int dmax, double tol, double* u,
                 double* v, int* k_out) {
  return guarded(ctx, [&] { trg_svd_run(&ctx->c, t, d1, d2, d3, d4, dmax, tol, u, v, k_out); });
}

int tnad_trg_svd_back(tnad_ctx* ctx, const double* t, int d1, int d2, int d3, int d4, int dmax, double tol,
                      const double* ubar, const double* vbar, int k, double* tbar) {
  return guarded(ctx, [&] { trg_svd_back_run(&ctx->c, t, d1, d2, d3, d4, dmax, tol, ubar, vbar, k, tbar); });
}

int tnad_contract(tnad_ctx* ctx, int precision, const char* ia, const long long* da, const double* A, const char* ib,
                  const long long* db, const double* B, const char* ic, const long long* dc, double alpha, double beta,
                  double* Cout) {
  return guarded(ctx, [&] {
    if (!ia || !ib || !ic || !da || !db || !dc) throw Error(TNAD_ERR_ARG, "contract: null argument");
    Ctx* c = &ctx->c;
    auto view = [](const char* idx, const long long* d) {
      std::string s(idx);
      std::vector<long long> dims(d, d + s.size());
      for (long long x : dims)
        if (x < 1) throw Error(TNAD_ERR_SHAPE, "contract: extents must be positive");
      return cm(s, dims);
    };
    TView va = view(ia, da), vb = view(ib, db), vc = view(ic, dc);
    auto count = [](const TView& v) {
      size_t n = 1;
      for (long long x : v.dim) n *= (size_t)x;
      return n;
    };
    const size_t na = count(va), nb = count(vb), ncount = count(vc);
    const Prec p = to_prec_enum(precision);
    TcDesc desc = contract_desc(va, vb, vc, alpha, beta);
    DevBuf a64(c, na * 8), b64(c, nb * 8), c64(c, ncount * 8);
    copy_in(c, a64.ptr, A, na * 8);
    copy_in(c, b64.ptr, B, nb * 8);
    PBuf ap(c, p, na), bp(c, p, nb), cp(c, p, ncount);
    to_prec(c, a64.as<double>(), ap);
    to_prec(c, b64.as<double>(), bp);
    if (beta != 0.0) {
      copy_in(c, c64.ptr, Cout, ncount * 8);
      to_prec(c, c64.as<double>(), cp);
    }
    tc_launch(c, p, desc, ap.ptr(), bp.ptr(), cp.ptr(), c->stream);
    from_prec(c, cp, c64.as<double>());
    copy_out(c, Cout, c64.ptr, ncount * 8);
  });
}

int tnad_time_contract(tnad_ctx* ctx, int precision, const char* ia, const long long* da, const char* ib,
                       const long long* db, const char* ic, const long long* dc, int reps, double* avg_ms) {
  return guarded(ctx, [&] {
    if (!ia || !ib || !ic || !da || !db || !dc || !avg_ms || reps < 1) throw Error(TNAD_ERR_ARG, "time_contract: bad argument");
    Ctx* c = &ctx->c;
    auto view = [](const char* idx, const long long* d) {
      std::string s(idx);
      return cm(s, std::vector<long long>(d, d + s.size()));
    };
    TView va = view(ia, da), vb = view(ib, db), vc = view(ic, dc);
    auto count = [](const TView& v) {
      size_t n = 1;
      for (long long x : v.dim) n *= (size_t)x;
      return n;
    };
    const Prec p = to_prec_enum(precision);
    TcDesc desc = contract_desc(va, vb, vc);
    DevBuf a64(c, count(va) * 8), b64(c, count(vb) * 8);
    k_random_matrix(a64.as<double>(), count(va), 11u, c->stream);
    k_random_matrix(b64.as<double>(), count(vb), 13u, c->stream);
    PBuf ap(c, p, count(va)), bp(c, p, count(vb)), cp(c, p, count(vc));
    to_prec(c, a64.as<double>(), ap);
    to_prec(c, b64.as<double>(), bp);
    tc_launch(c, p, desc, ap.ptr(), bp.ptr(), cp.ptr(), c->stream);   // warm-up launch
    hipEvent_t e0, e1;
    TNAD_HIP_TRY(hipEventCreate(&e0));
    TNAD_HIP_TRY(hipEventCreate(&e1));
    TNAD_HIP_TRY(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; ++r) tc_launch(c, p, desc, ap.ptr(), bp.ptr(), cp.ptr(), c->stream);
    TNAD_HIP_TRY(hipEventRecord(e1, c->stream));
    TNAD_HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    TNAD_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = (double)ms / reps;
  });
}

int tnad_time_jacobi(tnad_ctx* ctx, int m, int sweeps, int reps, double* avg_ms, int* rounds) {
  return guarded(ctx, [&] {
    if (!avg_ms || m < 2 || m > JACOBI_MAX || sweeps < 1 || sweeps > 60 || reps < 1)
      throw Error(TNAD_ERR_ARG, "time_jacobi: bad argument");
    Ctx* c = &ctx->c;
    DevBuf H(c, (size_t)m * m * 8), W(c, jacobi_work_doubles(m, sweeps) * 8), r(c, 16);
    k_random_matrix(H.as<double>(), (size_t)m * m, 17u, c->stream);   // symmetrised by the kernel on load
    *avg_ms = jacobi_rot_time(H.as<double>(), m, sweeps, reps, W.as<double>(), c->stream, r.as<int>());
    if (rounds) {
      int hr = 0;
      c->read_async(&hr, r.ptr, sizeof(int));
      c->sync();
      *rounds = hr;
    }
  });
}

int tnad_time_syev(tnad_ctx* ctx, int m, int reps, double* stage_ms) {
  return guarded(ctx, [&] {
    if (!stage_ms || m < 3 || m > JACOBI_MAX || reps < 1) throw Error(TNAD_ERR_ARG, "time_syev: bad argument");
    Ctx* c = &ctx->c;
    DevBuf H(c, (size_t)m * m * 8), W(c, (syev_dc_work_doubles(m) + (size_t)m * m + m) * 8);
    k_random_matrix(H.as<double>(), (size_t)m * m, 17u, c->stream);   // the upper triangle is read
    float ms[3];
    syev_dc_time(H.as<double>(), m, reps, W.as<double>(), c->stream, ms);
    for (int k = 0; k < 3; ++k) stage_ms[k] = ms[k];
  });
}

}  // extern "C"
