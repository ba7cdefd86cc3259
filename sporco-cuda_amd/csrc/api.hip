// Host driver and C ABI of the CBPDN ADMM solvers (include/sporco_amd.h).
//
// Replaces the reference host drivers cuda_wrapper_CBPDN
// (sporco_cuda/src/cbpdn.cu:86-616) and cuda_wrapper_CBPDN_GR
// (sporco_cuda/src/cbpdn_grd.cu:97-658).  Differences by design:
//  * a workspace (device buffers, FFT plans, twiddle tables) is cached per
//    device and shape instead of 14 allocations and 3 FFT plans per call;
//  * the iteration never waits for the host: residual norms, the stopping
//    test and AutoRho run on the device (k_reduce), and every kernel of a
//    later iteration returns at once after the stopping test has passed; the
//    host only polls the device flag every few iterations to stop queueing;
//  * errors are returned as status codes, the caller's device is restored.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/sporco_amd.h"
#include "fft_gen.h"
#include "scc_internal.h"
#include "scc_pow2.h"

namespace scc {
namespace {
// Message of the last failure on this thread: a fixed buffer, so that
// reporting an error allocates nothing.
thread_local char g_err[512];
}  // namespace

int set_error(int code, const char* what, const char* detail) {
  std::snprintf(g_err, sizeof(g_err), "%s%s%s", what ? what : "",
                detail ? ": " : "", detail ? detail : "");
  return code;
}

int set_error(int code, const std::string& msg) {
  return set_error(code, msg.c_str(), nullptr);
}

// Radix decomposition for the generic Stockham kernels (gen::plan_radices;
// pfa for the column transforms only).
bool make_plan(int L, int tw_stride, FftPlan* pl, bool pfa) {
  pl->L = L;
  pl->tw_stride = tw_stride;
  pl->nstages = gen::plan_radices(L, pl->radix, kMaxStages, pfa);
  return pl->nstages >= 0;
}

int make_geom(int R, int C, int M, Geom* g, bool* pow2, Pow2Config* p2) {
  g->R = R;
  g->C = C;
  g->M = M;
  g->H = C / 2 + 1;
  g->even = (C % 2 == 0) ? 1 : 0;
  g->Lr = g->even ? C / 2 : C;
  g->N = (long long)R * C;
  g->Hc = (long long)R * g->H;
  g->lgR = -1;
  if ((R & (R - 1)) == 0)
    for (g->lgR = 0; (1 << g->lgR) < R; ++g->lgR) {
    }
  *pow2 = pow2_select(R, C, M, p2);
  int TC;
  if (*pow2) {
    TC = p2->TC;
  } else {
    // 32 KiB of column buffers per block: narrower panels than a 64 KiB
    // budget allows, but twice the column blocks per CU (measured faster
    // on uneven sizes, whose column transforms are LDS- and latency-bound).
    TC = gen::col_panel_width(R, g->H, 32768);
  }
  g->TC = TC;
  g->P = (g->H + TC - 1) / TC;
  g->wlast = g->H - (g->P - 1) * TC;
  return SCC_OK;
}

namespace {

thread_local scc_info g_info;
thread_local bool g_hist_on = false;
thread_local std::vector<float> g_hist;
thread_local bool g_obj_on = false;
thread_local std::vector<float> g_obj;  // (Fnc, DFid, Regl1, Regl2g) rows
// Per-stage device timing of the ADMM iterations (events on the solve
// stream around every kernel): index = Stage.
enum Stage { kStColFwd = 0, kStSolve, kStColInv, kStRow, kStReduce, kStRowFwd,
             kNumStages };
thread_local bool g_stage_on = false;
// Graph replay of repeated fixed-iteration solves (scc_set_solve_graph), and
// whether the last solve on this thread ran from a graph.
thread_local bool g_graph_on = true;
thread_local int g_graph_used = 0;
thread_local double g_stage_ms[kNumStages];
thread_local int g_stage_n[kNumStages];

int fail(int code, const char* what, const char* detail = nullptr) {
  return set_error(code, what, detail);
}

#define SCC_HIP(call)                                                   \
  do {                                                                  \
    hipError_t e_ = (call);                                             \
    if (e_ != hipSuccess) {                                             \
      return fail(e_ == hipErrorOutOfMemory ? SCC_ERR_NOMEM : SCC_ERR_HIP, \
                  #call, hipGetErrorString(e_));                        \
    }                                                                   \
  } while (0)

// A device buffer: owned (allocated by ensure, freed with the workspace),
// or a view into memory the library does not own (owned = false).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool owned = true;
  ~DevBuf() {
    if (p && owned) (void)hipFree(p);
  }
};

// Makes b hold at least `bytes`.  A view that is too small is an error: a
// view is never replaced by an allocation.
int ensure(DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes && (b.p || bytes == 0)) return SCC_OK;
  if (!b.owned)
    return fail(SCC_ERR_UNSUPPORTED,
                "buffer needed beyond the fixed workspace of an "
                "allocation-free solve");
  if (b.p) {
    (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
  }
  if (bytes == 0) return SCC_OK;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    char what[64];
    std::snprintf(what, sizeof(what), "hipMalloc of %zu bytes", bytes);
    return fail(SCC_ERR_NOMEM, what, hipGetErrorString(e));
  }
  b.bytes = bytes;
  return SCC_OK;
}

// What a captured iteration loop depends on: the kernels' parameters and
// buffers (caller weight and mask pointers included) and the iteration
// count.  Value-initialised, so that padding bytes compare equal.
struct GraphKey {
  Params p;
  Buffers b;
  int maxit;
  int row_ct;  // compile-time-length row kernels in use
};

// Everything that depends only on (device, R, C, M): buffers, plans, tables.
// The entry points cache one per device; one whose buffers are views into
// memory the library does not own has owned = false (and no stream, events
// or pinned memory of its own).
struct Workspace {
  int device = -1;
  bool owned = true;
  Geom g{};
  FftPlan rowp{}, colp{};
  bool pow2 = false;
  Pow2Config p2{};
  DevBuf Y, U, F, Df, Sf, ghg, partials, blocksum, ctrl, hist, twR, twC;
  DevBuf sD, sS, sL1, sGW, sMask;  // staging for the host-pointer entry
  DevBuf oG, oZ, oE, oPart, oVal, oHist;  // objective evaluation
  DevBuf flag;                            // negative-weight / mask tests
  bool ghg_ready = false;
  hipStream_t stream = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  Ctrl* hctrl = nullptr;  // host copy of the control block (pinned if owned)
  Ctrl hctrl_view{};      // its storage in a workspace of views
  std::mutex mu;          // held for the whole of a solve in this workspace
  // Captured iteration loop of the last repeated fixed-iteration solve, and
  // the configuration of the last solve that was not yet captured.
  hipGraphExec_t gexec = nullptr;
  GraphKey gkey{};
  GraphKey seen{};
  bool seen_valid = false;
  ~Workspace() {
    if (!owned) return;
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (hctrl) (void)hipHostFree(hctrl);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// g_mu guards the cache only; a solve holds its workspace's own lock, so
// solves in different workspaces (devices, shapes) run concurrently.  An
// evicted workspace lives on until the solves holding it return.
std::mutex g_mu;
std::map<std::tuple<int, int, int, int>, std::shared_ptr<Workspace>> g_ws;

// Geometry and FFT plans of (R, C, M) into ws.
int init_geometry(Workspace* ws, int R, int C, int M) {
  make_geom(R, C, M, &ws->g, &ws->pow2, &ws->p2);
  if (!make_plan(ws->g.Lr, ws->g.even ? 2 : 1, &ws->rowp, false) ||
      !make_plan(R, 1, &ws->colp, true))
    return fail(SCC_ERR_UNSUPPORTED, "FFT length has too many factors");
  return SCC_OK;
}

// Number of row-kernel blocks (residual partial sums) of the geometry.
int ws_row_blocks(const Workspace* ws) {
  return ws->pow2 ? pow2_row_blocks(ws->g, ws->p2) : row_blocks(ws->g);
}

int get_workspace(int device, int R, int C, int M,
                  std::shared_ptr<Workspace>* out) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_tuple(device, R, C, M);
  auto it = g_ws.find(key);
  if (it != g_ws.end()) {
    *out = it->second;
    return SCC_OK;
  }
  // Keep at most one workspace per device: shapes of consecutive calls
  // usually repeat, and large shapes hold gigabytes.
  for (auto i = g_ws.begin(); i != g_ws.end();) {
    if (std::get<0>(i->first) == device)
      i = g_ws.erase(i);
    else
      ++i;
  }
  auto ws = std::make_shared<Workspace>();
  ws->device = device;
  int rc;
  if ((rc = init_geometry(ws.get(), R, C, M))) return rc;
  const Geom& g = ws->g;
  const size_t nreal = (size_t)M * g.N;
  const size_t ncplx = (size_t)M * g.Hc;
  // U is allocated by the first solve that needs it (the (Y, U) state or
  // AutoRho), not here.
  if ((rc = ensure(ws->Y, nreal * 4)) || (rc = ensure(ws->F, ncplx * 8)) ||
      (rc = ensure(ws->Df, ncplx * 8)) ||
      (rc = ensure(ws->Sf, (size_t)g.Hc * 8)) ||
      (rc = ensure(ws->ctrl, sizeof(Ctrl))) ||
      (rc = ensure(ws->twR, (size_t)R * 8)) ||
      (rc = ensure(ws->twC, (size_t)C * 8)))
    return rc;
  // Block sums of the first reduction pass (at most 64 blocks, in the
  // reduction kernel or in the power-of-two column kernel).
  if ((rc = ensure(ws->partials, (size_t)5 * ws_row_blocks(ws.get()) * 4)) ||
      (rc = ensure(ws->blocksum, (size_t)5 * 64 * 8)))
    return rc;
  SCC_HIP(hipStreamCreateWithFlags(&ws->stream, hipStreamNonBlocking));
  for (auto& e : ws->ev) SCC_HIP(hipEventCreate(&e));
  SCC_HIP(hipHostMalloc((void**)&ws->hctrl, sizeof(Ctrl)));
  SCC_HIP(launch_twiddles((float2*)ws->twR.p, R, ws->stream));
  SCC_HIP(launch_twiddles((float2*)ws->twC.p, C, ws->stream));
  SCC_HIP(hipStreamSynchronize(ws->stream));
  *out = ws;
  g_ws[key] = std::move(ws);
  return SCC_OK;
}

// GradReg spectrum |F[-1 1]|^2 along rows plus along columns, the closed
// form of the spectra the reference computes by FFT (cbpdn_grd.cu:295-314),
// on the solve stream.
int ensure_ghg(Workspace* ws, hipStream_t st) {
  if (ws->ghg_ready) return SCC_OK;
  int rc = ensure(ws->ghg, (size_t)ws->g.Hc * 4);
  if (rc) return rc;
  SCC_HIP(launch_ghg(ws->g, (float*)ws->ghg.p, st));
  ws->ghg_ready = ws->owned;  // a view is rebuilt on every call
  return SCC_OK;
}

struct Resolved {
  scc_algopt o;
  int l1_mode;   // 0 scalar, 1 per map, 2 full
  size_t n_l1;
  int gw_mode;   // 0 scalar, 1 per map
  bool mask;
  float xi;
};

// Apply defaults to a copy of the options and check them before any HIP
// call, so that invalid input fails the same way with or without a GPU.
int resolve(const scc_algopt* in, float lambda, int grad, Resolved* rs) {
  if (in == nullptr) return fail(SCC_ERR_INVALID, "opt is NULL");
  rs->o = *in;
  scc_default_opts(&rs->o);
  const scc_algopt& o = rs->o;
  const int R = o.IMG_ROW_SIZE, C = o.IMG_COL_SIZE, M = o.DICT_M_SIZE;
  if (R < 1 || C < 1 || M < 1 || o.DICT_ROW_SIZE < 1 || o.DICT_COL_SIZE < 1)
    return fail(SCC_ERR_INVALID, "image and dictionary sizes must be >= 1");
  if (o.DICT_ROW_SIZE > R || o.DICT_COL_SIZE > C)
    return fail(SCC_ERR_INVALID, "filters are larger than the image");
  if ((long long)R * C > (1LL << 30) || (long long)M * R * C > (1LL << 36))
    return fail(SCC_ERR_UNSUPPORTED, "problem too large");
  // LDS bounds of the generic kernels: two padded buffers of one column
  // (gld(R) complex values) must fit in 64 KiB, two of one row (gld(Lr))
  // in 60 KiB.
  if (!gen::sides_fit_lds(R, C))
    return fail(SCC_ERR_UNSUPPORTED, "image side too large");
  if (o.MaxMainIter < 0)
    return fail(SCC_ERR_INVALID, "MaxMainIter must be >= 0");
  if (!(o.rho > 0.f) || !std::isfinite(o.rho))
    return fail(SCC_ERR_INVALID, "rho must be positive and finite");
  if (!std::isfinite(lambda))
    return fail(SCC_ERR_INVALID, "lambda must be finite");
  if (o.AutoRho == 1 && o.AutoRhoPeriod < 1)
    return fail(SCC_ERR_INVALID, "AutoRhoPeriod must be >= 1");
  if (o.L1Weight == nullptr)
    return fail(SCC_ERR_INVALID, "L1Weight is NULL");
  const long long nl1 = (long long)std::max(o.L1_WEIGHT_M_SIZE, 0) *
                        std::max(o.L1_WEIGHT_ROW_SIZE, 0) *
                        std::max(o.L1_WEIGHT_COL_SIZE, 0);
  // Same dispatch order as cbpdn.cu:368-389.
  if (nl1 == M) {
    rs->l1_mode = 1;
  } else if (nl1 == 1) {
    rs->l1_mode = 0;
  } else if (nl1 == (long long)M * R * C) {
    rs->l1_mode = 2;
  } else {
    return fail(SCC_ERR_INVALID,
                "L1Weight must hold 1, M or M*R*C entries");
  }
  rs->n_l1 = (size_t)nl1;
  rs->gw_mode = 0;
  if (grad) {
    if (o.GrdWeight == nullptr)
      return fail(SCC_ERR_INVALID, "GrdWeight is NULL");
    if (o.WEIGHT_SIZE == M && M > 1)
      rs->gw_mode = 1;
    else if (o.WEIGHT_SIZE != 1 && o.WEIGHT_SIZE != M)
      return fail(SCC_ERR_INVALID, "GradWeight must hold 1 or M entries");
  }
  rs->mask = o.nWeight > 1;
  if (rs->mask) {
    if (o.Weight == nullptr)
      return fail(SCC_ERR_INVALID, "Weight (mask) is NULL");
    if ((long long)o.nWeight != (long long)R * C)
      return fail(SCC_ERR_INVALID, "mask must hold R*C entries");
  }
  float xi = o.RhoRsdlTarget;
  if (!(xi > 0.f))
    xi = o.StdResiduals == 1
             ? 1.f
             : (float)(1.0 + std::pow(18.3, std::log10((double)lambda) + 1.0));
  rs->xi = xi;
  return SCC_OK;
}

int check_device(int dev) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    return fail(SCC_ERR_NODEVICE, "no HIP device available");
  if (dev < 0 || dev >= n)
    return fail(SCC_ERR_NODEVICE, "device index out of range");
  return SCC_OK;
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

// 2-D transforms of nmaps maps with the workspace's kernels: forward real
// transform into spectra F, inverse into out times 1/N (F is destroyed).
hipError_t ws_fwd2(Workspace* ws, int nmaps, const float* src, float2* F,
                   hipStream_t st) {
  const Geom& g = ws->g;
  const float2* twR = (const float2*)ws->twR.p;
  const float2* twC = (const float2*)ws->twC.p;
  hipError_t e;
  if (ws->pow2) {
    e = pow2_row_fwd(g, ws->p2, nmaps, kFwdReal, src, nullptr, nullptr, F,
                     twC, nullptr, st);
    if (e == hipSuccess) e = pow2_col(g, ws->p2, nmaps, F, twR, 0, nullptr, st);
  } else {
    e = launch_row_fwd(g, nmaps, ws->rowp, kFwdReal, src, nullptr, nullptr, F,
                       twC, nullptr, st);
    if (e == hipSuccess) e = launch_col(g, nmaps, ws->colp, F, twR, 0, nullptr, st);
  }
  return e;
}

hipError_t ws_inv2(Workspace* ws, int nmaps, float2* F, float* out,
                   hipStream_t st) {
  const Geom& g = ws->g;
  const float2* twR = (const float2*)ws->twR.p;
  const float2* twC = (const float2*)ws->twC.p;
  const float s = 1.f / (float)g.N;
  hipError_t e;
  if (ws->pow2) {
    e = pow2_col(g, ws->p2, nmaps, F, twR, 1, nullptr, st);
    if (e == hipSuccess) e = pow2_row_inv(g, ws->p2, nmaps, F, out, twC, s, st);
  } else {
    e = launch_col(g, nmaps, ws->colp, F, twR, 1, nullptr, st);
    if (e == hipSuccess)
      e = launch_row_inv(g, nmaps, ws->rowp, F, out, twC, s, st);
  }
  return e;
}

// Objective of iteration iter (Fnc, DFid, Regl1, Regl2g) into ws->oVal and
// row iter - 1 of ws->oHist.  With aux off the variable is X, whose
// spectrum was copied to ws->oG right after the solve; with aux on it is Y.
int objective_step(Workspace* ws, const Buffers& b, const Params& p, bool aux,
                   int iter, hipStream_t st) {
  const Geom& g = ws->g;
  const int M = g.M;
  const int nf = obj_blocks(g);
  const int nl = 512, nm = p.m0 ? 256 : 0;
  double* pf = (double*)ws->oPart.p;
  double* pl = pf + 2 * nf;
  double* pm = pl + nl;
  float2* G = (float2*)ws->oG.p;
  float* Z = (float*)ws->oZ.p;
  float2* E = (float2*)ws->oE.p;
  float* e = (float*)(E + g.Hc);
  if (aux) {
    SCC_HIP(launch_obj_l1(g, p, b.Y, b.l1w, b.ctrl, pl, nl, iter, st));
    SCC_HIP(ws_fwd2(ws, M, b.Y, G, st));
    SCC_HIP(launch_obj_freq(g, b, p, G, p.m0 ? E : nullptr, pf, iter, st));
  } else {
    SCC_HIP(launch_obj_freq(g, b, p, G, p.m0 ? E : nullptr, pf, iter, st));
    SCC_HIP(ws_inv2(ws, M, G, Z, st));
    SCC_HIP(launch_obj_l1(g, p, Z, b.l1w, b.ctrl, pl, nl, iter, st));
  }
  if (p.m0) {
    SCC_HIP(ws_inv2(ws, 1, E, e, st));
    SCC_HIP(launch_obj_mask(g, e, b.mask, b.ctrl, pm, nm, iter, st));
  }
  SCC_HIP(launch_obj_final(pf, nf, pl, nl, pm, nm, p, 1.0 / (double)g.N,
                           b.ctrl, (float*)ws->oVal.p, (float*)ws->oHist.p,
                           iter, st));
  return SCC_OK;
}

// The ADMM solve on device-resident inputs.  D (M,Kh,Kw), S (R,C), weights
// and mask are device pointers; Y (M,R,C) receives the result on the device.
int solve_core(Workspace* ws, const Resolved& rs, const float* D,
               const float* S, float lambda, float mu, int grad,
               const float* l1w, const float* gw, const int* mask, float* Yout,
               hipStream_t st) {
  const scc_algopt& o = rs.o;
  const Geom& g = ws->g;
  const int M = g.M;
  g_graph_used = 0;
  Params p{};
  p.lambda = lambda;
  p.mu = mu;
  p.alpha = o.RelaxParam;
  p.abs_tol = o.AbsStopTol;
  p.rel_tol = o.RelStopTol;
  p.tau = o.RhoScaling;
  p.rsdl_ratio = o.RhoRsdlRatio;
  p.xi = rs.xi;
  p.sqrt_nc = std::sqrt((double)M * (double)g.N);
  p.std_residuals = o.StdResiduals == 1;
  p.autorho = o.AutoRho == 1;
  p.autoscaling = o.AutoRhoScaling == 1;
  p.period = std::max(o.AutoRhoPeriod, 1);
  p.nonneg = o.NonNegCoef == 1;
  p.nobndry = o.NoBndryCross == 1;
  p.Kh = o.DICT_ROW_SIZE;
  p.Kw = o.DICT_COL_SIZE;
  p.m0 = rs.mask ? 1 : 0;
  p.l1_mode = rs.l1_mode;
  p.grad = grad ? 1 : 0;
  p.gw_mode = rs.gw_mode;
  // A workspace of views records no history or objective (their buffers
  // would need allocation).
  p.track = (g_hist_on && ws->owned) ? 1 : 0;
  p.l1_scalar = 1.f;
  p.gw_scalar = 1.f;
  if (rs.l1_mode == 0)
    SCC_HIP(hipMemcpyAsync(&p.l1_scalar, l1w, 4, hipMemcpyDeviceToHost, st));
  if (grad && rs.gw_mode == 0)
    SCC_HIP(hipMemcpyAsync(&p.gw_scalar, gw, 4, hipMemcpyDeviceToHost, st));
  SCC_HIP(hipStreamSynchronize(st));
  int rc;
  if (grad && (rc = ensure_ghg(ws, st))) return rc;
  const int maxit = o.MaxMainIter;
  if (p.track && (rc = ensure(ws->hist, (size_t)3 * std::max(maxit, 1) * 4)))
    return rc;
  // Objective evaluation after every iteration (Verbose, or requested).
  const bool obj_on = (g_obj_on && ws->owned) || o.Verbose == 1;
  const bool aux = o.AuxVarObj == 1;
  // The state is V (one real array per coefficient) unless the objective
  // of every iteration needs Y, or a threshold is negative (shrink(0) != 0:
  // V cannot hold Y = U = 0).
  bool thr_ok = lambda >= 0.f && (rs.l1_mode != 0 || p.l1_scalar >= 0.f);
  if (!obj_on && thr_ok && rs.l1_mode != 0) {
    if ((rc = ensure(ws->flag, 4))) return rc;
    const long long nw = rs.l1_mode == 1 ? (long long)M : (long long)M * g.N;
    int neg = 0;
    SCC_HIP(launch_any_negative(l1w, nw, (int*)ws->flag.p, st));
    SCC_HIP(hipMemcpyAsync(&neg, ws->flag.p, 4, hipMemcpyDeviceToHost, st));
    SCC_HIP(hipStreamSynchronize(st));
    thr_ok = neg == 0;
  }
  p.vstate = (!obj_on && thr_ok) ? 1 : 0;
  // With AutoRho the V state decodes the AMS map after a rho update as
  // (1 - W) (Y + U / rsf), which is (Y, U / rsf) only where W is 0 or 1:
  // another mask value keeps the (Y, U) state.
  if (p.vstate && p.autorho && rs.mask) {
    if ((rc = ensure(ws->flag, 4))) return rc;
    int bad = 0;
    SCC_HIP(launch_any_nonbinary(mask, g.N, (int*)ws->flag.p, st));
    SCC_HIP(hipMemcpyAsync(&bad, ws->flag.p, 4, hipMemcpyDeviceToHost, st));
    SCC_HIP(hipStreamSynchronize(st));
    if (bad) p.vstate = 0;
  }
  // U is read only by the (Y, U) state and written by AutoRho steps.
  const bool need_u = !p.vstate || p.autorho;
  if (need_u && (rc = ensure(ws->U, (size_t)M * g.N * 4))) return rc;
  if (obj_on &&
      ((rc = ensure(ws->oG, (size_t)M * g.Hc * 8)) ||
       (rc = ensure(ws->oZ, aux ? 4 : (size_t)M * g.N * 4)) ||
       (rc = ensure(ws->oE, rs.mask ? (size_t)g.Hc * 8 + (size_t)g.N * 4
                                     : 8)) ||
       (rc = ensure(ws->oPart, ((size_t)2 * obj_blocks(g) + 512 + 256) * 8)) ||
       (rc = ensure(ws->oVal, 4 * 4)) ||
       (rc = ensure(ws->oHist, (size_t)4 * std::max(maxit, 1) * 4))))
    return rc;

  Buffers b{};
  b.Y = (float*)ws->Y.p;
  b.U = need_u ? (float*)ws->U.p : nullptr;
  b.F = (float2*)ws->F.p;
  b.Df = (float2*)ws->Df.p;
  b.Sf = (float2*)ws->Sf.p;
  b.ghg = (float*)ws->ghg.p;
  b.l1w = l1w;
  b.gw = gw;
  b.mask = mask;
  b.partials = (float*)ws->partials.p;
  b.blocksum = (double*)ws->blocksum.p;
  b.ctrl = (Ctrl*)ws->ctrl.p;
  b.hist = p.track ? (float*)ws->hist.p : nullptr;
  b.twR = (float2*)ws->twR.p;
  b.twC = (float2*)ws->twC.p;

  const size_t nreal = (size_t)M * g.N;
  const size_t ncplx = (size_t)M * g.Hc;
  if (ws->ev[0]) SCC_HIP(hipEventRecord(ws->ev[0], st));
  // Setup (cbpdn.cu:276-298): spectra of the padded filters and the image.
  SCC_HIP(launch_pad(D, b.Y, M, o.DICT_ROW_SIZE, o.DICT_COL_SIZE, g.R, g.C,
                     st));
  if (ws->pow2) {
    SCC_HIP(pow2_row_fwd(g, ws->p2, M, kFwdReal, b.Y, nullptr, nullptr, b.Df,
                         b.twC, nullptr, st));
    SCC_HIP(pow2_col(g, ws->p2, M, b.Df, b.twR, 0, nullptr, st));
    SCC_HIP(pow2_row_fwd(g, ws->p2, 1, kFwdReal, S, nullptr, nullptr, b.Sf,
                         b.twC, nullptr, st));
    SCC_HIP(pow2_col(g, ws->p2, 1, b.Sf, b.twR, 0, nullptr, st));
  } else {
    SCC_HIP(launch_row_fwd(g, M, ws->rowp, kFwdReal, b.Y, nullptr, nullptr,
                           b.Df, b.twC, nullptr, st));
    SCC_HIP(launch_col(g, M, ws->colp, b.Df, b.twR, 0, nullptr, st));
    SCC_HIP(launch_row_fwd(g, 1, ws->rowp, kFwdReal, S, nullptr, nullptr,
                           b.Sf, b.twC, nullptr, st));
    SCC_HIP(launch_col(g, 1, ws->colp, b.Sf, b.twR, 0, nullptr, st));
  }
  // Y = U = 0 and A-hat = 0 for the first iteration (cbpdn.cu:321-326).
  SCC_HIP(launch_zero(b.Y, nreal * 4, st));
  if (need_u) SCC_HIP(launch_zero(b.U, nreal * 4, st));
  SCC_HIP(launch_zero(b.F, ncplx * 8, st));
  Ctrl c0{};
  c0.rho = o.rho;
  c0.rsf = 1.f;
  *ws->hctrl = c0;
  SCC_HIP(hipMemcpyAsync(b.ctrl, ws->hctrl, sizeof(Ctrl),
                         hipMemcpyHostToDevice, st));
  if (ws->ev[1]) SCC_HIP(hipEventRecord(ws->ev[1], st));

  // Stage timing: one event before the first stage of an iteration and one
  // after every stage.
  std::vector<hipEvent_t> sev;
  std::vector<int> sid;
  for (int k = 0; k < kNumStages; ++k) {
    g_stage_ms[k] = 0.0;
    g_stage_n[k] = 0;
  }
  const bool stage_on = g_stage_on && ws->owned;
  auto mark = [&](int stage) -> hipError_t {
    if (!stage_on) return hipSuccess;
    hipEvent_t e;
    hipError_t err = hipEventCreate(&e);
    if (err != hipSuccess) return err;
    sev.push_back(e);
    sid.push_back(stage);
    return hipEventRecord(e, st);
  };
  struct EvFree {
    std::vector<hipEvent_t>* v;
    ~EvFree() {
      for (auto e : *v) (void)hipEventDestroy(e);
    }
  } evfree{&sev};

  // A stopping test can pass only with a positive tolerance.
  const bool can_stop = o.RelStopTol > 0.f || o.AbsStopTol > 0.f;
  const int poll = o.Verbose == 1 ? 1 : 16;
  int nblocks = 0;
  // With the register solve (power-of-two sizes, or M <= 256) the residual
  // reduction of an iteration that is not an AutoRho step, a host poll or
  // the last one is deferred into the next iteration (first pass in the
  // column kernel's blocks, second pass and control update in block 0 of
  // the solve, before the inverse transforms and the row kernel test the
  // stopping flag), saving two launches.
  bool pending = false;
  auto loop = [&]() -> int {
  for (int i = 1; i <= maxit; ++i) {
    const bool rho_step = p.autorho && i > 1 && i % p.period == 0;
    const int rmode = rho_step ? kRowUpdate : kRowFused;
    const bool poll_pt =
        (can_stop || o.Verbose == 1) && (i % poll == 0 || i == maxit);
    const bool defer = (ws->pow2 || solve_regs_fits(M)) && !rho_step &&
                       !poll_pt && i < maxit;
    SCC_HIP(mark(-1));
    if (ws->pow2) {
      SCC_HIP(pow2_col(g, ws->p2, M, b.F, b.twR, 0, b.ctrl, st,
                       pending ? b.partials : nullptr, pending ? nblocks : 0,
                       b.blocksum));
      SCC_HIP(mark(kStColFwd));
      SCC_HIP(pow2_solve(g, ws->p2, b, p, st, 1, pending ? i - 1 : 0,
                         pending ? pow2_defer_blocks(g, M) : 0));
      SCC_HIP(mark(kStSolve));
      if (obj_on && !aux)  // spectrum of X for the objective
        SCC_HIP(hipMemcpyAsync(ws->oG.p, b.F, (size_t)M * g.Hc * 8,
                               hipMemcpyDeviceToDevice, st));
      SCC_HIP(pow2_col(g, ws->p2, M, b.F, b.twR, 1, b.ctrl, st));
      SCC_HIP(mark(kStColInv));
      SCC_HIP(pow2_row(g, ws->p2, b, p, rmode, &nblocks, st));
    } else {
      SCC_HIP(launch_col(g, M, ws->colp, b.F, b.twR, 0, b.ctrl, st,
                         pending ? b.partials : nullptr, pending ? nblocks : 0,
                         b.blocksum));
      SCC_HIP(mark(kStColFwd));
      SCC_HIP(launch_solve(g, b, p, st, pending ? i - 1 : 0,
                           pending ? pow2_defer_blocks(g, M) : 0));
      SCC_HIP(mark(kStSolve));
      if (obj_on && !aux)  // spectrum of X for the objective
        SCC_HIP(hipMemcpyAsync(ws->oG.p, b.F, (size_t)M * g.Hc * 8,
                               hipMemcpyDeviceToDevice, st));
      SCC_HIP(launch_col(g, M, ws->colp, b.F, b.twR, 1, b.ctrl, st));
      SCC_HIP(mark(kStColInv));
      SCC_HIP(launch_row(g, ws->rowp, b, p, rmode, &nblocks, st));
    }
    SCC_HIP(mark(kStRow));
    pending = defer;
    if (!defer) {
      SCC_HIP(launch_reduce(b, p, nblocks, i, maxit, st));
      SCC_HIP(mark(kStReduce));
    }
    if (rho_step) {
      const int fm = p.vstate ? kFwdYUtoV : kFwdYU;
      if (ws->pow2)
        SCC_HIP(pow2_row_fwd(g, ws->p2, M, fm, nullptr, b.Y, b.U, b.F, b.twC,
                             b.ctrl, st, i));
      else
        SCC_HIP(launch_row_fwd(g, M, ws->rowp, fm, nullptr, b.Y, b.U, b.F,
                               b.twC, b.ctrl, st, i));
      SCC_HIP(mark(kStRowFwd));
    }
    if (obj_on && (rc = objective_step(ws, b, p, aux, i, st))) return rc;
    if ((can_stop || o.Verbose == 1) && (i % poll == 0 || i == maxit)) {
      SCC_HIP(hipMemcpyAsync(ws->hctrl, b.ctrl, sizeof(Ctrl),
                             hipMemcpyDeviceToHost, st));
      float ov[4] = {0.f, 0.f, 0.f, 0.f};
      if (o.Verbose == 1)
        SCC_HIP(hipMemcpyAsync(ov, ws->oVal.p, sizeof(ov),
                               hipMemcpyDeviceToHost, st));
      SCC_HIP(hipStreamSynchronize(st));
      if (o.Verbose == 1) {
        if (i == 1) {
          std::printf("Itn   Fnc       DFid      Regl1     %sr         s"
                      "         rho\n", grad ? "Regl2grd  " : "");
          std::printf("%s\n", grad ? "-----------------------------------"
                                     "-----------------------------------"
                                   : "-----------------------------------"
                                     "-------------------------");
        }
        if (!ws->hctrl->done || ws->hctrl->iter == i) {
          std::printf("%4d %9.2e %9.2e %9.2e ", ws->hctrl->iter, ov[0],
                      ov[1], ov[2]);
          if (grad) std::printf("%9.2e ", ov[3]);
          std::printf("%9.2e %9.2e %9.2e\n", ws->hctrl->r, ws->hctrl->s,
                      ws->hctrl->rho);
        }
      }
      if (ws->hctrl->done) break;
    }
  }
  return SCC_OK;
  };
  // Fixed-iteration solves (no host poll, objective or stage timing) run
  // their iterations from a captured HIP graph once the same configuration
  // (parameters, buffers, iteration count) has been solved before: the
  // second such call captures and instantiates it, later calls only replay
  // it, which removes the per-launch host cost (4 launches per iteration).
  // The graph holds kernel nodes only (no memset or copy nodes).
  const bool graphable = g_graph_on && ws->owned && !stage_on && !obj_on &&
                         !can_stop && o.Verbose != 1 && maxit > 0;
  if (graphable) {
    GraphKey key;
    std::memset(&key, 0, sizeof(key));
    std::memcpy(&key.p, &p, sizeof(p));
    std::memcpy(&key.b, &b, sizeof(b));
    key.maxit = maxit;
    key.row_ct = row_ct_enabled() ? 1 : 0;
    if (ws->gexec != nullptr && std::memcmp(&key, &ws->gkey, sizeof(key)) ==
                                    0) {
      SCC_HIP(hipGraphLaunch(ws->gexec, st));
      g_graph_used = 1;
    } else if (ws->seen_valid &&
               std::memcmp(&key, &ws->seen, sizeof(key)) == 0) {
      SCC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      rc = loop();
      hipGraph_t graph = nullptr;
      hipError_t ce = hipStreamEndCapture(st, &graph);
      if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      if (ce != hipSuccess)
        return fail(SCC_ERR_HIP, "hipStreamEndCapture", hipGetErrorString(ce));
      hipGraphExec_t exec = nullptr;
      hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess)
        return fail(SCC_ERR_HIP, "hipGraphInstantiate", hipGetErrorString(ie));
      if (ws->gexec) (void)hipGraphExecDestroy(ws->gexec);
      ws->gexec = exec;
      ws->gkey = key;
      SCC_HIP(hipGraphLaunch(ws->gexec, st));
      g_graph_used = 1;
    } else {
      ws->seen = key;
      ws->seen_valid = true;
      if ((rc = loop())) return rc;
    }
  } else {
    if ((rc = loop())) return rc;
  }
  if (ws->ev[2]) SCC_HIP(hipEventRecord(ws->ev[2], st));
  if (p.vstate)  // Y = shrink(V), into the output or in place
    SCC_HIP(launch_state_to_y(g, b, p, Yout != nullptr ? Yout : b.Y, st));
  else if (Yout != nullptr)
    SCC_HIP(hipMemcpyAsync(Yout, b.Y, nreal * 4, hipMemcpyDeviceToDevice, st));
  SCC_HIP(hipMemcpyAsync(ws->hctrl, b.ctrl, sizeof(Ctrl),
                         hipMemcpyDeviceToHost, st));
  SCC_HIP(hipStreamSynchronize(st));
  for (size_t k = 1; k < sev.size(); ++k) {
    if (sid[k] < 0) continue;
    float ms = 0.f;
    SCC_HIP(hipEventElapsedTime(&ms, sev[k - 1], sev[k]));
    g_stage_ms[sid[k]] += ms;
    g_stage_n[sid[k]] += 1;
  }
  const Ctrl& cf = *ws->hctrl;
  g_info = scc_info{};
  g_info.iterations = cf.iter;
  g_info.rho = cf.rho;
  g_info.r = cf.r;
  g_info.s = cf.s;
  g_info.epri = cf.epri;
  g_info.edua = cf.edua;
  float ms = 0.f;
  if (ws->ev[0] &&
      hipEventElapsedTime(&ms, ws->ev[0], ws->ev[1]) == hipSuccess)
    g_info.setup_ms = ms;
  if (ws->ev[1] &&
      hipEventElapsedTime(&ms, ws->ev[1], ws->ev[2]) == hipSuccess)
    g_info.solve_ms = ms;
  if (obj_on) {
    g_obj.resize((size_t)4 * cf.iter);
    if (cf.iter > 0)
      SCC_HIP(hipMemcpy(g_obj.data(), ws->oHist.p, g_obj.size() * 4,
                        hipMemcpyDeviceToHost));
  } else {
    g_obj.clear();
  }
  if (p.track) {
    g_hist.resize((size_t)3 * cf.iter);
    if (cf.iter > 0)
      SCC_HIP(hipMemcpy(g_hist.data(), b.hist, g_hist.size() * 4,
                        hipMemcpyDeviceToHost));
  } else {
    g_hist.clear();
  }
  return SCC_OK;
}

int solve_host(const float* D, const float* S, float lambda, float mu,
               int grad, const scc_algopt* opt, float* Y) {
  if (D == nullptr || S == nullptr || Y == nullptr)
    return fail(SCC_ERR_INVALID, "NULL array argument");
  Resolved rs;
  int rc = resolve(opt, lambda, grad, &rs);
  if (rc) return rc;
  const scc_algopt& o = rs.o;
  if ((rc = check_device(o.device))) return rc;
  DeviceGuard dg(o.device);
  std::shared_ptr<Workspace> wsp;
  const int R = o.IMG_ROW_SIZE, C = o.IMG_COL_SIZE, M = o.DICT_M_SIZE;
  if ((rc = get_workspace(o.device, R, C, M, &wsp))) return rc;
  Workspace* ws = wsp.get();
  std::lock_guard<std::mutex> lk(ws->mu);
  const size_t nd = (size_t)M * o.DICT_ROW_SIZE * o.DICT_COL_SIZE;
  const size_t nl1 = rs.n_l1;
  const size_t ngw = grad ? (size_t)std::max(o.WEIGHT_SIZE, 1) : 0;
  if ((rc = ensure(ws->sD, nd * 4)) || (rc = ensure(ws->sS, (size_t)R * C * 4)) ||
      (rc = ensure(ws->sL1, nl1 * 4)) || (rc = ensure(ws->sGW, ngw * 4)) ||
      (rc = ensure(ws->sMask, rs.mask ? (size_t)R * C * 4 : 0)))
    return rc;
  hipStream_t st = ws->stream;
  SCC_HIP(hipMemcpyAsync(ws->sD.p, D, nd * 4, hipMemcpyHostToDevice, st));
  SCC_HIP(hipMemcpyAsync(ws->sS.p, S, (size_t)R * C * 4,
                         hipMemcpyHostToDevice, st));
  SCC_HIP(hipMemcpyAsync(ws->sL1.p, o.L1Weight, nl1 * 4,
                         hipMemcpyHostToDevice, st));
  if (grad)
    SCC_HIP(hipMemcpyAsync(ws->sGW.p, o.GrdWeight, ngw * 4,
                           hipMemcpyHostToDevice, st));
  if (rs.mask)
    SCC_HIP(hipMemcpyAsync(ws->sMask.p, o.Weight, (size_t)R * C * 4,
                           hipMemcpyHostToDevice, st));
  rc = solve_core(ws, rs, (const float*)ws->sD.p, (const float*)ws->sS.p,
                  lambda, mu, grad, (const float*)ws->sL1.p,
                  (const float*)ws->sGW.p, (const int*)ws->sMask.p, nullptr,
                  st);
  if (rc) return rc;
  // Result return (cbpdn.cu:575): Y straight from the workspace.
  SCC_HIP(hipMemcpy(Y, ws->Y.p, (size_t)M * R * C * 4,
                    hipMemcpyDeviceToHost));
  return SCC_OK;
}

// --------------------------------------------------------------------------
// Allocation-free solves (scc_cbpdn_device_noalloc).  Their workspace lives
// in `scratch`, a file-scope device array: each GPU holds its own copy from
// the first call until the process ends.  In host code the array's name is
// only a stand-in for the symbol; the device address of each GPU's copy
// comes from hipGetSymbolAddress with that GPU current, and only that
// address is handed to kernels.
// --------------------------------------------------------------------------
constexpr size_t kScratchBytes = (size_t)SCC_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices];  // per device, once looked up
Workspace g_nws[kMaxDevices];               // per device workspace of views

// Device address of this device's scratch array (dev must be current).
int scratch_base(int dev, unsigned char** base) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
      return fail(SCC_ERR_HIP, "hipGetSymbolAddress(scratch)",
                  e != hipSuccess ? hipGetErrorString(e) : "null address");
    g_scratch_ptr[dev] = (unsigned char*)ptr;
  }
  *base = g_scratch_ptr[dev];
  return SCC_OK;
}

// Lays the buffers of one solve out from base (256-byte aligned each) and
// returns the bytes they take; base == nullptr only measures.  Y (the
// solver state and the result) is the caller's output array.
size_t carve_views(Workspace* ws, unsigned char* base, float* Y) {
  const Geom& g = ws->g;
  const size_t nreal = (size_t)g.M * g.N;
  const size_t ncplx = (size_t)g.M * g.Hc;
  size_t off = 0;
  auto take = [&](DevBuf& b, size_t bytes) {
    b.owned = false;
    b.p = base ? base + off : nullptr;
    b.bytes = bytes;
    off += (bytes + 255) & ~(size_t)255;
  };
  auto none = [](DevBuf& b) {
    b.owned = false;
    b.p = nullptr;
    b.bytes = 0;
  };
  for (DevBuf* b : {&ws->hist, &ws->sD, &ws->sS, &ws->sL1, &ws->sGW,
                    &ws->sMask, &ws->oG, &ws->oZ, &ws->oE, &ws->oPart,
                    &ws->oVal, &ws->oHist})
    none(*b);
  ws->Y.owned = false;
  ws->Y.p = Y;
  ws->Y.bytes = nreal * 4;
  take(ws->F, ncplx * 8);
  take(ws->Df, ncplx * 8);
  take(ws->U, nreal * 4);
  take(ws->Sf, (size_t)g.Hc * 8);
  take(ws->ghg, (size_t)g.Hc * 4);
  take(ws->twR, (size_t)g.R * 8);
  take(ws->twC, (size_t)g.C * 8);
  take(ws->partials, (size_t)5 * ws_row_blocks(ws) * 4);
  take(ws->blocksum, (size_t)5 * 64 * 8);
  take(ws->ctrl, sizeof(Ctrl));
  take(ws->flag, 4);
  return off;
}

int solve_noalloc(const float* D, const float* S, float lambda, float mu,
                  int grad, const scc_algopt* opt, float* Y, void* stream) {
  if (D == nullptr || S == nullptr || Y == nullptr)
    return fail(SCC_ERR_INVALID, "NULL array argument");
  Resolved rs;
  int rc = resolve(opt, lambda, grad, &rs);
  if (rc) return rc;
  const scc_algopt& o = rs.o;
  if (o.Verbose == 1)
    return fail(SCC_ERR_UNSUPPORTED,
                "Verbose needs per-iteration objective buffers, which the "
                "allocation-free solve does not have");
  {
    // Host-side size check first: no device work for a problem that does
    // not fit the array.
    Workspace probe;
    probe.owned = false;
    if ((rc = init_geometry(&probe, o.IMG_ROW_SIZE, o.IMG_COL_SIZE,
                            o.DICT_M_SIZE)))
      return rc;
    const size_t need = carve_views(&probe, nullptr, Y);
    if (need > kScratchBytes) {
      char what[160];
      std::snprintf(what, sizeof(what),
                    "allocation-free solve needs %zu bytes of workspace; the "
                    "fixed scratch array holds %zu",
                    need, kScratchBytes);
      return fail(SCC_ERR_UNSUPPORTED, what);
    }
  }
  if ((rc = check_device(o.device))) return rc;
  if (o.device >= kMaxDevices)
    return fail(SCC_ERR_UNSUPPORTED, "device index beyond the scratch table");
  Workspace* ws = &g_nws[o.device];
  std::lock_guard<std::mutex> lk(ws->mu);
  ws->owned = false;
  ws->device = o.device;
  ws->ghg_ready = false;
  ws->stream = nullptr;
  ws->hctrl = &ws->hctrl_view;
  if ((rc = init_geometry(ws, o.IMG_ROW_SIZE, o.IMG_COL_SIZE, o.DICT_M_SIZE)))
    return rc;
  DeviceGuard dg(o.device);
  unsigned char* base = nullptr;
  if ((rc = scratch_base(o.device, &base))) return rc;
  carve_views(ws, base, Y);
  hipStream_t st = (hipStream_t)stream;
  // The tables of the previous call's shape may still be in the array.
  SCC_HIP(launch_twiddles((float2*)ws->twR.p, ws->g.R, st));
  SCC_HIP(launch_twiddles((float2*)ws->twC.p, ws->g.C, st));
  return solve_core(ws, rs, D, S, lambda, mu, grad, o.L1Weight, o.GrdWeight,
                    o.Weight, nullptr, st);
}

}  // namespace
}  // namespace scc

using namespace scc;

extern "C" {

void scc_clear_opts(scc_algopt* opt) {
  if (opt == nullptr) return;
  std::memset(opt, 0, sizeof(*opt));
  opt->device = -1;
  opt->rho = -1;
  opt->Verbose = -1;
  opt->AutoRho = -1;
  opt->nWeight = -1;
  opt->AuxVarObj = -1;
  opt->AbsStopTol = -1;
  opt->NonNegCoef = -1;
  opt->RelaxParam = -1;
  opt->RhoScaling = -1;
  opt->RelStopTol = -1;
  opt->MaxMainIter = -1;
  opt->NoBndryCross = -1;
  opt->HighMemSolve = -1;
  opt->StdResiduals = -1;
  opt->RhoRsdlRatio = -1;
  opt->RhoRsdlTarget = -1;
  opt->AutoRhoPeriod = -1;
  opt->AutoRhoScaling = -1;
  opt->WEIGHT_SIZE = 1;
  opt->L1_WEIGHT_M_SIZE = 1;
  opt->L1_WEIGHT_ROW_SIZE = 1;
  opt->L1_WEIGHT_COL_SIZE = 1;
}

void scc_default_opts(scc_algopt* opt) {
  if (opt == nullptr) return;
  // Defaults of default_opts (sporco_cuda/src/utils.cu:12-64).
  if (opt->device == -1) opt->device = 0;
  if (opt->rho == -1) opt->rho = 0;
  if (opt->Verbose == -1) opt->Verbose = 0;
  if (opt->AutoRho == -1) opt->AutoRho = 1;
  if (opt->nWeight == -1) opt->nWeight = 1;
  if (opt->AuxVarObj == -1) opt->AuxVarObj = 0;
  if (opt->AbsStopTol == -1) opt->AbsStopTol = 0;
  if (opt->NonNegCoef == -1) opt->NonNegCoef = 0;
  if (opt->RelaxParam == -1) opt->RelaxParam = 1.8f;
  if (opt->RhoScaling == -1) opt->RhoScaling = 100;
  if (opt->RelStopTol == -1) opt->RelStopTol = 1e-4f;
  if (opt->MaxMainIter == -1) opt->MaxMainIter = 250;
  if (opt->NoBndryCross == -1) opt->NoBndryCross = 0;
  if (opt->HighMemSolve == -1) opt->HighMemSolve = 0;
  if (opt->StdResiduals == -1) opt->StdResiduals = 0;
  if (opt->RhoRsdlRatio == -1) opt->RhoRsdlRatio = 1.2f;
  if (opt->AutoRhoPeriod == -1) opt->AutoRhoPeriod = 1;
  if (opt->AutoRhoScaling == -1) opt->AutoRhoScaling = 1;
}

int scc_cbpdn(const float* D, const float* S, float lambda,
              const scc_algopt* opt, float* Y) {
  return solve_host(D, S, lambda, 0.f, 0, opt, Y);
}

int scc_cbpdn_grd(const float* D, const float* S, float lambda, float mu,
                  const scc_algopt* opt, float* Y) {
  return solve_host(D, S, lambda, mu, 1, opt, Y);
}

int scc_cbpdn_device(const float* D, const float* S, float lambda, float mu,
                     int grad, const scc_algopt* opt, float* Y,
                     void* stream) {
  if (D == nullptr || S == nullptr || Y == nullptr)
    return fail(SCC_ERR_INVALID, "NULL array argument");
  Resolved rs;
  int rc = resolve(opt, lambda, grad, &rs);
  if (rc) return rc;
  const scc_algopt& o = rs.o;
  if ((rc = check_device(o.device))) return rc;
  DeviceGuard dg(o.device);
  std::shared_ptr<Workspace> wsp;
  if ((rc = get_workspace(o.device, o.IMG_ROW_SIZE, o.IMG_COL_SIZE,
                          o.DICT_M_SIZE, &wsp)))
    return rc;
  Workspace* ws = wsp.get();
  std::lock_guard<std::mutex> lk(ws->mu);
  hipStream_t st = stream ? (hipStream_t)stream : ws->stream;
  return solve_core(ws, rs, D, S, lambda, mu, grad, o.L1Weight, o.GrdWeight,
                    o.Weight, Y, st);
}

int scc_cbpdn_device_noalloc(const float* D, const float* S, float lambda,
                             float mu, int grad, const scc_algopt* opt,
                             float* Y, void* stream) {
  return solve_noalloc(D, S, lambda, mu, grad, opt, Y, stream);
}

int scc_last_info(scc_info* info) {
  if (info == nullptr) return SCC_ERR_INVALID;
  *info = g_info;
  return SCC_OK;
}

const char* scc_last_error(void) { return g_err; }

void scc_set_history(int enabled) { g_hist_on = enabled != 0; }

void scc_set_objective(int enabled) { g_obj_on = enabled != 0; }

int scc_last_objective(float* rows, int cap) {
  const int n = (int)(g_obj.size() / 4);
  if (rows != nullptr && cap > 0)
    std::memcpy(rows, g_obj.data(), (size_t)std::min(n, cap) * 4 * 4);
  return n;
}

int scc_last_history(float* rsr, int cap) {
  const int n = (int)(g_hist.size() / 3);
  if (rsr != nullptr && cap > 0)
    std::memcpy(rsr, g_hist.data(), (size_t)std::min(n, cap) * 3 * 4);
  return n;
}

void scc_set_stage_timing(int enabled) { g_stage_on = enabled != 0; }

void scc_set_fixed_kernels(int on) { set_row_ct(on != 0); }

void scc_set_solve_graph(int enabled) { g_graph_on = enabled != 0; }

int scc_last_solve_graph(void) { return g_graph_used; }

int scc_last_stage_times(double* ms_total, int* launches, int n) {
  for (int k = 0; k < n && k < kNumStages; ++k) {
    if (ms_total) ms_total[k] = g_stage_ms[k];
    if (launches) launches[k] = g_stage_n[k];
  }
  return kNumStages;
}

int scc_get_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int scc_get_current_device(int* dev) {
  if (dev == nullptr) return SCC_ERR_INVALID;
  hipError_t e = hipGetDevice(dev);
  if (e != hipSuccess) return fail(SCC_ERR_NODEVICE, hipGetErrorString(e));
  return SCC_OK;
}

int scc_set_current_device(int dev) {
  int rc = check_device(dev);
  if (rc) return rc;
  hipError_t e = hipSetDevice(dev);
  if (e != hipSuccess) return fail(SCC_ERR_HIP, hipGetErrorString(e));
  return SCC_OK;
}

int scc_get_memory_info(size_t* free_bytes, size_t* total_bytes) {
  if (free_bytes == nullptr || total_bytes == nullptr) return SCC_ERR_INVALID;
  int rc = check_device(0);
  if (rc) return rc;
  hipError_t e = hipMemGetInfo(free_bytes, total_bytes);
  if (e != hipSuccess) return fail(SCC_ERR_HIP, hipGetErrorString(e));
  return SCC_OK;
}

const char* scc_get_device_name(int dev) {
  static thread_local hipDeviceProp_t prop;
  if (check_device(dev)) return nullptr;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return nullptr;
  // Some driver stacks leave the marketing name empty; the architecture
  // name still identifies the device.
  if (prop.name[0] == '\0') return prop.gcnArchName;
  return prop.name;
}

void scc_release_cache(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_ws.clear();
}

}  // extern "C"
