// Batched plain CBPDN over images of uneven sizes sharing one dictionary.
//
// Every stage of an ADMM iteration is one launch whose blocks walk a work
// list spanning all images: (panel, map) pairs for the column transforms,
// groups of rows for the row kernel, groups of frequencies for the X-solve.
// Item -> image is a binary search in a prefix table of item counts; each
// image carries its own geometry, FFT plans, twiddles, buffers, lambda and
// control block, so the images keep their own residuals, stopping test and
// AutoRho (the loop of the reference, cbpdn.cu:318-567, per image).  The
// stage bodies are those of the single-image generic kernels.
//
// The launch sequence of all iterations is captured once in a HIP graph and
// replayed; per call only the image descriptors (input / output pointers,
// lambda) and the control blocks are uploaded.  The first call with a new
// (device, M, sizes) signature times candidate panel widths, rows per row
// block, lanes per frequency of the solve and grid shapes on the batch
// itself and keeps the fastest.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sporco_amd.h"
#include "generic_common.h"
#include "scc_pow2.h"

namespace scc {

// Per-image descriptor (device resident).
struct BImg {
  Geom g;
  FftPlan rp, cp;
  int nr;          // rows per row-kernel block
  int nrow;        // row-kernel items of the M maps = partial sums
  float lambda;
  float rho0;      // initial rho
  float xi;        // AutoRho residual target
  double sqrt_nc;  // sqrt(M R C)
  const float* S;
  float* Y;
  float* U;
  float2* F;
  float2* Df;
  float2* Sf;
  float* partials;
  Ctrl* ctrl;
  float2* twR;
  float2* twC;
};

// One work list: item counts of every image as a prefix table.
struct BList {
  const BImg* im;
  const int* pre;  // nimg + 1 entries, pre[0] = 0
  int nimg;
  int total;
};

enum BSel { kSelDict = 0, kSelImage = 1, kSelWork = 2 };

// Image i with pre[i] <= item < pre[i + 1] (every image has >= 1 item).
__device__ inline int find_img(const int* pre, int n, int item) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pre[mid] <= item)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Image of a work item, made wavefront-uniform so that the descriptor is
// read with scalar loads.
__device__ __forceinline__ int img_of(const BList& L, int item) {
  return __builtin_amdgcn_readfirstlane(find_img(L.pre, L.nimg, item));
}

// Y, U (M maps) and the work spectra F of every image to zero; grid.y
// indexes the image.
__global__ void k_bzero(const BImg* ims) {
  const BImg& im = ims[blockIdx.y];
  const long long nr = (long long)im.g.M * im.g.N;
  const long long nc = (long long)im.g.M * im.g.Hc;
  const long long s = (long long)gridDim.x * blockDim.x;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nr;
       i += s) {
    im.Y[i] = 0.f;
    im.U[i] = 0.f;
  }
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nc;
       i += s)
    im.F[i] = make_float2(0.f, 0.f);
}

// Filters zero-padded at the origin (cbpdn_kernels.cu:602-615) into the Y
// buffer of every image, as the input of the dictionary spectra.
__global__ void k_bpad(const BImg* ims, const float* __restrict__ D, int Kh,
                       int Kw) {
  const BImg& im = ims[blockIdx.y];
  const int R = im.g.R, C = im.g.C;
  const long long RC = (long long)R * C;
  const long long n = (long long)im.g.M * RC;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long m = i / RC;
    const long long rc = i - m * RC;
    const int r = (int)(rc / C);
    const int c = (int)(rc - (long long)r * C);
    im.Y[i] = (r < Kh && c < Kw) ? D[(m * Kh + r) * Kw + c] : 0.f;
  }
}

// Forward row transforms: padded filters -> Df, images -> Sf, or (AutoRho)
// Y - U/rsf -> F.
template <bool WAVE>
__global__ __launch_bounds__(256) void k_brow_fwd(BList L, int sel) {
  extern __shared__ float2 lds[];
  for (int item = blockIdx.x; item < L.total; item += gridDim.x) {
    const int i = img_of(L, item);
    const BImg& im = L.im[i];
    // Geometry by value: the stores below would otherwise force its fields
    // to be reloaded inside the transform loops.  The plan is read once per
    // stage and stays in memory (a local copy indexed by stage would live in
    // scratch).
    const Geom g = geom_uni(im.g);
    const FftPlan& rp = im.rp;
    const long long blk = uni(item - L.pre[i]);
    const long long mr = (long long)g.M * g.R;
    const int nr = uni(im.nr);
    const float2* twC = uni(im.twC);
    auto body = [&](long long nrows, int mode, const float* src, float* Y,
                    float* U, float2* F, float rsf) {
      if constexpr (WAVE)
        row_fwd_body_wave(g, nrows, rp, mode, src, Y, U, F,
                          row_wave_table(g, lds, twC), rsf, nr, blk, lds);
      else
        row_fwd_body(g, nrows, rp, mode, src, Y, U, F, twC, rsf, nr, blk,
                     lds);
    };
    if (sel == kSelWork) {
      if (im.ctrl->done) continue;
      body(mr, kFwdYU, nullptr, uni(im.Y), uni(im.U), uni(im.F),
           uni(im.ctrl->rsf));
    } else if (sel == kSelDict) {
      body(mr, kFwdReal, uni(im.Y), nullptr, nullptr, uni(im.Df), 1.f);
    } else {
      body(g.R, kFwdReal, uni(im.S), nullptr, nullptr, uni(im.Sf), 1.f);
    }
    __syncthreads();  // LDS is reused by the next item
  }
}

// Column transforms of (panel, map) items, in place.
__global__ __launch_bounds__(256) void k_bcol(BList L, int sel, float dir) {
  extern __shared__ float2 lds[];
  for (int item = blockIdx.x; item < L.total; item += gridDim.x) {
    const int i = img_of(L, item);
    const BImg& im = L.im[i];
    if (sel == kSelWork && im.ctrl->done) continue;
    const Geom g = geom_uni(im.g);
    const FftPlan& cp = im.cp;
    const int local = uni(item - L.pre[i]);
    const int m = local / g.P;
    const int p = local - m * g.P;
    float2* F =
        uni(sel == kSelDict ? im.Df : (sel == kSelImage ? im.Sf : im.F));
    const float2* twR = uni(im.twR);
    col_body(g, cp, F, twR, dir, p, m, lds);
    __syncthreads();
  }
}

// Plain Sherman-Morrison X-solve (cbpdn_kernels.cu:15-117), G lanes per
// frequency holding MPT maps each in registers (maps gi, gi + G, ...):
//   b_m = conj(D_m) S + rho A_m,
//   x_m = (b_m - conj(D_m) [sum_j D_j b_j / rho] / [1 + sum_j |D_j|^2 / rho]) / rho.
// The sums over maps are shuffles across the G lanes of a frequency.
template <int G, int MPT>
__global__ __launch_bounds__(256) void k_bsolve(BList L) {
  constexpr int FPW = 64 / G;
  constexpr int FPB = 4 * FPW;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int gi = lane / FPW;
  const int fl = lane - gi * FPW;
  for (int item = blockIdx.x; item < L.total; item += gridDim.x) {
    const int i = img_of(L, item);
    const BImg& im = L.im[i];
    if (im.ctrl->done) continue;
    const long long Hc = im.g.Hc;
    const int M = im.g.M;
    const float2* __restrict__ Dfi = im.Df;
    const float2* __restrict__ Sfi = im.Sf;
    float2* Fi = im.F;
    const long long f = (long long)(item - L.pre[i]) * FPB + wv * FPW + fl;
    const bool valid = f < Hc;
    const float rho = im.ctrl->rho;
    const float irho = 1.f / rho;
    const float2 sh = valid ? Sfi[f] : make_float2(0.f, 0.f);
    float2 bm[MPT], dm[MPT];
    float2 num = make_float2(0.f, 0.f);
    float den = 0.f;
#pragma unroll
    for (int k = 0; k < MPT; ++k) {
      const int m = gi + G * k;
      bm[k] = make_float2(0.f, 0.f);
      dm[k] = make_float2(0.f, 0.f);
      if (valid && m < M) {
        const long long o = (long long)m * Hc + f;
        dm[k] = Dfi[o];
        const float2 a = Fi[o];
        bm[k] = cadd(cmul(cconj(dm[k]), sh), cscale(a, rho));
        num = cadd(num, cmul(dm[k], bm[k]));
        den += dm[k].x * dm[k].x + dm[k].y * dm[k].y;
      }
    }
#pragma unroll
    for (int o = FPW; o < 64; o <<= 1) {
      num.x += __shfl_xor(num.x, o, 64);
      num.y += __shfl_xor(num.y, o, 64);
      den += __shfl_xor(den, o, 64);
    }
    const float2 cb = cscale(num, irho / (1.f + den * irho));
#pragma unroll
    for (int k = 0; k < MPT; ++k) {
      const int m = gi + G * k;
      if (valid && m < M)
        Fi[(long long)m * Hc + f] =
            cscale(csub(bm[k], cmul(cconj(dm[k]), cb)), irho);
    }
  }
}

// Row kernel items: inverse row FFT, relax + shrink + dual update, residual
// partial sums, and in fused mode the forward row FFT of Y - U.
template <bool WAVE>
__global__ __launch_bounds__(256) void k_brow(BList L, Params p,
                                              const float* l1w, int mode) {
  extern __shared__ float2 lds[];
  __shared__ float red[5][4];
  for (int item = blockIdx.x; item < L.total; item += gridDim.x) {
    const int i = img_of(L, item);
    const BImg& im = L.im[i];
    if (im.ctrl->done) continue;
    const Geom g = geom_uni(im.g);
    const FftPlan& rp = im.rp;
    Buffers b{};
    b.Y = im.Y;
    b.U = im.U;
    b.F = im.F;
    b.l1w = l1w;
    b.partials = im.partials;
    b.ctrl = im.ctrl;
    b.twC = im.twC;
    Params q = p;
    q.lambda = im.lambda;
    if constexpr (WAVE)
      row_body_wave(g, rp, b, q, mode, im.nr, item - L.pre[i], im.nrow, lds,
                    row_wave_table(g, lds, im.twC), red);
    else
      row_body(g, rp, b, q, mode, im.nr, item - L.pre[i], im.nrow, lds, red);
    __syncthreads();
  }
}

// Residual norms, stopping test and AutoRho of every image: one block per
// image sums its partials in double in a fixed order, then runs the control
// logic of the single-image solver on its control block.
__global__ __launch_bounds__(256) void k_breduce(const BImg* ims, Params p,
                                                 int iter) {
  const BImg& im = ims[blockIdx.x];
  if (im.ctrl->done) return;
  __shared__ double red[5][4];
  const int n = im.nrow;
  double acc[5] = {0, 0, 0, 0, 0};
  for (int k = threadIdx.x; k < n; k += blockDim.x)
#pragma unroll
    for (int j = 0; j < 5; ++j) acc[j] += (double)im.partials[j * n + k];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    double v = acc[j];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) red[j][wv] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s5[5];
    for (int j = 0; j < 5; ++j)
      s5[j] = ((red[j][0] + red[j][1]) + red[j][2]) + red[j][3];
    Buffers b{};
    b.ctrl = im.ctrl;
    Params q = p;
    q.lambda = im.lambda;
    q.xi = im.xi;
    q.sqrt_nc = im.sqrt_nc;
    q.track = 0;
    control_update(b, q, s5, iter);
  }
}

// --------------------------------------------------------------------------
// Host driver
// --------------------------------------------------------------------------
namespace {

struct BBuf {
  void* p = nullptr;
  size_t n = 0;
  ~BBuf() {
    if (p) (void)hipFree(p);
  }
  int get(size_t bytes) {
    if (p && n >= bytes) return SCC_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (bytes == 0) return SCC_OK;
    if (hipMalloc(&p, bytes) != hipSuccess)
      return set_error(SCC_ERR_NOMEM, "hipMalloc of " +
                                          std::to_string(bytes) + " bytes");
    n = bytes;
    return SCC_OK;
  }
};

#define BHIP(call)                                                         \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess)                                                  \
      return set_error(                                                    \
          e_ == hipErrorOutOfMemory ? SCC_ERR_NOMEM : SCC_ERR_HIP,         \
          std::string(#call) + ": " + hipGetErrorString(e_));              \
  } while (0)

struct BCfg {
  int col_lds = 65536;
  int row_lds = 61440;  // block-wide row kernels
  int rpw = 2;          // rows per wavefront, wavefront-per-row kernels
  int mpt = 8;
  int persistent = 1;
  int branches = 0;     // 0: work-list kernels; k > 0: per-image chains
};

constexpr int kMaxBranches = 4;

// Per-image chains mode: every image runs the single-image kernels of its
// own size (power-of-two or generic, with the single-image geometry and
// panel layout) in a chain of launches; the chains of all images are
// captured on `branches` streams of one graph, so that the images of
// different branches overlap on the device.
struct BChain {
  Geom g;
  Pow2Config p2;
  bool pow2;
  int nblk;  // row-kernel blocks = residual partial sums
};

// Work lists: 0 row items of M maps, 1 row items of the image, 2 column
// items of M maps, 3 column items of the image, 4 solve items.
enum { kLRowM = 0, kLRow1, kLColM, kLCol1, kLSolve, kNumLists };

struct BWs {
  int device = -1;
  int M = 0;
  std::vector<int> sizes;  // R0, C0, R1, C1, ...
  int nimg = 0;
  BCfg cfg;
  BCfg tcfg;  // tuned (or default) configuration of this signature
  bool tuned = false;
  std::vector<BImg> h;
  std::vector<int> hpre;  // kNumLists x (nimg + 1)
  int total[kNumLists] = {};
  int lds_row = 0, lds_col = 0;
  bool wave = false;  // wavefront-per-row kernels (every row fits)
  int grid[kNumLists] = {};
  int G = 1;
  BBuf U, F, Df, Sf, part, ctrl, tw, desc, pre, l1w, bsum, flag;
  bool vs = false;  // chains keep the state as V (every threshold >= 0)
  // Chains mode: the iterations run on the state array Yw (same offsets as
  // U), turned into the caller's outputs Y afterwards, so that a captured
  // graph holds no per-call pointer.
  BBuf Yw;
  std::vector<BChain> ch;
  hipStream_t br[kMaxBranches] = {};
  hipEvent_t bev[kMaxBranches + 1] = {};
  std::vector<size_t> tw_off;
  hipStream_t cap = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<unsigned char> gkey;
  Ctrl* hctrl = nullptr;
  BImg* hdesc = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  ~BWs() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (cap) (void)hipStreamDestroy(cap);
    for (auto& s : br)
      if (s) (void)hipStreamDestroy(s);
    for (auto& e : bev)
      if (e) (void)hipEventDestroy(e);
    if (hctrl) (void)hipHostFree(hctrl);
    if (hdesc) (void)hipHostFree(hdesc);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

std::mutex b_mu;
std::unique_ptr<BWs> b_ws;
bool b_autotune = true;
bool b_graph = true;
int b_branches = -1;  // < 0: tuned mode; 0: work lists; k > 0: k chains
thread_local scc_batch_config b_last{};

BList list_of(const BWs* w, int k) {
  BList L;
  L.im = (const BImg*)w->desc.p;
  L.pre = (const int*)w->pre.p + (size_t)k * (w->nimg + 1);
  L.nimg = w->nimg;
  L.total = w->total[k];
  return L;
}

int solve_lanes(int M, int mpt) {
  int G = 1;
  while (G * mpt < M) G *= 2;
  return G;
}

int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// Geometry, plans, work lists, LDS sizes and grids of configuration c;
// uploads descriptors (with the current per-call fields) and prefix tables.
int configure(BWs* w, const BCfg& c, hipStream_t st) {
  w->cfg = c;
  w->G = solve_lanes(w->M, c.mpt);
  const int n = w->nimg;
  w->hpre.assign((size_t)kNumLists * (n + 1), 0);
  w->lds_row = 0;
  w->lds_col = 0;
  w->wave = true;
  for (int i = 0; i < n; ++i) w->wave = w->wave && row_wave_fits(w->h[i].g);
  const int fpb = 4 * (64 / w->G);
  for (int i = 0; i < n; ++i) {
    BImg& im = w->h[i];
    Geom& g = im.g;
    const int R = g.R;
    const int TC = gen::col_panel_width(R, g.H, c.col_lds);
    g.TC = TC;
    g.P = (g.H + TC - 1) / TC;
    g.wlast = g.H - (g.P - 1) * TC;
    im.nr = w->wave ? kRowWaves * c.rpw : row_nr_budget(g, c.row_lds);
    im.nrow = ceil_div((long long)w->M * R, im.nr);
    w->lds_row = std::max(w->lds_row,
                          w->wave ? row_wave_lds(g)
                                  : 2 * im.nr * gen::gld(g.Lr) * 8);
    w->lds_col = std::max(w->lds_col, 2 * TC * gen::gld(R) * 8);
    const int cnt[kNumLists] = {im.nrow, ceil_div(R, im.nr), g.P * w->M, g.P,
                                ceil_div(g.Hc, fpb)};
    for (int k = 0; k < kNumLists; ++k) {
      int* pre = w->hpre.data() + (size_t)k * (n + 1);
      pre[i + 1] = pre[i] + cnt[k];
    }
  }
  for (int k = 0; k < kNumLists; ++k)
    w->total[k] = w->hpre[(size_t)k * (n + 1) + n];
  // Grids: one block per item, or as many as stay resident.
  int ncu = 0;
  BHIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                             w->device));
  auto occ = [&](const void* fn, int lds) {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, fn, 256, lds) !=
            hipSuccess ||
        b < 1)
      b = 1;
    return b;
  };
  const int occ_row =
      occ(w->wave ? (const void*)k_brow<true> : (const void*)k_brow<false>,
          w->lds_row);
  const int occ_col = occ((const void*)k_bcol, w->lds_col);
  for (int k = 0; k < kNumLists; ++k) {
    int g = w->total[k];
    if (c.persistent) {
      const int o = (k == kLColM || k == kLCol1) ? occ_col
                    : (k == kLSolve)             ? 8
                                                 : occ_row;
      g = std::min(g, o * ncu);
    }
    w->grid[k] = std::max(g, 1);
  }
  BHIP(hipMemcpyAsync(w->pre.p, w->hpre.data(), w->hpre.size() * sizeof(int),
                      hipMemcpyHostToDevice, st));
  std::memcpy(w->hdesc, w->h.data(), (size_t)n * sizeof(BImg));
  BHIP(hipMemcpyAsync(w->desc.p, w->hdesc, (size_t)n * sizeof(BImg),
                      hipMemcpyHostToDevice, st));
  BHIP(hipStreamSynchronize(st));
  return SCC_OK;
}

int get_bws(int device, int M, const std::vector<int>& sizes, BWs** out) {
  if (b_ws && b_ws->device == device && b_ws->M == M && b_ws->sizes == sizes) {
    *out = b_ws.get();
    return SCC_OK;
  }
  b_ws.reset();
  auto w = std::make_unique<BWs>();
  w->device = device;
  w->M = M;
  w->sizes = sizes;
  const int n = (int)sizes.size() / 2;
  w->nimg = n;
  w->h.assign(n, BImg{});
  size_t nreal = 0, ncplx = 0, nimgc = 0, npart = 0, ntw = 0;
  std::vector<size_t> oreal(n), ocplx(n), oimgc(n), opart(n), otw(n);
  auto al = [](size_t x) { return (x + 63) / 64 * 64; };
  for (int i = 0; i < n; ++i) {
    BImg& im = w->h[i];
    Geom& g = im.g;
    g.R = sizes[2 * i];
    g.C = sizes[2 * i + 1];
    g.M = M;
    g.H = g.C / 2 + 1;
    g.even = (g.C % 2 == 0) ? 1 : 0;
    g.Lr = g.even ? g.C / 2 : g.C;
    g.N = (long long)g.R * g.C;
    g.Hc = (long long)g.R * g.H;
    g.lgR = -1;
    if (!make_plan(g.Lr, g.even ? 2 : 1, &im.rp, false) ||
        !make_plan(g.R, 1, &im.cp, true))
      return set_error(SCC_ERR_UNSUPPORTED, "FFT length has too many factors");
    im.sqrt_nc = std::sqrt((double)M * (double)g.N);
    oreal[i] = nreal;
    nreal += al((size_t)M * g.N);
    ocplx[i] = ncplx;
    ncplx += al((size_t)M * g.Hc);
    oimgc[i] = nimgc;
    nimgc += al((size_t)g.Hc);
    opart[i] = npart;
    npart += al((size_t)5 * M * g.R);  // enough for one row per block
    otw[i] = ntw;
    ntw += al((size_t)g.R + g.C);
  }
  int rc;
  if ((rc = w->U.get(nreal * 4)) || (rc = w->F.get(ncplx * 8)) ||
      (rc = w->Df.get(ncplx * 8)) || (rc = w->Sf.get(nimgc * 8)) ||
      (rc = w->part.get(npart * 4)) ||
      (rc = w->ctrl.get((size_t)n * sizeof(Ctrl))) ||
      (rc = w->tw.get(ntw * 8)) || (rc = w->desc.get((size_t)n * sizeof(BImg))) ||
      (rc = w->pre.get((size_t)kNumLists * (n + 1) * sizeof(int))) ||
      (rc = w->l1w.get((size_t)M * 4)) ||
      (rc = w->bsum.get((size_t)n * 5 * 64 * 8)))
    return rc;
  w->ch.resize(n);
  for (int i = 0; i < n; ++i) {
    BChain& c = w->ch[i];
    make_geom(sizes[2 * i], sizes[2 * i + 1], M, &c.g, &c.pow2, &c.p2);
    c.nblk = c.pow2 ? pow2_row_blocks(c.g, c.p2) : row_blocks(c.g);
  }
  for (auto& s2 : w->br)
    BHIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
  for (auto& e : w->bev)
    BHIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  BHIP(hipHostMalloc((void**)&w->hctrl, (size_t)n * sizeof(Ctrl)));
  BHIP(hipHostMalloc((void**)&w->hdesc, (size_t)n * sizeof(BImg)));
  BHIP(hipStreamCreateWithFlags(&w->cap, hipStreamNonBlocking));
  for (auto& e : w->ev) BHIP(hipEventCreate(&e));
  // Twiddle tables W_L^t = exp(-2 pi i t / L) in double, R then C per image.
  std::vector<float2> tw(ntw);
  for (int i = 0; i < n; ++i) {
    BImg& im = w->h[i];
    for (int pass = 0; pass < 2; ++pass) {
      const int L = pass == 0 ? im.g.R : im.g.C;
      float2* t = tw.data() + otw[i] + (pass == 0 ? 0 : im.g.R);
      for (int k = 0; k < L; ++k) {
        const double a = -2.0 * M_PI * (double)k / (double)L;
        t[k] = make_float2((float)std::cos(a), (float)std::sin(a));
      }
    }
    im.U = (float*)w->U.p + oreal[i];
    im.F = (float2*)w->F.p + ocplx[i];
    im.Df = (float2*)w->Df.p + ocplx[i];
    im.Sf = (float2*)w->Sf.p + oimgc[i];
    im.partials = (float*)w->part.p + opart[i];
    im.ctrl = (Ctrl*)w->ctrl.p + i;
    im.twR = (float2*)w->tw.p + otw[i];
    im.twC = (float2*)w->tw.p + otw[i] + im.g.R;
  }
  BHIP(hipMemcpy(w->tw.p, tw.data(), ntw * 8, hipMemcpyHostToDevice));
  *out = w.get();
  b_ws = std::move(w);
  return SCC_OK;
}

void launch_solve_b(BWs* w, hipStream_t st) {
  const BList L = list_of(w, kLSolve);
  const dim3 grid(w->grid[kLSolve]), blk(256);
#define X(g, t)                                                           \
  if (w->G == g && w->cfg.mpt == t) {                                     \
    hipLaunchKernelGGL((k_bsolve<g, t>), grid, blk, 0, st, L);            \
    return;                                                               \
  }
  X(1, 4) X(2, 4) X(4, 4) X(8, 4) X(16, 4) X(32, 4) X(64, 4)
  X(1, 8) X(2, 8) X(4, 8) X(8, 8) X(16, 8) X(32, 8) X(64, 8)
  X(1, 16) X(2, 16) X(4, 16) X(8, 16) X(16, 16) X(32, 16) X(64, 16)
#undef X
}

void launch_col_b(BWs* w, int list, int sel, float dir, hipStream_t st) {
  hipLaunchKernelGGL(k_bcol, dim3(w->grid[list]), dim3(256), w->lds_col, st,
                     list_of(w, list), sel, dir);
}

void launch_row_fwd_b(BWs* w, int list, int sel, hipStream_t st) {
  if (w->wave)
    hipLaunchKernelGGL(k_brow_fwd<true>, dim3(w->grid[list]), dim3(256),
                       w->lds_row, st, list_of(w, list), sel);
  else
    hipLaunchKernelGGL(k_brow_fwd<false>, dim3(w->grid[list]), dim3(256),
                       w->lds_row, st, list_of(w, list), sel);
}

void launch_row_b(BWs* w, const Params& p, int mode, hipStream_t st) {
  if (w->wave)
    hipLaunchKernelGGL(k_brow<true>, dim3(w->grid[kLRowM]), dim3(256),
                       w->lds_row, st, list_of(w, kLRowM), p,
                       (const float*)w->l1w.p, mode);
  else
    hipLaunchKernelGGL(k_brow<false>, dim3(w->grid[kLRowM]), dim3(256),
                       w->lds_row, st, list_of(w, kLRowM), p,
                       (const float*)w->l1w.p, mode);
}

// One ADMM iteration of every image (the single-image sequence of
// solve_core: columns, solve, inverse columns, row kernel, reduction, and
// on an AutoRho step the row transforms of Y - U/rsf).
int launch_iter(BWs* w, const Params& p, int iter, hipStream_t st) {
  const bool rho_step = p.autorho && iter > 1 && iter % p.period == 0;
  launch_col_b(w, kLColM, kSelWork, 1.f, st);
  launch_solve_b(w, st);
  launch_col_b(w, kLColM, kSelWork, -1.f, st);
  launch_row_b(w, p, rho_step ? kRowUpdate : kRowFused, st);
  hipLaunchKernelGGL(k_breduce, dim3(w->nimg), dim3(256), 0, st,
                     (const BImg*)w->desc.p, p, iter);
  if (rho_step) launch_row_fwd_b(w, kLRowM, kSelWork, st);
  BHIP(hipGetLastError());
  return SCC_OK;
}

void launch_zero(BWs* w, hipStream_t st) {
  hipLaunchKernelGGL(k_bzero, dim3(512, w->nimg), dim3(256), 0, st,
                     (const BImg*)w->desc.p);
}

// Control blocks: rho0 of every image (or `rho` when positive), rsf 1.
int upload_ctrl(BWs* w, float rho, hipStream_t st) {
  for (int i = 0; i < w->nimg; ++i) {
    w->hctrl[i] = Ctrl{};
    w->hctrl[i].rho = rho > 0.f ? rho : w->h[i].rho0;
    w->hctrl[i].rsf = 1.f;
  }
  BHIP(hipMemcpyAsync(w->ctrl.p, w->hctrl, (size_t)w->nimg * sizeof(Ctrl),
                      hipMemcpyHostToDevice, st));
  return SCC_OK;
}

// ---------------------------------------------------------------------------
// Per-image chains (BCfg::branches > 0)
// ---------------------------------------------------------------------------
Buffers chain_buffers(const BWs* w, int j) {
  const BImg& im = w->h[j];
  Buffers b{};
  b.Y = im.Y;
  b.U = im.U;
  b.F = im.F;
  b.Df = im.Df;
  b.Sf = im.Sf;
  b.l1w = (const float*)w->l1w.p;
  b.partials = im.partials;
  b.blocksum = (double*)w->bsum.p + (size_t)j * 5 * 64;
  b.ctrl = (Ctrl*)w->ctrl.p + j;
  b.twR = im.twR;
  b.twC = im.twC;
  return b;
}

// The chains keep the state as V (Params::vstate) when every threshold is
// >= 0.
Params chain_params(const BWs* w, int j, const Params& p0) {
  Params p = p0;
  p.lambda = w->h[j].lambda;
  p.xi = w->h[j].xi;
  p.sqrt_nc = w->h[j].sqrt_nc;
  p.vstate = w->vs ? 1 : 0;
  return p;
}

// Spectra of the padded filters (held in the image's Y) and of the image.
int chain_setup(BWs* w, int j, hipStream_t st) {
  const BChain& c = w->ch[j];
  const BImg& im = w->h[j];
  const Geom& g = c.g;
  const Buffers b = chain_buffers(w, j);
  if (c.pow2) {
    BHIP(pow2_row_fwd(g, c.p2, g.M, kFwdReal, b.Y, nullptr, nullptr, b.Df,
                      b.twC, nullptr, st));
    BHIP(pow2_col(g, c.p2, g.M, b.Df, b.twR, 0, nullptr, st));
    BHIP(pow2_row_fwd(g, c.p2, 1, kFwdReal, im.S, nullptr, nullptr, b.Sf,
                      b.twC, nullptr, st));
    BHIP(pow2_col(g, c.p2, 1, b.Sf, b.twR, 0, nullptr, st));
  } else {
    BHIP(launch_row_fwd(g, g.M, im.rp, kFwdReal, b.Y, nullptr, nullptr, b.Df,
                        b.twC, nullptr, st));
    BHIP(launch_col(g, g.M, im.cp, b.Df, b.twR, 0, nullptr, st));
    BHIP(launch_row_fwd(g, 1, im.rp, kFwdReal, im.S, nullptr, nullptr, b.Sf,
                        b.twC, nullptr, st));
    BHIP(launch_col(g, 1, im.cp, b.Sf, b.twR, 0, nullptr, st));
  }
  return SCC_OK;
}

// Iteration i of image j: the single-image sequence (solve_core of the
// single-image solver), with the power-of-two path's deferred residual
// reduction on every iteration that is neither an AutoRho step nor the
// last one.
bool chain_defers(const BChain& c, const Params& p, int i, int maxit) {
  const bool rho_step = p.autorho && i > 1 && i % p.period == 0;
  return (c.pow2 || solve_regs_fits(c.g.M)) && !rho_step && i < maxit;
}

int chain_iter(BWs* w, int j, const Params& p0, int i, int maxit,
               hipStream_t st) {
  const BChain& c = w->ch[j];
  const BImg& im = w->h[j];
  const Geom& g = c.g;
  const Buffers b = chain_buffers(w, j);
  const Params p = chain_params(w, j, p0);
  const bool rho_step = p.autorho && i > 1 && i % p.period == 0;
  const int rmode = rho_step ? kRowUpdate : kRowFused;
  const bool pending = i > 1 && chain_defers(c, p, i - 1, maxit);
  int nblocks = c.nblk;
  if (c.pow2) {
    BHIP(pow2_col(g, c.p2, g.M, b.F, b.twR, 0, b.ctrl, st,
                  pending ? b.partials : nullptr, pending ? c.nblk : 0,
                  b.blocksum));
    BHIP(pow2_solve(g, c.p2, b, p, st, 1, pending ? i - 1 : 0,
                    pending ? pow2_defer_blocks(g, g.M) : 0));
    BHIP(pow2_col(g, c.p2, g.M, b.F, b.twR, 1, b.ctrl, st));
    BHIP(pow2_row(g, c.p2, b, p, rmode, &nblocks, st));
  } else {
    BHIP(launch_col(g, g.M, im.cp, b.F, b.twR, 0, b.ctrl, st,
                    pending ? b.partials : nullptr, pending ? c.nblk : 0,
                    b.blocksum));
    BHIP(launch_solve(g, b, p, st, pending ? i - 1 : 0,
                      pending ? pow2_defer_blocks(g, g.M) : 0));
    BHIP(launch_col(g, g.M, im.cp, b.F, b.twR, 1, b.ctrl, st));
    BHIP(launch_row(g, im.rp, b, p, rmode, &nblocks, st));
  }
  if (!chain_defers(c, p, i, maxit))
    BHIP(launch_reduce(b, p, nblocks, i, maxit, st));
  if (rho_step) {
    if (c.pow2)
      BHIP(pow2_row_fwd(g, c.p2, g.M, kFwdYUtoV, nullptr, b.Y, b.U, b.F,
                        b.twC, b.ctrl, st, i));
    else
      BHIP(launch_row_fwd(g, g.M, im.rp, kFwdYUtoV, nullptr, b.Y, b.U, b.F,
                          b.twC, b.ctrl, st, i));
  }
  return SCC_OK;
}

// Iterations i0 .. i1 (of maxit) of every image: image j on branch
// j mod branches, forked from and joined to st.
int launch_chains(BWs* w, const Params& p, int i0, int i1, int maxit,
                  hipStream_t st) {
  const int K = std::min(std::max(w->cfg.branches, 1), kMaxBranches);
  int rc;
  if (K == 1) {
    for (int j = 0; j < w->nimg; ++j)
      for (int i = i0; i <= i1; ++i)
        if ((rc = chain_iter(w, j, p, i, maxit, st))) return rc;
    return SCC_OK;
  }
  BHIP(hipEventRecord(w->bev[kMaxBranches], st));
  for (int k = 0; k < K; ++k)
    BHIP(hipStreamWaitEvent(w->br[k], w->bev[kMaxBranches], 0));
  for (int j = 0; j < w->nimg; ++j)
    for (int i = i0; i <= i1; ++i)
      if ((rc = chain_iter(w, j, p, i, maxit, w->br[j % K]))) return rc;
  for (int k = 0; k < K; ++k) {
    BHIP(hipEventRecord(w->bev[k], w->br[k]));
    BHIP(hipStreamWaitEvent(st, w->bev[k], 0));
  }
  return SCC_OK;
}

// Average device time of `reps` runs of fn after one warm-up run.
template <class Fn>
int time_it(BWs* w, hipStream_t st, int reps, Fn fn, float* ms) {
  fn();
  BHIP(hipEventRecord(w->ev[0], st));
  for (int r = 0; r < reps; ++r) fn();
  BHIP(hipEventRecord(w->ev[1], st));
  BHIP(hipEventSynchronize(w->ev[1]));
  BHIP(hipGetLastError());
  float t = 0.f;
  BHIP(hipEventElapsedTime(&t, w->ev[0], w->ev[1]));
  *ms = t / reps;
  return SCC_OK;
}

// Stage-by-stage search on the batch's own buffers (zero data, rho = 1):
// panel width by the column transforms, rows per block by the row kernel,
// lanes per frequency by the solve, then the grid shape on a whole
// iteration.  The image descriptors must hold valid Y pointers.
int autotune(BWs* w, const Params& p0, hipStream_t st) {
  Params p = p0;
  p.autorho = 0;
  p.abs_tol = 0.f;
  p.rel_tol = 0.f;
  BCfg best = w->cfg;
  int rc;
  auto reset = [&]() -> int {
    launch_zero(w, st);
    return upload_ctrl(w, 1.f, st);
  };
  float bt = 1e30f, t;
  for (int lds : {16384, 32768, 65536}) {
    BCfg c = best;
    c.col_lds = lds;
    if ((rc = configure(w, c, st)) || (rc = reset())) return rc;
    if ((rc = time_it(w, st, 3,
                      [&] { launch_col_b(w, kLColM, kSelWork, 1.f, st); },
                      &t)))
      return rc;
    if (t < bt) {
      bt = t;
      best.col_lds = lds;
    }
  }
  bt = 1e30f;
  // Rows per wavefront (wavefront-per-row kernels) or an LDS budget of 15,
  // 30 or 60 KiB (block-wide kernels).
  for (int v : {1, 2, 4}) {
    BCfg c = best;
    if (w->wave)
      c.rpw = v;
    else
      c.row_lds = 15360 * v;
    if ((rc = configure(w, c, st)) || (rc = reset())) return rc;
    if ((rc = time_it(w, st, 3, [&] { launch_row_b(w, p, kRowFused, st); },
                      &t)))
      return rc;
    if (t < bt) {
      bt = t;
      best.rpw = c.rpw;
      best.row_lds = c.row_lds;
    }
  }
  bt = 1e30f;
  for (int mpt : {4, 8, 16}) {
    if (solve_lanes(w->M, mpt) > 64) continue;
    BCfg c = best;
    c.mpt = mpt;
    if ((rc = configure(w, c, st)) || (rc = reset())) return rc;
    if ((rc = time_it(w, st, 3, [&] { launch_solve_b(w, st); }, &t)))
      return rc;
    if (t < bt) {
      bt = t;
      best.mpt = mpt;
    }
  }
  bt = 1e30f;
  for (int pers : {0, 1}) {
    BCfg c = best;
    c.persistent = pers;
    if ((rc = configure(w, c, st)) || (rc = reset())) return rc;
    if ((rc = time_it(w, st, 2, [&] { launch_iter(w, p, 1, st); }, &t)))
      return rc;
    if (t < bt) {
      bt = t;
      best.persistent = pers;
    }
  }
  // Work lists against per-image chains on 1, 2 or 4 branches, over two
  // iterations (with the chains' deferred reduction).
  if ((rc = configure(w, best, st)) || (rc = reset())) return rc;
  if ((rc = time_it(w, st, 2,
                    [&] {
                      launch_iter(w, p, 1, st);
                      launch_iter(w, p, 2, st);
                    },
                    &bt)))
    return rc;
  for (int k : {1, 2, 4}) {
    BCfg c = best;
    c.branches = k;
    if ((rc = configure(w, c, st)) || (rc = reset())) return rc;
    int lrc = SCC_OK;
    if ((rc = time_it(w, st, 2,
                      [&] {
                        if (lrc == SCC_OK) lrc = launch_chains(w, p, 1, 2, 2, st);
                      },
                      &t)))
      return rc;
    if (lrc) return lrc;
    if (t < bt) {
      bt = t;
      best.branches = k;
    }
  }
  return configure(w, best, st);
}

}  // namespace
}  // namespace scc

using namespace scc;

extern "C" {

int scc_batched_set_autotune(int on) {
  std::lock_guard<std::mutex> lk(b_mu);
  b_autotune = on != 0;
  return SCC_OK;
}

int scc_batched_set_graph(int on) {
  std::lock_guard<std::mutex> lk(b_mu);
  b_graph = on != 0;
  return SCC_OK;
}

int scc_batched_set_branches(int k) {
  std::lock_guard<std::mutex> lk(b_mu);
  b_branches = k < 0 ? -1 : k;
  return SCC_OK;
}

int scc_batched_last_config(scc_batch_config* cfg) {
  if (cfg == nullptr) return SCC_ERR_INVALID;
  *cfg = b_last;
  return SCC_OK;
}

int scc_batched_cbpdn(const float* D, int M, int Kh, int Kw, int nimg,
                      const scc_batch_image* images, const scc_algopt* opt,
                      scc_batch_result* results, void* stream) {
  // Validation before any device call.
  if (D == nullptr || images == nullptr || opt == nullptr)
    return set_error(SCC_ERR_INVALID, "NULL argument");
  if (M < 1 || Kh < 1 || Kw < 1 || nimg < 1)
    return set_error(SCC_ERR_INVALID, "M, Kh, Kw and nimg must be >= 1");
  if (M > 4096) return set_error(SCC_ERR_UNSUPPORTED, "M too large");
  scc_algopt o = *opt;
  scc_default_opts(&o);
  if (o.MaxMainIter < 0)
    return set_error(SCC_ERR_INVALID, "MaxMainIter must be >= 0");
  for (int i = 0; i < nimg; ++i)
    if (!(images[i].rho > 0.f) && (!(o.rho > 0.f) || !std::isfinite(o.rho)))
      return set_error(SCC_ERR_INVALID, "rho must be positive and finite");
  if (o.AutoRho == 1 && o.AutoRhoPeriod < 1)
    return set_error(SCC_ERR_INVALID, "AutoRhoPeriod must be >= 1");
  if (o.L1Weight == nullptr)
    return set_error(SCC_ERR_INVALID, "L1Weight is NULL");
  const long long nl1 = (long long)std::max(o.L1_WEIGHT_M_SIZE, 0) *
                        std::max(o.L1_WEIGHT_ROW_SIZE, 0) *
                        std::max(o.L1_WEIGHT_COL_SIZE, 0);
  if (nl1 != 1 && nl1 != M)
    return set_error(nl1 == 0 ? SCC_ERR_INVALID : SCC_ERR_UNSUPPORTED,
                     "batched L1Weight must hold 1 or M entries");
  if (o.nWeight > 1)
    return set_error(SCC_ERR_UNSUPPORTED,
                     "masks are not supported by the batched solver");
  std::vector<int> sizes(2 * (size_t)nimg);
  for (int i = 0; i < nimg; ++i) {
    const scc_batch_image& b = images[i];
    if (b.S == nullptr || b.Y == nullptr)
      return set_error(SCC_ERR_INVALID, "NULL image or output pointer");
    if (b.R < 1 || b.C < 1 || Kh > b.R || Kw > b.C)
      return set_error(SCC_ERR_INVALID,
                       "image " + std::to_string(i) +
                           ": sizes must be >= 1 and >= the filter size");
    if (!gen::sides_fit_lds(b.R, b.C) ||
        (long long)M * b.R * b.C > (1LL << 36))
      return set_error(SCC_ERR_UNSUPPORTED,
                       "image " + std::to_string(i) + " too large");
    if (!std::isfinite(b.lambda))
      return set_error(SCC_ERR_INVALID, "lambda must be finite");
    if (!std::isfinite(b.rho))
      return set_error(SCC_ERR_INVALID, "rho must be finite");
    sizes[2 * i] = b.R;
    sizes[2 * i + 1] = b.C;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return set_error(SCC_ERR_NODEVICE, "no HIP device available");
  if (o.device < 0 || o.device >= ndev)
    return set_error(SCC_ERR_NODEVICE, "device index out of range");

  std::lock_guard<std::mutex> lk(b_mu);
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) prev = -1;
  struct Restore {
    int d;
    ~Restore() {
      if (d >= 0) (void)hipSetDevice(d);
    }
  } restore{prev};
  BHIP(hipSetDevice(o.device));
  BWs* w;
  int rc;
  if ((rc = get_bws(o.device, M, sizes, &w))) return rc;
  hipStream_t st = (hipStream_t)stream;

  Params p{};
  p.alpha = o.RelaxParam;
  p.abs_tol = o.AbsStopTol;
  p.rel_tol = o.RelStopTol;
  p.tau = o.RhoScaling;
  p.rsdl_ratio = o.RhoRsdlRatio;
  p.std_residuals = o.StdResiduals == 1;
  p.autorho = o.AutoRho == 1;
  p.autoscaling = o.AutoRhoScaling == 1;
  p.period = std::max(o.AutoRhoPeriod, 1);
  p.nonneg = o.NonNegCoef == 1;
  p.nobndry = o.NoBndryCross == 1;
  p.Kh = Kh;
  p.Kw = Kw;
  p.l1_mode = nl1 == 1 ? 0 : 1;
  p.l1_scalar = 1.f;
  p.gw_scalar = 1.f;
  int neg = 0;
  if (nl1 == 1) {
    BHIP(hipMemcpyAsync(&p.l1_scalar, o.L1Weight, 4, hipMemcpyDeviceToHost,
                        st));
  } else {
    BHIP(hipMemcpyAsync(w->l1w.p, o.L1Weight, (size_t)M * 4,
                        hipMemcpyDeviceToDevice, st));
    if ((rc = w->flag.get(4))) return rc;
    BHIP(launch_any_negative((const float*)w->l1w.p, M, (int*)w->flag.p, st));
    BHIP(hipMemcpyAsync(&neg, w->flag.p, 4, hipMemcpyDeviceToHost, st));
  }
  BHIP(hipStreamSynchronize(st));
  // V state needs every threshold (lambda_i / rho) w >= 0.
  w->vs = neg == 0 && (nl1 != 1 || p.l1_scalar >= 0.f);
  for (int i = 0; i < nimg; ++i) w->vs = w->vs && images[i].lambda >= 0.f;

  // Per-call descriptor fields.
  for (int i = 0; i < nimg; ++i) {
    BImg& im = w->h[i];
    im.S = images[i].S;
    im.Y = images[i].Y;
    im.lambda = images[i].lambda;
    im.rho0 = images[i].rho > 0.f ? images[i].rho : o.rho;
    float xi = o.RhoRsdlTarget;
    if (!(xi > 0.f))
      xi = o.StdResiduals == 1
               ? 1.f
               : (float)(1.0 + std::pow(18.3, std::log10((double)im.lambda) +
                                                  1.0));
    im.xi = xi;
  }
  b_last = scc_batch_config{};
  if (!w->tuned) {
    if (b_autotune) {
      BHIP(hipEventRecord(w->ev[2], st));
      if ((rc = configure(w, BCfg{}, st)) || (rc = autotune(w, p, st)))
        return rc;
      BHIP(hipEventRecord(w->ev[0], st));
      BHIP(hipEventSynchronize(w->ev[0]));
      float ms = 0.f;
      BHIP(hipEventElapsedTime(&ms, w->ev[2], w->ev[0]));
      b_last.tune_ms = ms;
    }
    w->tuned = true;
    w->tcfg = w->cfg;
    if (w->exec) {
      (void)hipGraphExecDestroy(w->exec);
      w->exec = nullptr;
    }
  }
  BCfg use = w->tcfg;
  if (b_branches >= 0) use.branches = std::min(b_branches, kMaxBranches);
  const bool chains = use.branches > 0;
  if (chains) {
    size_t nreal = 0;
    for (int i = 0; i < nimg; ++i)
      nreal = std::max(nreal, (size_t)(w->h[i].U - (float*)w->U.p) +
                                  (size_t)M * (size_t)w->h[i].g.N);
    if ((rc = w->Yw.get(nreal * 4))) return rc;
    for (int i = 0; i < nimg; ++i)
      w->h[i].Y = (float*)w->Yw.p + (w->h[i].U - (float*)w->U.p);
  }
  if ((rc = configure(w, use, st))) return rc;

  // Setup: spectra of the padded filters (per image size) and the images.
  BHIP(hipEventRecord(w->ev[0], st));
  hipLaunchKernelGGL(k_bpad, dim3(512, nimg), dim3(256), 0, st,
                     (const BImg*)w->desc.p, D, Kh, Kw);
  if (chains) {
    for (int j = 0; j < nimg; ++j)
      if ((rc = chain_setup(w, j, st))) return rc;
  } else {
    launch_row_fwd_b(w, kLRowM, kSelDict, st);
    launch_col_b(w, kLColM, kSelDict, 1.f, st);
    launch_row_fwd_b(w, kLRow1, kSelImage, st);
    launch_col_b(w, kLCol1, kSelImage, 1.f, st);
  }
  launch_zero(w, st);
  if ((rc = upload_ctrl(w, 0.f, st))) return rc;
  BHIP(hipGetLastError());
  BHIP(hipEventRecord(w->ev[1], st));

  const int maxit = o.MaxMainIter;
  bool graphed = false;
  if (b_graph && maxit > 0) {
    std::vector<unsigned char> key(sizeof(Params) + 3 * sizeof(int));
    std::memcpy(key.data(), &p, sizeof(Params));
    std::memcpy(key.data() + sizeof(Params), &maxit, sizeof(int));
    const int nl = (int)nl1;
    std::memcpy(key.data() + sizeof(Params) + sizeof(int), &nl, sizeof(int));
    std::memcpy(key.data() + sizeof(Params) + 2 * sizeof(int), &use.branches,
                sizeof(int));
    if (chains) {  // the chains' kernels take lambda and xi by value
      for (int i = 0; i < nimg; ++i) {
        const float v[2] = {w->h[i].lambda, w->h[i].xi};
        const unsigned char* b8 = (const unsigned char*)v;
        key.insert(key.end(), b8, b8 + sizeof(v));
      }
      key.push_back(w->vs ? 1 : 0);
      key.push_back(row_ct_enabled() ? 1 : 0);
    }
    if (w->exec == nullptr || key != w->gkey) {
      if (w->exec) {
        (void)hipGraphExecDestroy(w->exec);
        w->exec = nullptr;
      }
      BHIP(hipStreamBeginCapture(w->cap, hipStreamCaptureModeThreadLocal));
      int lrc = SCC_OK;
      if (chains)
        lrc = launch_chains(w, p, 1, maxit, maxit, w->cap);
      else
        for (int i = 1; i <= maxit && lrc == SCC_OK; ++i)
          lrc = launch_iter(w, p, i, w->cap);
      hipGraph_t graph = nullptr;
      const hipError_t ec = hipStreamEndCapture(w->cap, &graph);
      if (lrc) {
        if (graph) (void)hipGraphDestroy(graph);
        return lrc;
      }
      BHIP(ec);
      const hipError_t ei =
          hipGraphInstantiate(&w->exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      BHIP(ei);
      w->gkey = key;
    }
    BHIP(hipGraphLaunch(w->exec, st));
    graphed = true;
  } else if (chains) {
    if (maxit > 0 && (rc = launch_chains(w, p, 1, maxit, maxit, st)))
      return rc;
  } else {
    for (int i = 1; i <= maxit; ++i)
      if ((rc = launch_iter(w, p, i, st))) return rc;
  }
  BHIP(hipEventRecord(w->ev[2], st));
  if (chains)  // Y (= shrink(V) in the V state) into the outputs
    for (int i = 0; i < nimg; ++i) {
      if (w->vs)
        BHIP(launch_state_to_y(w->ch[i].g, chain_buffers(w, i),
                               chain_params(w, i, p), images[i].Y, st));
      else
        BHIP(hipMemcpyAsync(images[i].Y, w->h[i].Y,
                            (size_t)M * w->h[i].g.N * 4,
                            hipMemcpyDeviceToDevice, st));
    }
  BHIP(hipMemcpyAsync(w->hctrl, w->ctrl.p, (size_t)nimg * sizeof(Ctrl),
                      hipMemcpyDeviceToHost, st));
  BHIP(hipStreamSynchronize(st));
  if (results != nullptr)
    for (int i = 0; i < nimg; ++i) {
      const Ctrl& c = w->hctrl[i];
      results[i].iterations = c.iter;
      results[i].rho = c.rho;
      results[i].r = c.r;
      results[i].s = c.s;
      results[i].epri = c.epri;
      results[i].edua = c.edua;
    }
  float ms = 0.f;
  b_last.col_lds = w->cfg.col_lds;
  b_last.row_lds = w->lds_row;
  b_last.solve_lanes = w->G;
  b_last.solve_maps = w->cfg.mpt;
  b_last.persistent = w->cfg.persistent;
  b_last.branches = w->cfg.branches;
  b_last.graph = graphed ? 1 : 0;
  if (hipEventElapsedTime(&ms, w->ev[0], w->ev[1]) == hipSuccess)
    b_last.setup_ms = ms;
  if (hipEventElapsedTime(&ms, w->ev[1], w->ev[2]) == hipSuccess)
    b_last.solve_ms = ms;
  return SCC_OK;
}

}  // extern "C"
