// Register-resident kernels for power-of-two image sizes (kernels_pow2.hip).
#pragma once

#include "scc_internal.h"

namespace scc {

struct Pow2Config {
  int TC;        // panel width
  int rows;      // rows per row-kernel block
  int variant;   // kernel family index
};

// Chooses the power-of-two kernels for (R, C, M); false when the generic
// kernels must be used.
bool pow2_select(int R, int C, int M, Pow2Config* cfg);
int pow2_row_blocks(const Geom& g, const Pow2Config& cfg);
hipError_t pow2_row_fwd(const Geom& g, const Pow2Config& cfg, int nmaps,
                        int mode, const float* src, float* Y, float* U,
                        float2* F, const float2* twC, const Ctrl* ctrl,
                        hipStream_t st, int iter = 0);
// Inverse real row transforms of nmaps maps of half spectra into real
// (nmaps, R, C) rows, times scale.
hipError_t pow2_row_inv(const Geom& g, const Pow2Config& cfg, int nmaps,
                        const float2* F, float* out, const float2* twC,
                        float scale, hipStream_t st);
// With nparts > 0 the first pow2_defer_blocks(g, nmaps) blocks of the
// column kernel also sum slices of the previous row kernel's nparts
// residual partials into bsum (5 doubles each): the first pass of a
// deferred reduction.
int pow2_defer_blocks(const Geom& g, int nmaps);
hipError_t pow2_col(const Geom& g, const Pow2Config& cfg, int nmaps,
                    float2* F, const float2* twR, int inverse,
                    const Ctrl* ctrl, hipStream_t st,
                    const float* parts = nullptr, int nparts = 0,
                    double* bsum = nullptr);
// nimg same-size images (maps of image i at F + i M Hc, spectrum at
// Sf + i Hc) share the dictionary spectra.  ctl_iter > 0: block 0 first
// finishes the deferred reduction of iteration ctl_iter from the nbsum
// block sums in b.blocksum and runs the control update.
hipError_t pow2_solve(const Geom& g, const Pow2Config& cfg, const Buffers& b,
                      const Params& p, hipStream_t st, int nimg = 1,
                      int ctl_iter = 0, int nbsum = 0);
hipError_t pow2_row(const Geom& g, const Pow2Config& cfg, const Buffers& b,
                    const Params& p, int mode, int* nblocks, hipStream_t st);
// The register-resident X-solve for any image size (one image, M <= 256).
bool solve_regs_fits(int M);
hipError_t solve_regs(const Geom& g, const Buffers& b, const Params& p,
                      hipStream_t st, int ctl_iter = 0, int nbsum = 0);

}  // namespace scc
