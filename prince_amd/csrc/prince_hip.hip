// prince_hip.hip -- MI355X (gfx950) kernels and C ABI for the ETD2
// propagation hot path.  See include/prince_hip.h for the data layout.
//
// Kernels:
//   fold_gemm_mfma   FP64 MFMA GEMM folding the rate kernel B against the
//                    photon vectors of every refresh window at once.
//   gather_rows      row gather building the block-band ordered fold kernel.
//   etd2_stage<S2>   one fused ETD2 stage: off-diagonal coupling product in
//                    block-band layout, FD stencil on kappa*x with wave
//                    shuffles, diagonal phi factors, source term and the
//                    stage update.  One wave64 per (species block, rhs);
//                    lanes are energy bins.  Rows with many coupling blocks
//                    are split into parts whose partial sums are combined,
//                    in part order, by the last part to finish.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/prince_hip.h"

namespace {

thread_local std::string g_last_error = "no error";

int32_t set_error(const std::string& msg) {
  g_last_error = msg;
  return 1;
}

int32_t check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    return set_error(std::string(what) + ": " + hipGetErrorString(e));
  }
  return 0;
}

constexpr int WAVE = 64;
constexpr int WAVES_PER_BLOCK = 4;
constexpr int THREADS = WAVE * WAVES_PER_BLOCK;

// phi factor thresholds of the reference step (solvers/etd2.py:43-46)
constexpr double PHI1_SMALL = 1e-6;
constexpr double PHI2_SMALL = 1e-3;

typedef double double4_t __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Fold GEMM: out[w, r] = sum_j phot[w, j] * kern[r, j]
// One wave owns 16 kernel rows and keeps their B fragments in registers for
// the whole window loop; each 16-window tile is 16x16 outputs accumulated by
// v_mfma_f64_16x16x4 over the photon axis.
// Operand layout of the 16x16x4 f64 MFMA (lane l):
//   A (16 x 4): A[l % 16][l / 16];  B (4 x 16): B[l / 16][l % 16];
//   D (16 x 16): D[l / 16 + 4 * v][l % 16], v = 0..3 (rows interleaved by 4).
// ---------------------------------------------------------------------------
constexpr int FOLD_MAX_KSTEPS = 32;  // dph <= 128

__global__ void __launch_bounds__(THREADS) fold_gemm_mfma(const double* __restrict__ kern,
                                                          int64_t nrows, int dph,
                                                          const double* __restrict__ phot,
                                                          int nwin, double* __restrict__ out,
                                                          int64_t out_stride) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t wave = (int64_t)blockIdx.x * WAVES_PER_BLOCK + (threadIdx.x / WAVE);
  const int64_t r0 = wave * 16;
  if (r0 >= nrows) return;
  const int col = lane & 15;  // row of the kernel tile / column of D
  const int kq = lane >> 4;   // 0..3
  const int64_t r = r0 + col;
  const int nk = (dph + 3) / 4;

  double bfrag[FOLD_MAX_KSTEPS];
#pragma unroll
  for (int t = 0; t < FOLD_MAX_KSTEPS; ++t) {
    const int j = 4 * t + kq;
    bfrag[t] = (t < nk && j < dph && r < nrows) ? kern[r * dph + j] : 0.0;
  }

  for (int w0 = 0; w0 < nwin; w0 += 16) {
    double4_t acc = {0.0, 0.0, 0.0, 0.0};
    const int wa = w0 + col;  // window row of the A fragment
#pragma unroll
    for (int t = 0; t < FOLD_MAX_KSTEPS; ++t) {
      if (t < nk) {
        const int j = 4 * t + kq;
        const double a = (wa < nwin && j < dph) ? phot[(int64_t)wa * dph + j] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bfrag[t], acc, 0, 0, 0);
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int w = w0 + kq + 4 * v;
      if (w < nwin && r < nrows) out[(int64_t)w * out_stride + r] = acc[v];
    }
  }
}

__global__ void gather_rows_kernel(const double* __restrict__ src, const int64_t* __restrict__ sel,
                                   int64_t nrows, int ncols, double* __restrict__ dst) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = nrows * (int64_t)ncols;
  if (idx >= total) return;
  const int64_t r = idx / ncols;
  const int c = (int)(idx - r * ncols);
  const int64_t s = sel[r];
  dst[idx] = (s >= 0) ? src[s * ncols + c] : 0.0;
}

// ---------------------------------------------------------------------------
// ETD2 stage kernel
// ---------------------------------------------------------------------------
__device__ __forceinline__ double shfl_or_zero(double v, int src, int d) {
  const double r = __shfl(v, src < 0 ? 0 : (src > 63 ? 63 : src), WAVE);
  return (src >= 0 && src < d) ? r : 0.0;
}

template <bool STAGE2>
__global__ void __launch_bounds__(THREADS) etd2_stage(prh_operator op, prh_schedule sch,
                                                      prh_state st, int step) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t wave = (int64_t)blockIdx.x * WAVES_PER_BLOCK + (threadIdx.x / WAVE);
  const int K = op.K;
  const int d = op.d;
  if (wave >= (int64_t)op.nitems * K) return;
  const int item = (int)(wave / K);
  const int k = (int)(wave - (int64_t)item * K);
  const int* it = op.items + 6 * item;
  const int da = it[0];
  const int b_begin = it[1];
  const int b_end = it[2];
  const int part = it[3];
  const int nparts = it[4];
  const int part_base = it[5];
  const bool active = lane < d;

  const double* __restrict__ xin = STAGE2 ? st.a : st.x;
  const int w = sch.win[step];
  const double* __restrict__ mv = op.mvals + (int64_t)w * op.mvals_stride;

  // ---- off-diagonal coupling: sum over blocks of the row, in column order
  double acc = 0.0;
  for (int b = b_begin; b < b_end; ++b) {
    const int* blk = op.blocks + 4 * b;
    const int mo = blk[0];
    const int olo = blk[1];
    const int nslot = blk[2];
    const int64_t voff = blk[3];
    const double xv = active ? xin[((int64_t)mo * K + k) * d + lane] : 0.0;
    if (nslot == 1 && olo == 0) {
      if (active) acc += mv[voff + lane] * xv;
    } else {
      for (int s = 0; s < nslot; ++s) {
        const double xs = shfl_or_zero(xv, lane + olo + s, d);
        if (active) acc += mv[voff + (int64_t)s * d + lane] * xs;
      }
    }
  }

  if (nparts > 1) {
    // write this part's row sums; the last part to arrive combines them
    double* pbuf = st.partials + ((int64_t)(part_base + part) * K + k) * d;
    if (active) pbuf[lane] = acc;
    __threadfence();
    int prev = 0;
    int* ctr = st.counters + (int64_t)part_base * K + k;
    if (lane == 0) prev = atomicAdd(ctr, 1);
    prev = __shfl(prev, 0, WAVE);
    if (prev != nparts - 1) return;
    __threadfence();
    acc = 0.0;
    for (int p = 0; p < nparts; ++p) {
      const volatile double* q = st.partials + ((int64_t)(part_base + p) * K + k) * d;
      acc += active ? q[lane] : 0.0;
    }
    if (lane == 0) atomicExch(ctr, 0);
  }

  // ---- epilogue (row block da, rhs k, bin = lane)
  const int64_t own = ((int64_t)da * K + k) * d + lane;
  const double xown = active ? xin[own] : 0.0;
  const double dl = sch.dldz[step];
  const double hh = sch.h[step];
  const bool pdiff = (op.flags & PRH_F_PARTIAL_DIFF) != 0;
  double kap = 0.0;
  if (active) {
    kap = sch.hadia[step] * op.egrid[lane];
    kap = kap + op.pscale[da] * op.rpair[(int64_t)w * d + lane];
  }
  double out = acc;
  if (pdiff) {
    const double kx = kap * xown;
    double dx = 0.0;
#pragma unroll
    for (int o = -3; o <= 3; ++o) {
      if (o == 0) continue;
      const double kxo = shfl_or_zero(kx, lane + o, d);
      const double c = active ? op.dcoef[lane * 7 + (o + 3)] : 0.0;
      if (c != 0.0) dx += c * kxo;
    }
    out = out + dx;
  }
  out = out * dl;
  if ((op.flags & PRH_F_INJECTION) && op.src_slot[da] >= 0 && active) {
    const int slot = op.src_slot[da];
    const int J = op.nsrc;
    double bsum = 0.0;
    for (int j = 0; j < J; ++j) {
      const double c = sch.src_coef[((int64_t)step * J + j) * K + k];
      const double g = op.src_grid[(((int64_t)slot * J + j) * K + k) * d + lane];
      bsum = (j == 0) ? c * g : bsum + c * g;
    }
    out = out + sch.finj[step] * bsum;
  }
  if (!active) return;

  // ---- diagonal and phi factors (etd2.py:329-390)
  double dval = mv[op.zslots + (int64_t)da * d + lane] * dl;
  if (pdiff) {
    double kxd = kap * op.ddiag[lane];
    kxd = kxd * dl;
    dval = dval + kxd;
  }
  const double hd = dval * hh;
  const double e = exp(hd);
  const double ahd = fabs(hd);

  if (!STAGE2) {
    double phi1;
    if (ahd > PHI1_SMALL) {
      phi1 = (e - 1.0) / hd;
    } else {
      phi1 = (hd * (1.0 / 6.0) + 0.5) * hd + 1.0;
    }
    double a = e * xown;
    double sc = phi1 * out;
    sc = sc * hh;
    a = a + sc;
    st.a[own] = a;
    st.fphi[own] = out;
  } else {
    double phi2;
    if (ahd > PHI2_SMALL) {
      phi2 = ((e - 1.0) - hd) / (hd * hd);
    } else {
      phi2 = (hd * (1.0 / 24.0) + (1.0 / 6.0)) * hd + 0.5;
    }
    double sc = out - st.fphi[own];
    sc = sc * hh;
    sc = sc * phi2;
    st.x[own] = xown + sc;
  }
}

}  // namespace

extern "C" {

int32_t prh_version(void) { return 1; }

const char* prh_last_error(void) { return g_last_error.c_str(); }

int32_t prh_fold_gemm(const double* kern, int64_t nrows, int32_t dph, const double* phot,
                      int32_t nwin, double* out, int64_t out_stride, void* stream) {
  if (nrows < 0 || nwin < 0) return set_error("prh_fold_gemm: negative size");
  if (dph <= 0 || dph > 4 * FOLD_MAX_KSTEPS) return set_error("prh_fold_gemm: dph must be in [1, 128]");
  if (out_stride < nrows) return set_error("prh_fold_gemm: out_stride < nrows");
  if (nrows == 0 || nwin == 0) return 0;
  if (!kern || !phot || !out) return set_error("prh_fold_gemm: null pointer");
  const int64_t waves = (nrows + 15) / 16;
  const int64_t blocks = (waves + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
  hipLaunchKernelGGL(fold_gemm_mfma, dim3((unsigned)blocks), dim3(THREADS), 0,
                     (hipStream_t)stream, kern, nrows, (int)dph, phot, (int)nwin, out, out_stride);
  return check_hip(hipGetLastError(), "prh_fold_gemm launch");
}

int32_t prh_gather_rows(const double* src, const int64_t* sel, int64_t nrows, int32_t ncols,
                        double* dst, void* stream) {
  if (nrows < 0 || ncols <= 0) return set_error("prh_gather_rows: bad size");
  if (nrows == 0) return 0;
  if (!src || !sel || !dst) return set_error("prh_gather_rows: null pointer");
  const int64_t total = nrows * (int64_t)ncols;
  const int64_t blocks = (total + 255) / 256;
  hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     src, sel, nrows, (int)ncols, dst);
  return check_hip(hipGetLastError(), "prh_gather_rows launch");
}

int32_t prh_etd2_steps(const prh_operator* op, const prh_schedule* sch, const prh_state* st,
                       int32_t k_begin, int32_t k_end, void* stream) {
  if (!op || !sch || !st) return set_error("prh_etd2_steps: null argument");
  if (op->d <= 0 || op->d > WAVE) return set_error("prh_etd2_steps: d must be in [1, 64]");
  if (op->K <= 0) return set_error("prh_etd2_steps: K must be >= 1");
  if (k_begin < 0 || k_end > sch->nsteps || k_begin > k_end)
    return set_error("prh_etd2_steps: step range outside the schedule");
  if (op->nitems <= 0 || k_begin == k_end) return 0;
  const int64_t waves = (int64_t)op->nitems * op->K;
  const int64_t blocks = (waves + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
  if (blocks > 0x7fffffff) return set_error("prh_etd2_steps: grid too large");
  hipStream_t s = (hipStream_t)stream;
  for (int k = k_begin; k < k_end; ++k) {
    hipLaunchKernelGGL(etd2_stage<false>, dim3((unsigned)blocks), dim3(THREADS), 0, s, *op, *sch,
                       *st, k);
    hipLaunchKernelGGL(etd2_stage<true>, dim3((unsigned)blocks), dim3(THREADS), 0, s, *op, *sch,
                       *st, k);
  }
  return check_hip(hipGetLastError(), "prh_etd2_steps launch");
}

}  // extern "C"
