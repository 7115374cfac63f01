// prince_hip.hip -- MI355X (gfx950) kernels and C ABI for the ETD2
// propagation hot path.  See include/prince_hip.h for the data layout.
//
// Kernels:
//   fold_gemm_mfma   FP64 MFMA GEMM folding the rate kernel B against the
//                    photon vectors of every refresh window at once.
//   gather_rows      row gather building the block-band ordered fold kernel.
//   etd2_stage<S2,KT> one fused ETD2 stage: off-diagonal coupling product in
//                    block-band layout, FD stencil on kappa*x with wave
//                    shuffles, diagonal phi factors, source term and the
//                    stage update.  One wave64 per (species-row part, KT
//                    right-hand sides); lanes are energy bins, so each
//                    coupling value loaded is used for KT right-hand sides.
//                    The parts of a long row share a 16-wave workgroup and
//                    are summed in part order through LDS.
//   step graphs      hipGraph of a chunk of steps whose kernels read the
//                    current step from device memory, replayed per chunk.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
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

// Index of this wave in its workgroup, as a wave-uniform (scalar) value: the
// compiler cannot prove threadIdx.x / 64 uniform by itself, and values read
// through it (task fields, block ranges) would otherwise live in vector
// registers behind exec-masked branches.
__device__ __forceinline__ int wave_id() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
}
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
// v_mfma_f64_16x16x4 over the photon axis.  The 16 waves of a workgroup share
// each block of 32 windows' photon rows through LDS (635 -> 510 us for the
// 288 windows of the nuclear run: 0.31 -> 0.38 of FP64 MFMA peak).
// Operand layout of the 16x16x4 f64 MFMA (lane l):
//   A (16 x 4): A[l % 16][l / 16];  B (4 x 16): B[l / 16][l % 16];
//   D (16 x 16): D[l / 16 + 4 * v][l % 16], v = 0..3 (rows interleaved by 4).
// ---------------------------------------------------------------------------
constexpr int FOLD_MAX_KSTEPS = 32;  // dph <= 128

constexpr int FOLD_WAVES = 16;  // waves sharing one staged window block
constexpr int FOLD_THREADS = FOLD_WAVES * WAVE;

__global__ void __launch_bounds__(FOLD_THREADS) fold_gemm_mfma(const double* __restrict__ kern,
                                                               int64_t nrows, int dph,
                                                               const double* __restrict__ phot,
                                                               int nwin, double* __restrict__ out,
                                                               int64_t out_stride) {
  // the photon rows of the current 32 windows, staged in LDS once per
  // workgroup (row stride dph + 1): the A fragments are LDS reads instead of
  // L2 gathers, and two 16-window tiles run as independent MFMA chains
  __shared__ double sp[32 * (4 * FOLD_MAX_KSTEPS + 1)];
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t wave = (int64_t)blockIdx.x * FOLD_WAVES + wave_id();
  const int64_t r0 = wave * 16;
  const bool rows = r0 < nrows;  // waves past the last row still take part in the barriers
  const int col = lane & 15;  // row of the kernel tile / column of D
  const int kq = lane >> 4;   // 0..3
  const int64_t r = r0 + col;
  const int nk = (dph + 3) / 4;
  const int ps = dph + 1;

  double bfrag[FOLD_MAX_KSTEPS];
#pragma unroll
  for (int t = 0; t < FOLD_MAX_KSTEPS; ++t) {
    const int j = 4 * t + kq;
    bfrag[t] = (rows && t < nk && j < dph && r < nrows) ? kern[r * dph + j] : 0.0;
  }

  for (int w0 = 0; w0 < nwin; w0 += 32) {
    __syncthreads();  // the previous windows are consumed
    for (int i = threadIdx.x; i < 32 * dph; i += FOLD_THREADS) {
      const int wr = i / dph, j = i % dph;
      sp[wr * ps + j] = w0 + wr < nwin ? phot[(int64_t)(w0 + wr) * dph + j] : 0.0;
    }
    __syncthreads();
    if (!rows) continue;
    double4_t acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = acc0;
#pragma unroll
    for (int t = 0; t < FOLD_MAX_KSTEPS; ++t) {
      if (t < nk) {
        const int j = 4 * t + kq;
        const double a0 = j < dph ? sp[col * ps + j] : 0.0;
        const double a1 = j < dph ? sp[(16 + col) * ps + j] : 0.0;
        acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bfrag[t], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bfrag[t], acc1, 0, 0, 0);
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int w = w0 + kq + 4 * v;
      if (w < nwin && r < nrows) out[(int64_t)w * out_stride + r] = acc0[v];
      if (w + 16 < nwin && r < nrows) out[(int64_t)(w + 16) * out_stride + r] = acc1[v];
    }
  }
}

// Same product for a contraction length above 128 (the cascade tensors: 400
// and 256 target-photon points): the kernel rows are re-read from cache per
// 16-window tile instead of held in registers.
__global__ void __launch_bounds__(THREADS) contract_rows_mfma(const double* __restrict__ kern,
                                                              int64_t nrows, int kdim,
                                                              const double* __restrict__ vec,
                                                              int nvec, double* __restrict__ out,
                                                              int64_t out_stride) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t wave = (int64_t)blockIdx.x * WAVES_PER_BLOCK + wave_id();
  const int64_t r0 = wave * 16;
  if (r0 >= nrows) return;
  const int col = lane & 15;
  const int kq = lane >> 4;
  const int64_t r = r0 + col;
  const double* kr = kern + (r < nrows ? r : nrows - 1) * (int64_t)kdim;
  for (int w0 = 0; w0 < nvec; w0 += 16) {
    double4_t acc = {0.0, 0.0, 0.0, 0.0};
    const int wa = w0 + col;
    const double* va = vec + (int64_t)(wa < nvec ? wa : nvec - 1) * kdim;
    for (int k0 = 0; k0 < kdim; k0 += 16) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0 + 4 * u + kq;
        a[u] = (wa < nvec && k < kdim) ? va[k] : 0.0;
        b[u] = (r < nrows && k < kdim) ? kr[k] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int w = w0 + kq + 4 * v;
      if (w < nvec && r < nrows) out[(int64_t)w * out_stride + r] = acc[v];
    }
  }
}

// Rows of the rate kernel expanded from their Toeplitz descriptors (one
// thread per output value; the same products in the same order as the host
// construction, so the rows equal the dense kernel's bit for bit).
__global__ void toeplitz_rows_kernel(prh_toeplitz tz, const int64_t* __restrict__ sel,
                                     int64_t nsel, double* __restrict__ out) {
  // every product rounded on its own, as the host construction does (no
  // fused multiply-add: fm * gain - fm * loss must round twice)
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int dph = tz.dph;
  if (t >= nsel * dph) return;
  const int64_t r = t / dph;
  const int j = (int)(t - r * dph);
  const int64_t s = sel[r];
  if (s < 0 || s >= tz.nrows) {
    out[t] = 0.0;
    return;
  }
  const int32_t* ds = tz.desc + 5 * s;
  const int kind = ds[0], tg = ds[1], tl = ds[2], i = ds[3], k = ds[4];
  const int d = tz.d;
  const int64_t L = d + dph - 1;
  const int w = i + j;
  const double fm = tz.fmo[i];
  double base;
  if (kind == 0) {
    base = tg >= 0 ? fm * tz.tab1d[tg * L + w] : 0.0;
  } else {
    const double v = tz.fdiff[i * d + k] * tz.tab2d[(tg * L + w) * (2 * d - 1) + (k - i + d - 1)];
    base = v >= 0.0 ? v : 0.0;
  }
  const bool loss = tl >= 0 && (kind == 0 || i == k);
  out[t] = loss ? base - fm * tz.tab1d[tl * L + w] : base;
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
constexpr int SWG = PRH_WG_WAVES;     // waves per stage workgroup
constexpr int STHREADS = SWG * WAVE;  // 1024

// waves per workgroup of the narrow stage kernels for this operator
inline int wg_waves_of(const prh_operator& op) { return op.wg_waves == 8 ? 8 : 16; }

__device__ __forceinline__ double shfl_or_zero(double v, int src, int d) {
  const double r = __shfl(v, src < 0 ? 0 : (src > 63 ? 63 : src), WAVE);
  return (src >= 0 && src < d) ? r : 0.0;
}

// Whole-wave lane shifts by one (DPP wave_shl:1 / wave_shr:1, VALU moves, no
// LDS): lane i receives lane i+1 (lane_next; lane 63 receives 0) or lane i-1
// (lane_prev; lane 0 receives 0).  Callers keep lanes >= d at zero, so a chain
// of shifts equals shfl_or_zero at offsets of one sign.  All 64 lanes must be
// active.
constexpr int DPP_WAVE_SHL1 = 0x130;
constexpr int DPP_WAVE_SHR1 = 0x138;
template <int CTRL>
__device__ __forceinline__ double lane_shift1(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double lane_next(double v) { return lane_shift1<DPP_WAVE_SHL1>(v); }
__device__ __forceinline__ double lane_prev(double v) { return lane_shift1<DPP_WAVE_SHR1>(v); }
// value of lane l (wave-uniform l) as a scalar
__device__ __forceinline__ double readlane_f64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// Buffer resource over `bytes` bytes from base (raw, no swizzle) and 64-bit
// loads / stores at a per-lane byte offset voff plus a wave-uniform soff: the
// address arithmetic stays in scalar registers.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t tile_rsrc(const double* base, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(base), (short)0, bytes, 0x00020000);
}
__device__ __forceinline__ double tload(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ void tstore(double v, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
}

// Blocks whose loads are issued together before any of them is used: runs
// of diagonal blocks and of band blocks with at most NSMAX diagonals.  With
// two photon-field components every block has two value sets.
constexpr int NSMAX = 8;
template <int KT, bool TWO>
struct Batch {
  static constexpr int DIAG = TWO ? (KT <= 2 ? 4 : 2) : (KT <= 2 ? 8 : (KT == 4 ? 4 : (KT == 8 ? 2 : 1)));
  static constexpr int BAND = TWO ? (KT <= 2 ? 1 : 1) : (KT == 1 ? 4 : (KT == 2 ? 2 : 1));
};

enum PartKind { PART_DIAG = 0, PART_BAND = 1, PART_GENERIC = 2 };

// Coupling product of one row part over its blocks, in column order:
//   accC[kk] = sum_b mC_b (shifted) y_b,   accE[kk] = sum_b mE_b (shifted) y_b
// for the KT right-hand sides k0 .. k0+KT-1 (accE only with TWO).  BACK adds
// the DPP walk for bands left of the main diagonal (the transposed operator).
template <int KT, bool TWO, bool BACK = false>
__device__ __forceinline__ void couple_part(const prh_operator& op, const double* __restrict__ y,
                                            const double* __restrict__ mvC,
                                            const double* __restrict__ mvE, int b_begin,
                                            int b_end, int kind, int k0, int lane, bool active,
                                            double (&accC)[KT], double (&accE)[KT]) {
  const int K = op.K;
  const int d = op.d;
  const int64_t Kd = (int64_t)K * d;
  // lanes past the last bin read the last bin (every load in bounds, no
  // branch); diagonal blocks leave their products in those lanes, band blocks
  // zero them before lanes mix
  const int lb = (active ? lane : d - 1) * 8;
  // the window's coupling values and each mother's [K][d] panel as buffers;
  // value offsets and right-hand-side columns (clamped to K - 1) go in the
  // scalar offset
  // buffer loads wherever the wave index is scalar (the two-field kernels
  // with more than 4 right-hand sides per wave keep plain loads)
  constexpr bool BUF = !TWO || KT <= 4;
  const int vbytes = (int)(op.mvals_stride * 8);
  const auto rmC = tile_rsrc(mvC, vbytes);
  const auto rmE = tile_rsrc(TWO ? mvE : mvC, vbytes);
  const int pbytes = (int)(Kd * 8);
  int coff[KT];
#pragma unroll
  for (int kk = 0; kk < KT; ++kk) {
    coff[kk] = (k0 + kk < K ? k0 + kk : K - 1) * d * 8;
    accC[kk] = 0.0;
    accE[kk] = 0.0;
  }
  // the first 64 block descriptors of the part, one per lane, broadcast
  // with readlane instead of a dependent load per block
  const int dbase = b_begin;
  int dmo = 0, dol = 0, dns = 0, dvo = 0;
  if (b_begin + lane < b_end) {
    const int* blk = op.blocks + 4 * (b_begin + lane);
    dmo = blk[0];
    dol = blk[1];
    dns = blk[2];
    dvo = blk[3];
  }
  if (kind == PART_DIAG) {
    constexpr int UB = Batch<KT, TWO>::DIAG;
    for (int b = b_begin; b < b_end; b += UB) {
      double mc[UB], me[UB];
      double xv[UB][KT];
#pragma unroll
      for (int u = 0; u < UB; ++u) {
        const bool valid = b + u < b_end;
        const int bb = valid ? b + u : b_end - 1;
        int mo, voffi;
        if (bb - dbase < WAVE) {
          mo = __builtin_amdgcn_readlane(dmo, bb - dbase);
          voffi = __builtin_amdgcn_readlane(dvo, bb - dbase);
        } else {
          mo = op.blocks[4 * bb];
          voffi = op.blocks[4 * bb + 3];
        }
        if constexpr (BUF) {
          mc[u] = tload(rmC, lb, voffi * 8);
          me[u] = TWO ? tload(rmE, lb, voffi * 8) : 0.0;
          const auto ry = tile_rsrc(y + (int64_t)mo * Kd, pbytes);
#pragma unroll
          for (int kk = 0; kk < KT; ++kk) xv[u][kk] = tload(ry, lb, coff[kk]);
        } else {
          const int64_t voff = voffi;
          mc[u] = active ? mvC[voff + lane] : 0.0;
          me[u] = (TWO && active) ? mvE[voff + lane] : 0.0;
#pragma unroll
          for (int kk = 0; kk < KT; ++kk) {
            const int k = k0 + kk;
            xv[u][kk] = (active && k < K) ? y[((int64_t)mo * K + k) * d + lane] : 0.0;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < UB; ++u) {
        if (b + u < b_end) {
#pragma unroll
          for (int kk = 0; kk < KT; ++kk) {
            accC[kk] += mc[u] * xv[u][kk];
            if (TWO) accE[kk] += me[u] * xv[u][kk];
          }
        }
      }
    }
  } else if (kind == PART_BAND) {
    constexpr int UB = Batch<KT, TWO>::BAND;
    for (int b = b_begin; b < b_end; b += UB) {
      double mc[UB][NSMAX], me[UB][TWO ? NSMAX : 1];
      double xv[UB][KT];
      int olo[UB], ns[UB];
#pragma unroll
      for (int u = 0; u < UB; ++u) {
        const bool valid = b + u < b_end;
        const int bb = valid ? b + u : b_end - 1;
        int mo, ol, nsl, voffi;
        if (bb - dbase < WAVE) {
          mo = __builtin_amdgcn_readlane(dmo, bb - dbase);
          ol = __builtin_amdgcn_readlane(dol, bb - dbase);
          nsl = __builtin_amdgcn_readlane(dns, bb - dbase);
          voffi = __builtin_amdgcn_readlane(dvo, bb - dbase);
        } else {
          mo = op.blocks[4 * bb];
          ol = op.blocks[4 * bb + 1];
          nsl = op.blocks[4 * bb + 2];
          voffi = op.blocks[4 * bb + 3];
        }
        olo[u] = ol;
        ns[u] = valid ? nsl : 0;
        if constexpr (BUF) {
#pragma unroll
          for (int s2 = 0; s2 < NSMAX; ++s2) {
            const int sc = s2 < nsl ? s2 : nsl - 1;
            mc[u][s2] = tload(rmC, lb, (voffi + sc * d) * 8);
            if (TWO) me[u][TWO ? s2 : 0] = tload(rmE, lb, (voffi + sc * d) * 8);
          }
          const auto ry = tile_rsrc(y + (int64_t)mo * Kd, pbytes);
#pragma unroll
          for (int kk = 0; kk < KT; ++kk) {
            const double v = tload(ry, lb, coff[kk]);
            xv[u][kk] = active ? v : 0.0;
          }
        } else {
          const int64_t voff = voffi;
#pragma unroll
          for (int s2 = 0; s2 < NSMAX; ++s2) {
            const int sc = s2 < nsl ? s2 : nsl - 1;
            mc[u][s2] = active ? mvC[voff + (int64_t)sc * d + lane] : 0.0;
            if (TWO) me[u][TWO ? s2 : 0] = active ? mvE[voff + (int64_t)sc * d + lane] : 0.0;
          }
#pragma unroll
          for (int kk = 0; kk < KT; ++kk) {
            const int k = k0 + kk;
            xv[u][kk] = (active && k < K) ? y[((int64_t)mo * K + k) * d + lane] : 0.0;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < UB; ++u) {
        if (olo[u] >= 0) {
          // diagonal s2 reads lane + olo + s2: one shuffle to the first
          // diagonal (none when olo = 0), then one-lane DPP shifts
          if (olo[u] != 0) {
#pragma unroll
            for (int kk = 0; kk < KT; ++kk) xv[u][kk] = shfl_or_zero(xv[u][kk], lane + olo[u], d);
          }
#pragma unroll
          for (int s2 = 0; s2 < NSMAX; ++s2) {
            if (s2 < ns[u]) {
#pragma unroll
              for (int kk = 0; kk < KT; ++kk) {
                accC[kk] += mc[u][s2] * xv[u][kk];
                if (TWO) accE[kk] += me[u][TWO ? s2 : 0] * xv[u][kk];
              }
              if (s2 + 1 < ns[u]) {
#pragma unroll
                for (int kk = 0; kk < KT; ++kk) xv[u][kk] = lane_next(xv[u][kk]);
              }
            }
          }
        } else if (BACK && olo[u] + ns[u] <= 1) {
          // every diagonal at or left of the main one (transposed bands):
          // shuffle to the last diagonal, then walk back with wave_shr:1
          const int top = olo[u] + ns[u] - 1;
          if (top != 0) {
#pragma unroll
            for (int kk = 0; kk < KT; ++kk) xv[u][kk] = shfl_or_zero(xv[u][kk], lane + top, d);
          }
#pragma unroll
          for (int t = 0; t < NSMAX; ++t) {
            const int s2 = NSMAX - 1 - t;
            if (s2 < ns[u]) {
#pragma unroll
              for (int kk = 0; kk < KT; ++kk) {
                accC[kk] += mc[u][s2] * xv[u][kk];
                if (TWO) accE[kk] += me[u][TWO ? s2 : 0] * xv[u][kk];
              }
              if (s2 > 0) {
#pragma unroll
                for (int kk = 0; kk < KT; ++kk) xv[u][kk] = lane_prev(xv[u][kk]);
              }
            }
          }
        } else {
#pragma unroll
          for (int s2 = 0; s2 < NSMAX; ++s2) {
            if (s2 < ns[u]) {
              const int src = lane + olo[u] + s2;
#pragma unroll
              for (int kk = 0; kk < KT; ++kk) {
                const double xs = shfl_or_zero(xv[u][kk], src, d);
                accC[kk] += mc[u][s2] * xs;
                if (TWO) accE[kk] += me[u][TWO ? s2 : 0] * xs;
              }
            }
          }
        }
      }
    }
  } else {
    for (int b = b_begin; b < b_end; ++b) {
      const int* blk = op.blocks + 4 * b;
      const int mo = blk[0];
      const int olo = blk[1];
      const int nslot = blk[2];
      const int64_t voff = blk[3];
      double xv[KT];
#pragma unroll
      for (int kk = 0; kk < KT; ++kk) {
        const int k = k0 + kk;
        xv[kk] = (active && k < K) ? y[((int64_t)mo * K + k) * d + lane] : 0.0;
      }
      for (int s2 = 0; s2 < nslot; ++s2) {
        const int src = lane + olo + s2;
        const double mC = active ? mvC[voff + (int64_t)s2 * d + lane] : 0.0;
        const double mE = (TWO && active) ? mvE[voff + (int64_t)s2 * d + lane] : 0.0;
#pragma unroll
        for (int kk = 0; kk < KT; ++kk) {
          const double xs = shfl_or_zero(xv[kk], src, d);
          accC[kk] += mC * xs;
          if (TWO) accE[kk] += mE * xs;
        }
      }
    }
  }
}

// Coupling product of two consecutive rows da, da+1 whose blocks are all
// diagonal: blocks [b0, b1) of row da and [b1, b2) of row da+1, each in
// mother order.  The walk merges the two mother lists, so a mother panel that
// both rows read is loaded once and feeds both rows' FMAs; each row adds its
// own blocks in the same order as a one-row walk (same bits).
template <int KT>
__device__ __forceinline__ void couple_pair(const prh_operator& op, const double* __restrict__ y,
                                            const double* __restrict__ mv, int b0, int b1,
                                            int b2, int k0, int lane, bool active,
                                            double (&accA)[KT], double (&accB)[KT]) {
  const int K = op.K;
  const int d = op.d;
  const int64_t Kd = (int64_t)K * d;
  const int lb = (active ? lane : d - 1) * 8;
  const auto rm = tile_rsrc(mv, (int)(op.mvals_stride * 8));
  const int pbytes = (int)(Kd * 8);
  int coff[KT];
#pragma unroll
  for (int kk = 0; kk < KT; ++kk) {
    coff[kk] = (k0 + kk < K ? k0 + kk : K - 1) * d * 8;
    accA[kk] = 0.0;
    accB[kk] = 0.0;
  }
  // descriptors of the first 64 blocks of the pair, one per lane
  int dmo = 0, dvo = 0;
  if (b0 + lane < b2) {
    const int* blk = op.blocks + 4 * (b0 + lane);
    dmo = blk[0];
    dvo = blk[3];
  }
  auto mother = [&](int b) -> int {
    return b - b0 < WAVE ? __builtin_amdgcn_readlane(dmo, b - b0) : op.blocks[4 * b];
  };
  auto voff = [&](int b) -> int {
    return b - b0 < WAVE ? __builtin_amdgcn_readlane(dvo, b - b0) : op.blocks[4 * b + 3];
  };
  constexpr int SPENT = 0x7fffffff;
  // merged steps whose loads are issued together: one more than a one-row
  // walk at 8 right-hand sides (554 vs 563 ms per K = 512 solve; the one-row
  // walk is slower with 3)
  constexpr int UB = KT == 8 ? 3 : Batch<KT, false>::DIAG;
  int i = b0, j = b1;
  while (i < b1 || j < b2) {
    double mA[UB], mB[UB];
    double xv[UB][KT];
    bool uA[UB], uB[UB];
#pragma unroll
    for (int u = 0; u < UB; ++u) {
      const int ma = i < b1 ? mother(i) : SPENT;
      const int mb = j < b2 ? mother(j) : SPENT;
      const int m = ma < mb ? ma : mb;
      uA[u] = m != SPENT && ma == m;
      uB[u] = m != SPENT && mb == m;
      mA[u] = 0.0;
      mB[u] = 0.0;
      if (uA[u]) mA[u] = tload(rm, lb, voff(i) * 8);
      if (uB[u]) mB[u] = tload(rm, lb, voff(j) * 8);
      // a spent slot of the last batch reads the first block's panel (in
      // bounds, unused)
      const int mm = m != SPENT ? m : mother(b0);
      const auto ry = tile_rsrc(y + (int64_t)mm * Kd, pbytes);
#pragma unroll
      for (int kk = 0; kk < KT; ++kk) xv[u][kk] = tload(ry, lb, coff[kk]);
      i += uA[u] ? 1 : 0;
      j += uB[u] ? 1 : 0;
    }
#pragma unroll
    for (int u = 0; u < UB; ++u) {
      if (uA[u]) {
#pragma unroll
        for (int kk = 0; kk < KT; ++kk) accA[kk] += mA[u] * xv[u][kk];
      }
      if (uB[u]) {
#pragma unroll
        for (int kk = 0; kk < KT; ++kk) accB[kk] += mB[u] * xv[u][kk];
      }
    }
  }
}

// Sum the parts of split rows through LDS, in part order.  Every wave of the
// workgroup must call this (it contains the workgroup barrier).
template <int KT>
__device__ __forceinline__ void reduce_parts(double (*red)[KT][WAVE], int wv, int lane, int lead,
                                             int nparts, double (&acc)[KT]) {
  if (nparts > 1) {
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) red[wv][kk][lane] = acc[kk];
  }
  __syncthreads();
  if (nparts > 1 && wv == lead) {
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) {
      double sum = red[lead][kk][lane];
      for (int p = 1; p < nparts; ++p) sum += red[lead + p][kk][lane];
      acc[kk] = sum;
    }
  }
}

// workgroup -> (row group, right-hand-side group), XCD-aware when the number
// of right-hand-side groups is a multiple of 8 (workgroup L runs on XCD L % 8):
// XCD x owns the right-hand-side groups x, x+8, ... and walks them for one row
// group before the next, so the coupling values of a few row groups and the
// mother panels they read stay in that XCD's L2 (1.5x over walking all row
// groups of one right-hand-side group first, measured at K = 512).
__device__ __forceinline__ void wg_coords(int ntasks, int K, int kt, int& gx, int& gy,
                                          int wgw = SWG) {
  const int nx = ntasks / wgw;
  const int ny = (K + kt - 1) / kt;
  const int L = blockIdx.x;
  if (ny % 8 == 0) {
    const int j = L >> 3;
    const int per = ny / 8;
    gy = (L & 7) + 8 * (j % per);
    gx = j / per;
    (void)nx;
  } else {
    gx = L % nx;
    gy = L / nx;
  }
}

// diagonal phi factors (solvers/etd2.py:329-390)
__device__ __forceinline__ double phi1_of(double hd, double e) {
  return (fabs(hd) > PHI1_SMALL) ? (e - 1.0) / hd : (hd * (1.0 / 6.0) + 0.5) * hd + 1.0;
}
__device__ __forceinline__ double phi2_of(double hd, double e) {
  return (fabs(hd) > PHI2_SMALL) ? ((e - 1.0) - hd) / (hd * hd)
                                 : (hd * (1.0 / 24.0) + (1.0 / 6.0)) * hd + 0.5;
}

// Row-local inputs of step `step` for row da (bin = lane).
struct RowIn {
  double dl, kadia, pscale, rc, re, mdc, mde, ddiag;
};
template <bool TWO>
__device__ __forceinline__ RowIn row_in(const prh_operator& op, const prh_schedule& sch, int step,
                                        int w, int da, int lane) {
  RowIn r;
  const int d = op.d;
  r.dl = sch.dldz[step];
  r.kadia = sch.hadia[step] * op.egrid[lane];
  r.pscale = op.pscale[da];
  r.rc = op.rpair[(int64_t)w * d + lane];
  r.re = TWO ? op.rpair_e[(int64_t)w * d + lane] : 0.0;
  r.mdc = op.mvals[(int64_t)w * op.mvals_stride + op.diag_off + (int64_t)da * d + lane];
  r.mde = TWO ? op.mvals_e[(int64_t)w * op.mvals_stride + op.diag_off + (int64_t)da * d + lane]
              : 0.0;
  r.ddiag = op.ddiag[lane];
  return r;
}
struct RowCoef;
template <bool TWO>
__device__ __forceinline__ RowCoef row_coef_from(const RowIn& ri, double s, bool pdiff);

// Row-local coefficients of step `step` for row da, rhs scale s:
// kappa, kappa's EBL part, and the diagonal d (propagation.py:1389-1422).
struct RowCoef {
  double kap, kapE, dval, dE;
};
template <bool TWO>
__device__ __forceinline__ RowCoef row_coef(const prh_operator& op, const prh_schedule& sch,
                                            int step, int w, int da, int lane, double s,
                                            bool pdiff) {
  RowCoef r;
  const int d = op.d;
  const double dl = sch.dldz[step];
  const double E = op.egrid[lane];
  const double rc = op.rpair[(int64_t)w * d + lane];
  const double re = TWO ? op.rpair_e[(int64_t)w * d + lane] : 0.0;
  const double* mv = op.mvals + (int64_t)w * op.mvals_stride;
  const double mdc = mv[op.diag_off + (int64_t)da * d + lane];
  const double mde = TWO ? op.mvals_e[(int64_t)w * op.mvals_stride + op.diag_off +
                                      (int64_t)da * d + lane]
                         : 0.0;
  r.kap = sch.hadia[step] * E;
  r.kap = r.kap + op.pscale[da] * (TWO ? rc + s * re : rc);
  r.kapE = TWO ? op.pscale[da] * re : 0.0;
  r.dval = (TWO ? mdc + s * mde : mdc) * dl;
  if (pdiff) {
    double kxd = r.kap * op.ddiag[lane];
    kxd = kxd * dl;
    r.dval = r.dval + kxd;
  }
  r.dE = TWO ? (mde + r.kapE * op.ddiag[lane]) * dl : 0.0;
  return r;
}

template <bool TWO>
__device__ __forceinline__ RowCoef row_coef_from(const RowIn& ri, double s, bool pdiff) {
  RowCoef r;
  r.kap = ri.kadia + ri.pscale * (TWO ? ri.rc + s * ri.re : ri.rc);
  r.kapE = TWO ? ri.pscale * ri.re : 0.0;
  r.dval = (TWO ? ri.mdc + s * ri.mde : ri.mdc) * ri.dl;
  if (pdiff) {
    double kxd = r.kap * ri.ddiag;
    kxd = kxd * ri.dl;
    r.dval = r.dval + kxd;
  }
  r.dE = TWO ? (ri.mde + r.kapE * ri.ddiag) * ri.dl : 0.0;
  return r;
}

// One stage for the tasks `tasks` (op.tasks, or op.fin_tasks for the dense
// rows' epilogues).  step_local: step index (use_ctl = 0) or offset from
// step_ctl[0] (use_ctl = 1, graph replay; steps at or beyond step_ctl[1] do
// nothing).
// Dense-row handling compiled into a stage kernel: none, the mother-major
// pass's chunk sums in the epilogue, chunk tasks storing to dpart, or the
// dense rows' epilogue launch.
enum DenseMode { DM_NONE = 0, DM_PASS = 1, DM_CHUNKS = 2, DM_FINISH = 3 };

// Epilogue of one row block da (bins = lanes) for the KT right-hand sides
// k0 .. k0+KT-1: kappa, d, the phi factors, the FD stencil on kappa x, the
// injection term and the stage update.  acc holds the coupling sums (the
// second field already scaled in), xo / fo the row's own state and, in stage
// 2, F(x) (loaded by the caller for the buffer-addressed kernels).
template <bool STAGE2, int KT, bool TWO>
__device__ __forceinline__ void stage_epilogue(const prh_operator& op, const prh_schedule& sch,
                                               const prh_state& st, int step, int ridx,
                                               bool record, int w, int da, int k0, int lane,
                                               bool active, const double* __restrict__ xin,
                                               const double (&acc)[KT], const double (&sc_)[KT],
                                               const double (&xo)[KT], const double (&fo)[KT]) {
  constexpr bool EBUF = !TWO || KT <= 4;
  const int K = op.K;
  const int d = op.d;
  const int64_t rowoff = (int64_t)da * K * d;
  const int pb = (int)((int64_t)K * d * 8);
  const int lo = lane * 8;
  const auto r_a = tile_rsrc(st.a + rowoff, pb);
  const auto r_f = tile_rsrc(st.fphi + rowoff, pb);
  const auto r_x = tile_rsrc(st.x + rowoff, pb);
  const double dl = sch.dldz[step];
  const double hh = sch.h[step];
  const bool pdiff = (op.flags & PRH_F_PARTIAL_DIFF) != 0;
  const bool inj = (op.flags & PRH_F_INJECTION) != 0 && op.src_slot[da] >= 0;
  double dcf[7];
#pragma unroll
  for (int o = 0; o < 7; ++o) dcf[o] = (active && pdiff) ? op.dcoef[lane * 7 + o] : 0.0;
  const double finj = sch.finj[step];
  const int slot = inj ? op.src_slot[da] : 0;
  const int J = op.nsrc;
  const int64_t nk = (int64_t)op.nspec * K * d;

  // row-local inputs, loaded once (the stores below may alias them)
  RowIn ri{};
  if (active) ri = row_in<TWO>(op, sch, step, w, da, lane);
  RowCoef rc0{0.0, 0.0, 0.0, 0.0};
  if (!TWO && active) rc0 = row_coef_from<TWO>(ri, 0.0, pdiff);
  const double hd0 = rc0.dval * hh;
  const double e0 = exp(hd0);
  const double phi0 = STAGE2 ? phi2_of(hd0, e0) : phi1_of(hd0, e0);
#pragma unroll
  for (int kk = 0; kk < KT; ++kk) {
    const int k = k0 + kk;
    if (k >= K) break;  // uniform over the wave
    RowCoef rc = rc0;
    if (TWO && active) rc = row_coef_from<TWO>(ri, sc_[kk], pdiff);
    const int64_t own = ((int64_t)da * K + k) * d + lane;
    const int co = k * d * 8;
    double xown;
    if constexpr (EBUF) {
      xown = xo[kk];
    } else {
      xown = active ? xin[own] : 0.0;
    }
    double out = acc[kk];
    if (pdiff) {
      const double kx = rc.kap * xown;
      // stencil neighbours kx[lane + o], o = -3..3 (zero outside [0, d)),
      // added in the order o = -3, -2, -1, 1, 2, 3 (a zero coefficient adds
      // an exact zero: no per-term test)
      const double km1 = lane_prev(kx);
      const double km2 = lane_prev(km1);
      const double km3 = lane_prev(km2);
      double dx = 0.0;
      dx += dcf[0] * km3;
      dx += dcf[1] * km2;
      dx += dcf[2] * km1;
      double kxo = kx;
#pragma unroll
      for (int o = 1; o <= 3; ++o) {
        kxo = lane_next(kxo);
        dx += dcf[o + 3] * kxo;
      }
      out = out + dx;
    }
    out = out * dl;
    if (inj && active) {
      double bsum = 0.0;
      for (int j = 0; j < J; ++j) {
        const double c = sch.src_coef[((int64_t)step * J + j) * K + k];
        const double g = op.src_grid[(((int64_t)slot * J + j) * K + k) * d + lane];
        bsum = (j == 0) ? c * g : bsum + c * g;
      }
      out = out + finj * bsum;
    }
    if (!active) continue;
    const double hd = rc.dval * hh;
    const double e = TWO ? exp(hd) : e0;
    if (!STAGE2) {
      const double phi = TWO ? phi1_of(hd, e) : phi0;
      double a = e * xown;
      double sc = phi * out;
      sc = sc * hh;
      a = a + sc;
      if constexpr (EBUF) {
        tstore(a, r_a, lo, co);
        tstore(out, r_f, lo, co);
      } else {
        st.a[own] = a;
        st.fphi[own] = out;
      }
      if (record) {
        st.rec_x[ridx * nk + own] = xown;
        st.rec_a[ridx * nk + own] = a;
        st.rec_fx[ridx * nk + own] = out;
      }
    } else {
      const double phi = TWO ? phi2_of(hd, e) : phi0;
      double sc = out - (EBUF ? fo[kk] : st.fphi[own]);
      sc = sc * hh;
      sc = sc * phi;
      if constexpr (EBUF)
        tstore(xown + sc, r_x, lo, co);
      else
        st.x[own] = xown + sc;
      if (record) st.rec_fa[ridx * nk + own] = out;
    }
  }
}

// One stage for a row-pair task (one-field kernels without dense-row modes):
// the two rows' coupling in one merged walk, both rows' own state (and F(x))
// issued before the workgroup barrier every wave of the workgroup takes once,
// then the two epilogues.
template <bool STAGE2, int KT>
__device__ __forceinline__ void stage_pair(const prh_operator& op, const prh_schedule& sch,
                                           const prh_state& st, const int* __restrict__ tk,
                                           int step, int ridx, bool record, int w, int k0,
                                           int lane, bool active, const double* __restrict__ xin,
                                           const double* __restrict__ mv) {
  const int da = tk[0];
  const int b0 = tk[1];
  const int b1 = b0 + (tk[5] >> PRH_TASK_PAIR_SHIFT);
  double accA[KT], accB[KT];
  couple_pair<KT>(op, xin, mv, b0, b1, tk[2], k0, lane, active, accA, accB);
  const int K = op.K;
  const int d = op.d;
  const int pb = (int)((int64_t)K * d * 8);
  const int lb = (active ? lane : d - 1) * 8;
  double xo[2][KT], fo[2][KT];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int64_t rowoff = (int64_t)(da + r) * K * d;
    const auto r_in = tile_rsrc(xin + rowoff, pb);
    const auto r_f = tile_rsrc(st.fphi + rowoff, pb);
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) {
      const int co = (k0 + kk < K ? k0 + kk : K - 1) * d * 8;
      const double v = tload(r_in, lb, co);
      xo[r][kk] = active ? v : 0.0;
      fo[r][kk] = STAGE2 ? tload(r_f, lb, co) : 0.0;
    }
  }
  __syncthreads();  // the part-reduction barrier of the other waves
  double sc0[KT];
#pragma unroll
  for (int kk = 0; kk < KT; ++kk) sc0[kk] = 0.0;
  stage_epilogue<STAGE2, KT, false>(op, sch, st, step, ridx, record, w, da, k0, lane, active,
                                    xin, accA, sc0, xo[0], fo[0]);
  stage_epilogue<STAGE2, KT, false>(op, sch, st, step, ridx, record, w, da + 1, k0, lane, active,
                                    xin, accB, sc0, xo[1], fo[1]);
}

template <bool STAGE2, int KT, bool TWO, int WGW, int DM>
__device__ __forceinline__ void stage_body(const prh_operator& op, const prh_schedule& sch,
                                           const prh_state& st, const int* __restrict__ tasks,
                                           int ntasks, int step_local, int use_ctl) {
  __shared__ double red[WGW][KT][WAVE];
  int step = step_local;
  int ridx = -1;
  if (use_ctl) {
    step = st.step_ctl[0] + step_local;
    if (step >= st.step_ctl[1]) return;  // uniform over the grid
    ridx = step_local;
  } else {
    ridx = step - st.rec_first;
  }
  const bool record = st.rec_x != nullptr && ridx >= 0 && ridx < st.rec_count;
  const int lane = threadIdx.x & (WAVE - 1);
  // the two-field kernels with 8 or 16 right-hand sides per wave keep the
  // vector form: with a scalar wave index the compiler pipelines their wider
  // batches past the register budget (spills)
  const int wv = (TWO && KT > 4) ? (int)(threadIdx.x / WAVE) : wave_id();
  const int K = op.K;
  const int d = op.d;
  int gx, gy;
  wg_coords(ntasks, K, KT, gx, gy, WGW);
  const int k0 = gy * KT;
  const int* tk = tasks + 6 * ((int64_t)gx * WGW + wv);
  const int da = tk[0];
  const bool active = lane < d;

  const double* __restrict__ xin = STAGE2 ? st.a : st.x;
  const int w = sch.win[step];
  const double* __restrict__ mvC = op.mvals + (int64_t)w * op.mvals_stride;
  const double* __restrict__ mvE = TWO ? op.mvals_e + (int64_t)w * op.mvals_stride : nullptr;

  if constexpr (!TWO && DM == DM_NONE && KT <= 8) {
    if (da >= 0 && (tk[5] & PRH_TASK_PAIR)) {
      stage_pair<STAGE2, KT>(op, sch, st, tk, step, ridx, record, w, k0, lane, active, xin, mvC);
      return;
    }
  }
  double acc[KT], accE[KT];
  if (da >= 0) {
    couple_part<KT, TWO>(op, xin, mvC, mvE, tk[1], tk[2], tk[5] & PRH_TASK_KIND_MASK, k0, lane,
                         active, acc, accE);
  } else {
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) acc[kk] = accE[kk] = 0.0;
  }
  double sc_[KT];
#pragma unroll
  for (int kk = 0; kk < KT; ++kk) {
    const int k = k0 + kk < K ? k0 + kk : K - 1;
    sc_[kk] = TWO ? op.escale[k] : 0.0;
    if (TWO) acc[kk] = acc[kk] + sc_[kk] * accE[kk];
  }
  if (DM == DM_FINISH && da >= 0) {
    // dense row: this part's share of the mother-chunk partial sums, in chunk
    // order (the parts' shares are then added in part order)
    const int r = op.dense_row[da];
    const int np = tk[4];
    const int pi = wv - tk[3];
    const int c1 = (int)((int64_t)(pi + 1) * op.ndchunk / np);
    // the loads of CB chunks issued together, then added in chunk order
    constexpr int CB = 16 / KT > 0 ? 16 / KT : 1;
    for (int cb = (int)((int64_t)pi * op.ndchunk / np); cb < c1; cb += CB) {
      double pv[CB][KT];
#pragma unroll
      for (int u = 0; u < CB; ++u) {
        const int c = cb + u < c1 ? cb + u : c1 - 1;
        const double* dp = st.dpart + ((int64_t)c * op.ndense + r) * K * d;
#pragma unroll
        for (int kk = 0; kk < KT; ++kk) {
          const int k = k0 + kk;
          pv[u][kk] = (k < K && active) ? dp[(int64_t)k * d + lane] : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < CB; ++u) {
        if (cb + u < c1) {
#pragma unroll
          for (int kk = 0; kk < KT; ++kk) acc[kk] += pv[u][kk];
        }
      }
    }
  }
  // one-field kernels: this row's [K][d] panels of the stage arrays as
  // buffers (column offset in the scalar offset, bin offset per lane); a row's
  // lead wave issues its own state (and, in stage 2, F(x)) for all KT columns
  // before the part reduction so the loads overlap the wait
  constexpr bool EBUF = !TWO || KT <= 4;
  const bool erow = da >= 0 && wv == tk[3] && !(DM == DM_CHUNKS && (tk[5] & PRH_TASK_TO_DPART));
  const int64_t rowoff = (int64_t)(da >= 0 ? da : 0) * K * d;
  const int pb = (int)((int64_t)K * d * 8);
  const int lb = (active ? lane : d - 1) * 8;
  const auto r_in = tile_rsrc(xin + rowoff, pb);
  const auto r_f = tile_rsrc(st.fphi + rowoff, pb);
  double xo[KT], fo[KT];
  if constexpr (EBUF) {
    if (erow) {
#pragma unroll
      for (int kk = 0; kk < KT; ++kk) {
        const int co = (k0 + kk < K ? k0 + kk : K - 1) * d * 8;
        const double v = tload(r_in, lb, co);
        xo[kk] = active ? v : 0.0;
        fo[kk] = STAGE2 ? tload(r_f, lb, co) : 0.0;
      }
    }
  }
  reduce_parts<KT>(red, wv, lane, tk[3], tk[4], acc);
  if (DM == DM_CHUNKS && da >= 0 && (tk[5] & PRH_TASK_TO_DPART)) {
    // one mother chunk of a dense row: store its partial sum (chunk tk[3])
    double* dp = st.dpart + ((int64_t)tk[3] * op.ndense + op.dense_row[da]) * K * d;
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) {
      const int k = k0 + kk;
      if (k < K && active) dp[(int64_t)k * d + lane] = acc[kk];
    }
    return;
  }
  if (da < 0 || wv != tk[3]) return;
  if (DM == DM_PASS) {
    // dense rows: add the mother-chunk partial sums of dense_pass in chunk order
    const int r = op.dense_row[da];
    if (r >= 0) {
#pragma unroll
      for (int kk = 0; kk < KT; ++kk) {
        const int k = k0 + kk;
        if (k >= K || !active) continue;
        double sum = 0.0;
#pragma unroll 8
        for (int c = 0; c < op.ndchunk; ++c)
          sum += st.dpart[(((int64_t)c * op.ndense + r) * K + k) * d + lane];
        acc[kk] = acc[kk] + sum;
      }
    }
  }

  stage_epilogue<STAGE2, KT, TWO>(op, sch, st, step, ridx, record, w, da, k0, lane, active, xin,
                                  acc, sc_, xo, fo);
}

template <bool STAGE2, int KT, bool TWO, int WGW, int DM>
__global__ void __launch_bounds__(WGW * WAVE)
    etd2_stage(prh_operator op, prh_schedule sch, prh_state st, int step_local, int use_ctl) {
  stage_body<STAGE2, KT, TWO, WGW, DM>(op, sch, st, op.tasks, op.ntasks, step_local, use_ctl);
}

// Epilogues of the dense rows whose mother chunks the stage kernel stored in
// dpart: the row's nparts waves each add a share of the chunks, in order.
template <bool STAGE2, int KT, int WGW>
__global__ void __launch_bounds__(WGW * WAVE) dense_finish(prh_operator op, prh_schedule sch,
                                                           prh_state st, int step_local,
                                                           int use_ctl) {
  stage_body<STAGE2, KT, false, WGW, DM_FINISH>(op, sch, st, op.fin_tasks, op.nfin, step_local,
                                                use_ctl);
}

// ---------------------------------------------------------------------------
// Adjoint ETD2 stages (no counterpart in the reference).  Forward step:
//   a = e x + h phi1 F(x),  x' = a + h phi2 L (a - x),  F(y) = L y + b,
//   L y = dl (M y + D (kappa y)).
// Given lam' = dJ/dx' and u = h phi2 lam':
//   stage 1: lam_a = lam' + L^T u,  v = h phi1 lam_a - u,
//            dJ/db += h phi1 lam_a (source accumulators),
//            dJ/ds += dl <(M_E + D kappa_E)^T u, a> + <dJ/dd, dd/ds>;
//   stage 2: lam = e lam_a + L^T v,  dJ/ds += dl <(M_E + D kappa_E)^T v, x>,
//            and u for the previous step.
// L^T y = dl (M^T y + kappa (D^T y)); the operator passed here is the
// transposed block layout and the transposed FD stencil.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double dphi1(double z, double e, double phi1) {
  // d phi1 / dz = (e - phi1) / z, series near 0
  return (fabs(z) < 1e-2) ? 0.5 + z * (1.0 / 3.0 + z * (1.0 / 8.0 + z * (1.0 / 30.0)))
                          : (e - phi1) / z;
}
__device__ __forceinline__ double dphi2(double z, double phi1, double phi2) {
  // d phi2 / dz = (phi1 - 2 phi2) / z, series near 0
  return (fabs(z) < 1e-2) ? 1.0 / 6.0 + z * (1.0 / 12.0 + z * (1.0 / 40.0 + z * (1.0 / 180.0)))
                          : (phi1 - 2.0 * phi2) / z;
}

template <bool ASTAGE2, int KT, bool TWO>
__global__ void __launch_bounds__(STHREADS) etd2_adjoint(prh_operator op, prh_schedule sch,
                                                         prh_adjoint adj, int step) {
  __shared__ double red[SWG][KT][WAVE];
  const int lane = threadIdx.x & (WAVE - 1);
  const int wv = (TWO && KT > 4) ? (int)(threadIdx.x / WAVE) : wave_id();  // see stage_body
  const int K = op.K;
  const int d = op.d;
  int gx, gy;
  wg_coords(op.ntasks, K, KT, gx, gy);
  const int k0 = gy * KT;
  const int* tk = op.tasks + 6 * ((int64_t)gx * SWG + wv);
  const int r = tk[0];
  const bool active = lane < d;
  const double* __restrict__ yin = ASTAGE2 ? adj.v : adj.u;
  const int w = sch.win[step];
  const double* __restrict__ mvC = op.mvals + (int64_t)w * op.mvals_stride;
  const double* __restrict__ mvE = TWO ? op.mvals_e + (int64_t)w * op.mvals_stride : nullptr;

  double acc[KT], accE[KT];
  if (r >= 0) {
    couple_part<KT, TWO, true>(op, yin, mvC, mvE, tk[1], tk[2], tk[5], k0, lane, active, acc,
                               accE);
  } else {
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) acc[kk] = accE[kk] = 0.0;
  }
  reduce_parts<KT>(red, wv, lane, tk[3], tk[4], acc);
  if (TWO) {
    __syncthreads();  // the lead waves have read red before it is reused
    reduce_parts<KT>(red, wv, lane, tk[3], tk[4], accE);
  }
  if (r < 0 || wv != tk[3]) return;

  const double dl = sch.dldz[step];
  const double hh = sch.h[step];
  const bool pdiff = (op.flags & PRH_F_PARTIAL_DIFF) != 0;
  double dcf[7];
#pragma unroll
  for (int o = 0; o < 7; ++o) dcf[o] = (active && pdiff) ? op.dcoef[lane * 7 + o] : 0.0;
  const int64_t nk = (int64_t)op.nspec * K * d;
  const int ridx = step - adj.rec_first;
  const bool want_s = TWO && adj.gdot != nullptr;
  const bool inj = !ASTAGE2 && adj.gsrc0 != nullptr && (op.flags & PRH_F_INJECTION) != 0 &&
                   op.src_slot[r] >= 0;
  // row-local inputs of this step (and of the previous one for u), loaded once
  RowIn ri{}, rip{};
  if (active) ri = row_in<TWO>(op, sch, step, w, r, lane);
  if (ASTAGE2 && step > 0 && active) rip = row_in<TWO>(op, sch, step - 1, sch.win[step - 1], r, lane);
  // up to 4 columns per wave: the row's own adjoint state and lam (stage 1) /
  // lam_a (stage 2) for all columns, issued together before the column loop
  // (with 8 the registers are spoken for)
  constexpr bool PRE = KT <= 4;
  double ypre[KT], lpre[KT];
  if constexpr (PRE) {
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) {
      const int kc = k0 + kk < K ? k0 + kk : K - 1;
      const int64_t own = ((int64_t)r * K + kc) * d + lane;
      ypre[kk] = active ? yin[own] : 0.0;
      lpre[kk] = active ? (ASTAGE2 ? adj.lam_a[own] : adj.lam[own]) : 0.0;
    }
  }

#pragma unroll
  for (int kk = 0; kk < KT; ++kk) {
    const int k = k0 + kk;
    if (k >= K) break;
    const double s = TWO ? op.escale[k] : 0.0;
    RowCoef rc{0.0, 0.0, 0.0, 0.0};
    if (active) rc = row_coef_from<TWO>(ri, s, pdiff);
    const int64_t own = ((int64_t)r * K + k) * d + lane;
    const double yown = PRE ? ypre[kk] : (active ? yin[own] : 0.0);
    // transposed FD stencil D^T y at this row, then kappa-scaled
    double q = 0.0;
    if (pdiff) {
      // neighbours yown[lane + o], o = -3..3 (zero outside [0, d)), added in
      // the order o = -3, -2, -1, 1, 2, 3
      const double ym1 = lane_prev(yown);
      const double ym2 = lane_prev(ym1);
      const double ym3 = lane_prev(ym2);
      q += dcf[0] * ym3;
      q += dcf[1] * ym2;
      q += dcf[2] * ym1;
      double yo = yown;
#pragma unroll
      for (int o = 1; o <= 3; ++o) {
        yo = lane_next(yo);
        q += dcf[o + 3] * yo;
      }
    }
    const double mty = TWO ? acc[kk] + s * accE[kk] : acc[kk];
    const double t = (mty + rc.kap * q) * dl;
    if (!active) continue;
    const double hd = rc.dval * hh;
    const double e = exp(hd);
    const double p1 = phi1_of(hd, e);
    const double p2 = phi2_of(hd, e);
    double gs = 0.0;
    if (!ASTAGE2) {
      const double lamp = PRE ? lpre[kk] : adj.lam[own];
      const double lam_a = lamp + t;
      const double wv_ = hh * p1 * lam_a;
      adj.lam_a[own] = lam_a;
      adj.v[own] = wv_ - yown;
      if (inj) {
        const int slot = op.src_slot[r];
        const double f = sch.finj[step];
        const double lz = adj.lnz1[step];
        for (int j = 0; j < op.nsrc; ++j) {
          const double c = sch.src_coef[((int64_t)step * op.nsrc + j) * K + k];
          const int64_t gi = (((int64_t)slot * op.nsrc + j) * K + k) * d + lane;
          adj.gsrc0[gi] += f * c * wv_;
          adj.gsrc1[gi] += f * c * lz * wv_;
        }
      }
      if (want_s) {
        const double xa = adj.rec_a[ridx * nk + own];
        const double xx = adj.rec_x[ridx * nk + own];
        const double fx = adj.rec_fx[ridx * nk + own];
        const double fa = adj.rec_fa[ridx * nk + own];
        gs = dl * (accE[kk] + rc.kapE * q) * xa;
        const double dJdd = lam_a * (hh * e * xx + hh * hh * dphi1(hd, e, p1) * fx) +
                            lamp * (hh * hh * dphi2(hd, p1, p2) * (fa - fx));
        gs += dJdd * rc.dE;
      }
    } else {
      const double lam_a = PRE ? lpre[kk] : adj.lam_a[own];
      const double lam = e * lam_a + t;
      adj.lam[own] = lam;
      if (want_s) {
        const double xx = adj.rec_x[ridx * nk + own];
        gs = dl * (accE[kk] + rc.kapE * q) * xx;
      }
      // u of the previous step: h_{k-1} phi2(h_{k-1} d_{k-1}) lam
      if (step > 0) {
        const int sp = step - 1;
        const RowCoef rp = row_coef_from<TWO>(rip, s, pdiff);
        const double hdp = rp.dval * sch.h[sp];
        const double ep = exp(hdp);
        adj.u[own] = sch.h[sp] * phi2_of(hdp, ep) * lam;
      }
    }
    if (want_s) adj.gdot[own] += gs;
  }
}

// u = h phi2(h d) lam for the last step of the sweep (start of the adjoint).
template <bool TWO>
__global__ void __launch_bounds__(STHREADS) adjoint_seed(prh_operator op, prh_schedule sch,
                                                         prh_adjoint adj, int step) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int wv = wave_id();
  const int K = op.K;
  const int d = op.d;
  const int* tk = op.tasks + 6 * ((int64_t)blockIdx.x * SWG + wv);
  const int r = tk[0];
  if (r < 0 || wv != tk[3] || lane >= d) return;
  const bool pdiff = (op.flags & PRH_F_PARTIAL_DIFF) != 0;
  const int w = sch.win[step];
  for (int k = 0; k < K; ++k) {
    const double s = TWO ? op.escale[k] : 0.0;
    const RowCoef rc = row_coef<TWO>(op, sch, step, w, r, lane, s, pdiff);
    const double hd = rc.dval * sch.h[step];
    const int64_t own = ((int64_t)r * K + k) * d + lane;
    adj.u[own] = sch.h[step] * phi2_of(hd, exp(hd)) * adj.lam[own];
  }
}

__global__ void step_ctl_set(int32_t* ctl, int32_t first, int32_t end) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    ctl[0] = first;
    ctl[1] = end;
  }
}

__global__ void step_ctl_advance(int32_t* ctl, int32_t chunk) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctl[0] = ctl[0] + chunk;
}

// Step of a launch: step_local itself, or step_ctl[0] + step_local under graph
// replay (use_ctl), where steps at or beyond step_ctl[1] do nothing.
__device__ __forceinline__ void resolve_step(const prh_state& st, int step_local, int use_ctl,
                                             int& step, bool& live) {
  step = step_local;
  live = true;
  if (use_ctl) {
    step = st.step_ctl[0] + step_local;
    live = step < st.step_ctl[1];
  }
}

// ---------------------------------------------------------------------------
// EM sector (photons, e+-): two launches per step after stage 2 -- the
// Bethe-Heitler pair weights, then the photon ETD2 step and the cascade
// transfer as one product with the per-step matrix T_gamma P(t).
// ---------------------------------------------------------------------------

// Bethe-Heitler weights w[j][k] = bhw[win][j] * sum_s cbh[s] x[s][k][j]:
// workgroup = one right-hand side, lanes = bins, so every species panel is one
// coalesced 512-byte load; the 4 waves sum interleaved quarters of the species
// (8 loads in flight each) and combine them in wave order.
__global__ void __launch_bounds__(256) em_bh_weights_bins(prh_operator op, prh_schedule sch,
                                                          prh_state st, prh_em em,
                                                          int step_local, int use_ctl) {
  __shared__ double part[4][WAVE];
  int step;
  bool live;
  resolve_step(st, step_local, use_ctl, step, live);
  if (!live) return;
  const int lane = threadIdx.x & (WAVE - 1);
  const int wv = wave_id();
  const int K = op.K;
  const int d = op.d;
  const int k = blockIdx.x;
  const bool act = lane < d;
  const int ns = op.nspec;
  const int64_t sS = (int64_t)K * d;
  const double* xp = st.x + (int64_t)k * d + (act ? lane : 0);
  double acc = 0.0;
  int s = wv;
  for (; s + 28 < ns; s += 32) {
    double xv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) xv[u] = act ? xp[(int64_t)(s + 4 * u) * sS] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += em.cbh[s + 4 * u] * xv[u];
  }
  for (; s < ns; s += 4) acc += em.cbh[s] * (act ? xp[(int64_t)s * sS] : 0.0);
  part[wv][lane] = acc;
  __syncthreads();
  if (wv != 0 || !act) return;
  const double tot = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
  const int w = sch.win[step];
  em.wbh[(int64_t)lane * K + k] = em.bhw[(int64_t)w * d + lane] * tot;
}

// Per-step photon transfer of the shared-schedule EM step:
//   qstep[t] = T_gamma[win(t)] P(t),
// P(t) the photon ETD2 step of step t as a matrix.  The photon stages are
// linear in the photons g:  a = e g + h phi1 F g,  g2 = a + h phi2 (F a - F g)
// with F the FD stencil on kappa g (F[i][j] = dl dcoef[i][j - i] kappa_j,
// kappa = H cm2sec E, 0 < |j - i| <= 3) and e, phi1, phi2 of h d, d = dl kappa
// ddiag, so  P = A + h diag(phi2) F (A - I),  A = diag(e) + h diag(phi1) F,
// a band of 13 diagonals.  One workgroup per step: the band of P in LDS, then
// every entry of T_gamma P (13 terms each).
__global__ void __launch_bounds__(256) em_qstep_kernel(prh_operator op, prh_schedule sch,
                                                      prh_em em, double* __restrict__ qstep) {
  __shared__ double pb[PRH_EM_MAX][13];
  __shared__ double se[PRH_EM_MAX], sp1[PRH_EM_MAX], sp2[PRH_EM_MAX], skap[PRH_EM_MAX];
  const int t = blockIdx.x;
  const int de = em.d_em;
  const double hh = sch.h[t];
  const double dl = sch.dldz[t];
  const double had = sch.hadia[t];
  const bool pdiff = (op.flags & PRH_F_PARTIAL_DIFF) != 0;
  for (int i = threadIdx.x; i < de; i += blockDim.x) {
    const double kap = had * em.egrid[i];
    double dv = 0.0;
    if (pdiff) {
      double kx = kap * em.ddiag[i];
      dv = kx * dl;
    }
    const double hd = dv * hh;
    const double e = exp(hd);
    se[i] = e;
    sp1[i] = phi1_of(hd, e);
    sp2[i] = phi2_of(hd, e);
    skap[i] = kap;
  }
  __syncthreads();
  // F[a][b] (zero off the stencil and without the partial-differential term)
  auto F = [&](int a, int b) -> double {
    const int o = b - a;
    if (!pdiff || o == 0 || o < -3 || o > 3 || b < 0 || b >= de) return 0.0;
    return dl * (em.dcoef[a * 7 + o + 3] * skap[b]);
  };
  for (int r = threadIdx.x; r < de * 13; r += blockDim.x) {
    const int m = r / 13, o = r % 13 - 6, j = m + o;
    double v = 0.0;
    if (j >= 0 && j < de) {
      v = (o == 0 ? se[m] : 0.0) + hh * sp1[m] * F(m, j);
      double s = 0.0;
      for (int on = -3; on <= 3; ++on) {
        const int n = m + on;
        if (on == 0 || n < 0 || n >= de) continue;
        const int oj = j - n;
        if (oj < -3 || oj > 3) continue;
        const double amd = oj == 0 ? se[n] - 1.0 : hh * sp1[n] * F(n, j);
        s += F(m, n) * amd;
      }
      v += hh * sp2[m] * s;
    }
    pb[m][o + 6] = v;
  }
  __syncthreads();
  const int w = sch.win[t];
  const double* T = em.tg + (int64_t)w * de * de;
  double* Q = qstep + (int64_t)t * de * de;
  for (int r = threadIdx.x; r < de * de; r += blockDim.x) {
    const int i = r / de, j = r % de;
    const int m0 = j - 6 > 0 ? j - 6 : 0, m1 = j + 6 < de - 1 ? j + 6 : de - 1;
    double q = 0.0;
    for (int m = m0; m <= m1; ++m) q += T[(int64_t)i * de + m] * pb[m][j - m + 6];
    Q[r] = q;
  }
}

// The EM step's transfer as one FP64 MFMA product per 16 x 16 tile:
//   out[i][k] = sum_c S[i][c] y[c][k],  S = [qstep[t] | 2 T_electron R_BH],
//   y = [photons before the step; Bethe-Heitler weights wbh],
// contraction length d_em + d split over the 4 waves of the tile's
// workgroup (summed through LDS in wave order).  The photon state alternates:
// step t reads gam and writes gam2 for even t, the reverse for odd t.
__global__ void __launch_bounds__(256) em_step_mfma(prh_operator op, prh_schedule sch,
                                                   prh_state st, prh_em em, int step_local,
                                                   int use_ctl) {
  __shared__ double red[4][4][WAVE];
  int step;
  bool live;
  resolve_step(st, step_local, use_ctl, step, live);
  if (!live) return;
  const int K = op.K;
  const int de = em.d_em;
  const int d = op.d;
  const int kd = de + d;
  const int lane = threadIdx.x & (WAVE - 1);
  const int wv = wave_id();
  const int nct = (K + 15) / 16;
  const int rt = blockIdx.x / nct, ct = blockIdx.x % nct;
  const int col = lane & 15, kq = lane >> 4;
  const int i = rt * 16 + col;
  const int ic = i < de ? i : de - 1;
  const int kc = ct * 16 + col;
  const int kcc = kc < K ? kc : K - 1;
  const int w = sch.win[step];
  const double* Qr = em.qstep + ((int64_t)step * de + ic) * de;
  const double* Sr = em.tstack + ((int64_t)w * de + ic) * kd;
  const double* gin = (step & 1) ? em.gam2 : em.gam;
  double* gout = (step & 1) ? em.gam : em.gam2;
  // this wave's k-steps (4 contraction indices each), loads first
  constexpr int KS = 13;
  const int nks = (kd + 3) / 4;
  const int s0 = wv * nks / 4, s1 = (wv + 1) * nks / 4;
  double a[KS], b[KS];
#pragma unroll
  for (int u = 0; u < KS; ++u) {
    const int c = 4 * (s0 + u) + kq;
    a[u] = 0.0;
    b[u] = 0.0;
    if (s0 + u < s1 && i < de && c < kd) a[u] = c < de ? Qr[c] : Sr[c];
    if (s0 + u < s1 && c < kd) b[u] = c < de ? gin[(int64_t)c * K + kcc] : em.wbh[(int64_t)(c - de) * K + kcc];
  }
  double4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int u = 0; u < KS; ++u)
    if (s0 + u < s1) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], acc, 0, 0, 0);
#pragma unroll
  for (int v = 0; v < 4; ++v) red[wv][v][lane] = acc[v];
  __syncthreads();
  if (wv != 0 || kc >= K) return;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int row = rt * 16 + kq + 4 * v;
    const double sum = ((red[0][v][lane] + red[1][v][lane]) + red[2][v][lane]) + red[3][v][lane];
    if (row < de) gout[(int64_t)row * K + kc] = sum;
  }
}

// [T_gamma | 2 T_electron R_BH] per window (one thread per output value,
// the T_electron R_BH sum over the EM bins in order).
__global__ void em_stack_kernel(const double* __restrict__ tg, const double* __restrict__ te,
                                const double* __restrict__ rbh, int nwin, int de, int d,
                                double* __restrict__ out) {
  const int kd = de + d;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nwin * de * kd) return;
  const int64_t w = t / ((int64_t)de * kd);
  const int rem = (int)(t - w * de * kd);
  const int i = rem / kd, c = rem % kd;
  if (c < de) {
    out[t] = tg[(w * de + i) * de + c];
    return;
  }
  const int j = c - de;
  const double* ter = te + (w * de + i) * de;
  const double* rb = rbh + w * de * d + j;
  double s = 0.0;
  for (int m = 0; m < de; ++m) s += ter[m] * rb[(int64_t)m * d];
  out[t] = s + s;
}

void launch_em(const prh_operator& op, const prh_schedule& sch, const prh_state& st,
               int step_local, int use_ctl, hipStream_t s) {
  const prh_em& em = *st.em;
  hipLaunchKernelGGL(em_bh_weights_bins, dim3((unsigned)op.K), dim3(256), 0, s, op, sch, st, em,
                     step_local, use_ctl);
  const int64_t tiles = (int64_t)((em.d_em + 15) / 16) * ((op.K + 15) / 16);
  hipLaunchKernelGGL(em_step_mfma, dim3((unsigned)tiles), dim3(256), 0, s, op, sch, st, em,
                     step_local, use_ctl);
}

int32_t check_uneven(const prh_operator* op, const prh_uneven* un, const prh_state* st,
                     const char* who) {
  const std::string w(who);
  if (!op || !un || !st) return set_error(w + ": null argument");
  if (op->d <= 0 || op->d > WAVE) return set_error(w + ": d must be in [1, 64]");
  if (op->K <= 0) return set_error(w + ": K must be >= 1");
  if (op->ntasks <= 0 || op->ntasks % SWG != 0)
    return set_error(w + ": ntasks must be a positive multiple of 16");
  if (!(op->kt == 1 || op->kt == 2 || op->kt == 4)) return set_error(w + ": kt must be 1, 2 or 4");
  if (op->mvals_e) return set_error(w + ": one-field operators only");
  if (op->wg_waves == 8) return set_error(w + ": the uneven kernel runs 16-wave workgroups");
  if (op->ndense > 0) return set_error(w + ": the uneven kernel has no dense-row pass");
  if (op->flags & PRH_F_ROW_PAIRS) return set_error(w + ": the uneven kernel has no row-pair tasks");
  if (!un->h || !un->dldz || !un->hadia || !un->finj || !un->win || !un->rpair || !un->mrhs)
    return set_error(w + ": null schedule buffer");
  if (un->nslot != 1 && un->nslot != 2) return set_error(w + ": nslot must be 1 or 2");
  if ((int64_t)op->mvals_stride * 8 > INT32_MAX || (int64_t)op->K * op->d * 8 > INT32_MAX)
    return set_error(w + ": value rows and state panels must stay below 2 GiB");
  if ((op->flags & PRH_F_INJECTION) && !un->src_coef) return set_error(w + ": injection needs src_coef");
  return 0;
}

int32_t validate_em(const prh_operator* op, const prh_state* st, const char* who) {
  if (!st->em) return 0;
  const prh_em* em = st->em;
  if (em->d_em < 1 || em->d_em > PRH_EM_MAX)
    return set_error(std::string(who) + ": d_em must be in [1, 144]");
  if (!em->egrid || !em->dcoef || !em->ddiag || !em->cbh || !em->bhw || !em->rbh || !em->tg ||
      !em->te || !em->gam || !em->gam2 || !em->wbh || !em->tstack || !em->qstep)
    return set_error(std::string(who) + ": EM sector with a null buffer");
  if (em->nwin < 1) return set_error(std::string(who) + ": EM sector without windows");
  (void)op;
  return 0;
}

// ---------------------------------------------------------------------------
// Uneven batch: every right-hand side has its own z grid (its own z_max), hence
// its own refresh windows, coupling values and per-step scalars; columns
// retire after their last step.  Narrow layout [species][rhs][bin], lanes = bins.
// ---------------------------------------------------------------------------

// Photon-field tables for a list of (z, s): out[w][i] = sum_m W[i][m] *
// (n_CMB(P[i][m], z_w) + s_w (1 + z_w)^zpow E0[i][m]) -- the CMB Planck
// spectrum evaluated here (photonfields.py:114-126) plus a z-separable component
// tabulated at z = 0.  One thread per (w, i).
__global__ void field_tables_kernel(const double* __restrict__ zs, const double* __restrict__ ss,
                                    int nw, const double* __restrict__ P,
                                    const double* __restrict__ W, const double* __restrict__ E0,
                                    int ni, int nm, double planck_pref, double e_cmb, double zpow,
                                    double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nw * ni) return;
  const int w = (int)(t / ni);
  const int i = (int)(t % ni);
  const double z = zs[w];
  const double s = ss[w];
  const double zp1 = 1.0 + z;
  const double pref = planck_pref * (zp1 * zp1);
  const double sz = s * pow(zp1, zpow);
  double acc = 0.0;
  for (int m = 0; m < nm; ++m) {
    const int64_t q = (int64_t)i * nm + m;
    const double Ered = P[q] / zp1;
    const double ratio = Ered / e_cmb;
    double n = 0.0;
    if (ratio < 709.0) n = pref * (Ered * Ered) / (exp(ratio) - 1.0);
    n = n + sz * E0[q];
    acc += W[q] * n;
  }
  out[t] = acc;
}

// Rate fold of listed columns: out[orow[c]][r] = sum_j kern[r][j] phot[irow[c]][j]
// (the fold_gemm_mfma tiling with gathered input and scattered output rows).
__global__ void __launch_bounds__(FOLD_THREADS) fold_rows_mfma(const double* __restrict__ kern,
                                                               int64_t nrows, int dph,
                                                               const double* __restrict__ phot,
                                                               const int32_t* __restrict__ irow,
                                                               const int32_t* __restrict__ orow,
                                                               int ncol, double* __restrict__ out,
                                                               int64_t out_stride) {
  // the gathered photon rows of 32 listed columns at a time, staged in LDS
  // once per workgroup as in fold_gemm_mfma; two 16-column MFMA chains per wave
  __shared__ double sp[32 * (4 * FOLD_MAX_KSTEPS + 1)];
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t wave = (int64_t)blockIdx.x * FOLD_WAVES + wave_id();
  const int64_t r0 = wave * 16;
  const bool rows = r0 < nrows;  // waves past the last row still take part in the barriers
  const int col = lane & 15;
  const int kq = lane >> 4;
  const int64_t r = r0 + col;
  const int nk = (dph + 3) / 4;
  const int ps = dph + 1;
  double bfrag[FOLD_MAX_KSTEPS];
#pragma unroll
  for (int t = 0; t < FOLD_MAX_KSTEPS; ++t) {
    const int j = 4 * t + kq;
    bfrag[t] = (rows && t < nk && j < dph && r < nrows) ? kern[r * dph + j] : 0.0;
  }
  for (int c0 = 0; c0 < ncol; c0 += 32) {
    __syncthreads();  // the previous columns are consumed
    for (int i = threadIdx.x; i < 32 * dph; i += FOLD_THREADS) {
      const int cr = i / dph, j = i % dph;
      sp[cr * ps + j] = c0 + cr < ncol ? phot[(int64_t)irow[c0 + cr] * dph + j] : 0.0;
    }
    __syncthreads();
    if (!rows) continue;
    double4_t acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = acc0;
#pragma unroll
    for (int t = 0; t < FOLD_MAX_KSTEPS; ++t) {
      if (t < nk) {
        const int j = 4 * t + kq;
        const double a0 = j < dph ? sp[col * ps + j] : 0.0;
        const double a1 = j < dph ? sp[(16 + col) * ps + j] : 0.0;
        acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bfrag[t], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bfrag[t], acc1, 0, 0, 0);
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int c = c0 + kq + 4 * v;
      if (c < ncol && r < nrows) out[(int64_t)orow[c] * out_stride + r] = acc0[v];
      if (c + 16 < ncol && r < nrows) out[(int64_t)orow[c + 16] * out_stride + r] = acc1[v];
    }
  }
}

// Coupling product of one row part for KT columns of the uneven batch, each
// with its own window values (vbase[kk]: the column's value row in mrhs):
//   acc[kk] = sum_b m_{b,k} (shifted) y_b[k]
// in block order, as couple_part.  Runs of diagonal blocks issue the loads of
// DIAG blocks before any product, a band block those of half its diagonals,
// through buffer resources (one per column's value row, one per mother panel).
// Two diagonal blocks per batch at KT = 2: 4.87 s for the K = 1024 march
// against 4.93 s with six.
template <int KT>
struct UnevenBatch {
  static constexpr int DIAG = KT == 1 ? 4 : (KT == 2 ? 2 : 1);
};

template <int KT>
__device__ __forceinline__ void couple_cols(const prh_operator& op, const double* __restrict__ y,
                                            const double* const (&vbase)[KT], const int (&kc)[KT],
                                            int b_begin, int b_end, int kind, int lane,
                                            bool active, double (&acc)[KT]) {
  const int d = op.d;
  const int64_t Kd = (int64_t)op.K * d;
  const int lb = (active ? lane : d - 1) * 8;
  const int vbytes = (int)(op.mvals_stride * 8);
  const int pbytes = (int)(Kd * 8);
  __amdgpu_buffer_rsrc_t rv[KT];
  int coff[KT];
#pragma unroll
  for (int kk = 0; kk < KT; ++kk) {
    rv[kk] = tile_rsrc(vbase[kk], vbytes);
    coff[kk] = kc[kk] * d * 8;
    acc[kk] = 0.0;
  }
  const int dbase = b_begin;
  int dmo = 0, dol = 0, dns = 0, dvo = 0;
  if (b_begin + lane < b_end) {
    const int* blk = op.blocks + 4 * (b_begin + lane);
    dmo = blk[0];
    dol = blk[1];
    dns = blk[2];
    dvo = blk[3];
  }
  if (kind == PART_DIAG) {
    constexpr int UB = UnevenBatch<KT>::DIAG;
    for (int b = b_begin; b < b_end; b += UB) {
      double mv[UB][KT], xv[UB][KT];
#pragma unroll
      for (int u = 0; u < UB; ++u) {
        const int bb = b + u < b_end ? b + u : b_end - 1;
        int mo, voffi;
        if (bb - dbase < WAVE) {
          mo = __builtin_amdgcn_readlane(dmo, bb - dbase);
          voffi = __builtin_amdgcn_readlane(dvo, bb - dbase);
        } else {
          mo = op.blocks[4 * bb];
          voffi = op.blocks[4 * bb + 3];
        }
        const auto ry = tile_rsrc(y + (int64_t)mo * Kd, pbytes);
#pragma unroll
        for (int kk = 0; kk < KT; ++kk) {
          mv[u][kk] = tload(rv[kk], lb, voffi * 8);
          xv[u][kk] = tload(ry, lb, coff[kk]);
        }
      }
#pragma unroll
      for (int u = 0; u < UB; ++u) {
        if (b + u < b_end) {
#pragma unroll
          for (int kk = 0; kk < KT; ++kk) acc[kk] += mv[u][kk] * xv[u][kk];
        }
      }
    }
  } else if (kind == PART_BAND) {
    // one block at a time, its diagonals in two halves of NSMAX / 2 (loads of
    // both columns issued before the products of each half)
    constexpr int NH = NSMAX / 2;
    for (int b = b_begin; b < b_end; ++b) {
      int mo, ol, nsl, voffi;
      if (b - dbase < WAVE) {
        mo = __builtin_amdgcn_readlane(dmo, b - dbase);
        ol = __builtin_amdgcn_readlane(dol, b - dbase);
        nsl = __builtin_amdgcn_readlane(dns, b - dbase);
        voffi = __builtin_amdgcn_readlane(dvo, b - dbase);
      } else {
        mo = op.blocks[4 * b];
        ol = op.blocks[4 * b + 1];
        nsl = op.blocks[4 * b + 2];
        voffi = op.blocks[4 * b + 3];
      }
      const auto ry = tile_rsrc(y + (int64_t)mo * Kd, pbytes);
      double xv[KT];
#pragma unroll
      for (int kk = 0; kk < KT; ++kk) {
        const double v = tload(ry, lb, coff[kk]);
        xv[kk] = active ? v : 0.0;
      }
      if (ol >= 0 && ol != 0) {
#pragma unroll
        for (int kk = 0; kk < KT; ++kk) xv[kk] = shfl_or_zero(xv[kk], lane + ol, d);
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (h * NH >= nsl) break;
        double mv[KT][NH];
#pragma unroll
        for (int s = 0; s < NH; ++s) {
          const int s2 = h * NH + s;
          const int sc = s2 < nsl ? s2 : nsl - 1;
#pragma unroll
          for (int kk = 0; kk < KT; ++kk) mv[kk][s] = tload(rv[kk], lb, (voffi + sc * d) * 8);
        }
#pragma unroll
        for (int s = 0; s < NH; ++s) {
          const int s2 = h * NH + s;
          if (s2 < nsl) {
            if (ol >= 0) {
              // diagonal s2 reads lane + ol + s2: one shuffle to the first
              // diagonal (above), then one-lane DPP shifts
#pragma unroll
              for (int kk = 0; kk < KT; ++kk) acc[kk] += mv[kk][s] * xv[kk];
              if (s2 + 1 < nsl) {
#pragma unroll
                for (int kk = 0; kk < KT; ++kk) xv[kk] = lane_next(xv[kk]);
              }
            } else {
              const int src = lane + ol + s2;
#pragma unroll
              for (int kk = 0; kk < KT; ++kk) acc[kk] += mv[kk][s] * shfl_or_zero(xv[kk], src, d);
            }
          }
        }
      }
    }
  } else {
    for (int b = b_begin; b < b_end; ++b) {
      const int* blk = op.blocks + 4 * b;
      const int mo = blk[0];
      const int olo = blk[1];
      const int nslot = blk[2];
      const int voffi = blk[3];
      const auto ry = tile_rsrc(y + (int64_t)mo * Kd, pbytes);
      double xv[KT];
#pragma unroll
      for (int kk = 0; kk < KT; ++kk) {
        const double v = tload(ry, lb, coff[kk]);
        xv[kk] = active ? v : 0.0;
      }
      for (int s2 = 0; s2 < nslot; ++s2) {
        const int src = lane + olo + s2;
#pragma unroll
        for (int kk = 0; kk < KT; ++kk)
          acc[kk] += tload(rv[kk], lb, (voffi + s2 * d) * 8) * shfl_or_zero(xv[kk], src, d);
      }
    }
  }
}

// One stage of the uneven batch for the live columns [0, ncols) of step
// `step`.  Workgroup L runs on XCD L % 8; XCD x takes the column groups x,
// x + 8, ... and walks them for one row group before the next (4.87 vs 5.37 s
// for the K = 1024 march with row groups walked first; the coupling values are
// read once per stage either way).
template <bool STAGE2, int KT>
__global__ void __launch_bounds__(STHREADS, KT <= 2 ? 8 : 4)
    etd2_uneven(prh_operator op, prh_uneven un, prh_state st, int step, int ncols) {
  __shared__ double red[SWG][KT][WAVE];
  const int lane = threadIdx.x & (WAVE - 1);
  const int wv = wave_id();
  const int K = op.K;
  const int d = op.d;
  const int nx = op.ntasks / SWG;
  const int ny = (ncols + KT - 1) / KT;
  const int L = blockIdx.x;
  const int j = L >> 3;
  const int per = (ny + 7) / 8;
  const int gy = (L & 7) + 8 * (j % per);
  const int gx = j / per;
  if (gy >= ny) return;  // padding of the column groups to a multiple of 8
  const int k0 = gy * KT;
  const int kend = ncols;
  // per-column window of this step (-1: finished); a column outside the
  // launched range reads the group's first column (in bounds, result unused)
  int wk[KT], kc[KT];
  const double* vb[KT];
  bool any = false;
#pragma unroll
  for (int kk = 0; kk < KT; ++kk) {
    const int k = k0 + kk;
    wk[kk] = (k < kend) ? un.win[(int64_t)step * K + k] : -1;
    kc[kk] = wk[kk] >= 0 ? k : k0;
    const int slot = (un.nslot == 2 && wk[kk] >= 0) ? (wk[kk] & 1) : 0;
    vb[kk] = un.mrhs + ((int64_t)slot * K + kc[kk]) * op.mvals_stride;
    any = any || wk[kk] >= 0;
  }
  if (!any) return;  // uniform over the workgroup, before any barrier
  const int* tk = op.tasks + 6 * ((int64_t)gx * SWG + wv);
  const int da = tk[0];
  const bool active = lane < d;
  const double* __restrict__ xin = STAGE2 ? st.a : st.x;
  double acc[KT];
  if (da >= 0) {
    couple_cols<KT>(op, xin, vb, kc, tk[1], tk[2], tk[5] & PRH_TASK_KIND_MASK, lane, active, acc);
  } else {
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) acc[kk] = 0.0;
  }
  reduce_parts<KT>(red, wv, lane, tk[3], tk[4], acc);
  if (da < 0 || wv != tk[3]) return;

  const bool pdiff = (op.flags & PRH_F_PARTIAL_DIFF) != 0;
  const bool inj = (op.flags & PRH_F_INJECTION) != 0 && op.src_slot[da] >= 0;
  double dcf[7];
#pragma unroll
  for (int o = 0; o < 7; ++o) dcf[o] = (active && pdiff) ? op.dcoef[lane * 7 + o] : 0.0;
  const int slot = inj ? op.src_slot[da] : 0;
  const int J = op.nsrc;
  const double E = active ? op.egrid[lane] : 0.0;
  const double ddg = active ? op.ddiag[lane] : 0.0;
  const double psc = op.pscale[da];
#pragma unroll
  for (int kk = 0; kk < KT; ++kk) {
    const int k = k0 + kk;
    if (wk[kk] < 0) continue;  // uniform over the wave
    const int64_t sk = (int64_t)step * K + k;
    const double dl = un.dldz[sk];
    const double hh = un.h[sk];
    double kap = un.hadia[sk] * E;
    kap = kap + psc * (active ? un.rpair[(int64_t)wk[kk] * d + lane] : 0.0);
    const double mdc = active ? vb[kk][op.diag_off + (int64_t)da * d + lane] : 0.0;
    double dval = mdc * dl;
    if (pdiff) {
      double kxd = kap * ddg;
      kxd = kxd * dl;
      dval = dval + kxd;
    }
    const double hd = dval * hh;
    const double e = exp(hd);
    const int64_t own = ((int64_t)da * K + k) * d + lane;
    const double xown = active ? xin[own] : 0.0;
    double out = acc[kk];
    if (pdiff) {
      const double kx = kap * xown;
      // stencil neighbours kx[lane + o], o = -3..3 (zero outside [0, d)),
      // added in the order o = -3, -2, -1, 1, 2, 3 (a zero coefficient adds
      // an exact zero: no per-term test)
      const double km1 = lane_prev(kx);
      const double km2 = lane_prev(km1);
      const double km3 = lane_prev(km2);
      double dx = 0.0;
      dx += dcf[0] * km3;
      dx += dcf[1] * km2;
      dx += dcf[2] * km1;
      double kxo = kx;
#pragma unroll
      for (int o = 1; o <= 3; ++o) {
        kxo = lane_next(kxo);
        dx += dcf[o + 3] * kxo;
      }
      out = out + dx;
    }
    out = out * dl;
    if (inj && active) {
      double bsum = 0.0;
      for (int j2 = 0; j2 < J; ++j2) {
        const double c = un.src_coef[((int64_t)step * J + j2) * K + k];
        const double g = op.src_grid[(((int64_t)slot * J + j2) * K + k) * d + lane];
        bsum = (j2 == 0) ? c * g : bsum + c * g;
      }
      out = out + un.finj[sk] * bsum;
    }
    if (!active) continue;
    if (!STAGE2) {
      const double phi = phi1_of(hd, e);
      double a = e * xown;
      double sc = phi * out;
      sc = sc * hh;
      a = a + sc;
      st.a[own] = a;
      st.fphi[own] = out;
    } else {
      const double phi = phi2_of(hd, e);
      double sc = out - st.fphi[own];
      sc = sc * hh;
      sc = sc * phi;
      st.x[own] = xown + sc;
    }
  }
}

template <bool S2>
void launch_uneven(const prh_operator& op, const prh_uneven& un, const prh_state& st, int step,
                   int ncols, hipStream_t s) {
  const int64_t ny = (ncols + op.kt - 1) / op.kt;
  const int64_t nwg = (int64_t)(op.ntasks / SWG) * ((ny + 7) / 8 * 8);
  const dim3 grid((unsigned)nwg);
  switch (op.kt) {
    case 1:
      hipLaunchKernelGGL((etd2_uneven<S2, 1>), grid, dim3(STHREADS), 0, s, op, un, st, step, ncols);
      break;
    case 2:
      hipLaunchKernelGGL((etd2_uneven<S2, 2>), grid, dim3(STHREADS), 0, s, op, un, st, step, ncols);
      break;
    default:
      hipLaunchKernelGGL((etd2_uneven<S2, 4>), grid, dim3(STHREADS), 0, s, op, un, st, step, ncols);
      break;
  }
}

// EM sector of the uneven batch: one workgroup per column k that has step
// `step` (win >= 0), after both nuclear stages.  The column's own window w
// selects T_gamma, T_electron, R_BH and the Bethe-Heitler weights; its own
// h, dl/dz and H cm2sec drive the photon ETD2 step.  The photon state
// [d_em][K] is updated in place (the whole column lives in one workgroup).
// The Bethe-Heitler weights sum the species as em_bh_weights_bins does.
__global__ void __launch_bounds__(256) em_uneven_kernel(prh_operator op, prh_uneven un,
                                                       prh_state st, prh_em em, int step) {
  __shared__ double part[4][WAVE];
  __shared__ double wb[WAVE];
  __shared__ double g0[PRH_EM_MAX], ga[PRH_EM_MAX], lep[PRH_EM_MAX];
  const int K = op.K;
  const int d = op.d;
  const int de = em.d_em;
  const int k = blockIdx.x;
  const int w = un.win[(int64_t)step * K + k];
  if (w < 0) return;  // uniform per workgroup: the column has finished
  const int tid = threadIdx.x;
  const int lane = tid & (WAVE - 1);
  const int wv = wave_id();
  // Bethe-Heitler weights of the column: 4 waves over interleaved species
  {
    const bool act = lane < d;
    const int ns = op.nspec;
    const int64_t sS = (int64_t)K * d;
    const double* xp = st.x + (int64_t)k * d + (act ? lane : 0);
    double acc = 0.0;
    int s = wv;
    for (; s + 28 < ns; s += 32) {
      double xv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) xv[u] = act ? xp[(int64_t)(s + 4 * u) * sS] : 0.0;
#pragma unroll
      for (int u = 0; u < 8; ++u) acc += em.cbh[s + 4 * u] * xv[u];
    }
    for (; s < ns; s += 4) acc += em.cbh[s] * (act ? xp[(int64_t)s * sS] : 0.0);
    part[wv][lane] = acc;
  }
  const double hh = un.h[(int64_t)step * K + k];
  const double dl = un.dldz[(int64_t)step * K + k];
  const double had = un.hadia[(int64_t)step * K + k];
  const bool pdiff = (op.flags & PRH_F_PARTIAL_DIFF) != 0;
  for (int i = tid; i < de; i += blockDim.x) g0[i] = em.gam[(int64_t)i * K + k];
  __syncthreads();
  if (wv == 0 && lane < d) {
    const double tot = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
    wb[lane] = em.bhw[(int64_t)w * d + lane] * tot;
  }
  auto stencil = [&](const double* y, int i) {
    double f = 0.0;
    if (pdiff) {
#pragma unroll
      for (int o = -3; o <= 3; ++o) {
        const int j = i + o;
        if (o == 0 || j < 0 || j >= de) continue;
        const double c = em.dcoef[i * 7 + o + 3];
        if (c != 0.0) f += c * ((had * em.egrid[j]) * y[j]);
      }
    }
    return f * dl;
  };
  // photon ETD2 step, stage 1
  double f0 = 0.0, p2 = 0.0;
  const int i = tid;
  if (i < de) {
    double dv = 0.0;
    if (pdiff) {
      double kx = (had * em.egrid[i]) * em.ddiag[i];
      dv = kx * dl;
    }
    const double hd = dv * hh;
    const double e = exp(hd);
    const double p1 = phi1_of(hd, e);
    p2 = phi2_of(hd, e);
    f0 = stencil(g0, i);
    double a = e * g0[i];
    double sc = p1 * f0;
    sc = sc * hh;
    ga[i] = a + sc;
  }
  __syncthreads();
  double xn = 0.0;
  if (i < de) {
    const double fa = stencil(ga, i);
    double sc = fa - f0;
    sc = sc * hh;
    sc = sc * p2;
    xn = ga[i] + sc;
    // Bethe-Heitler pairs: e_src = R_BH w, lep = e- + e+ = 2 e_src
    const double* R = em.rbh + ((int64_t)w * de + i) * d;
    double es = 0.0;
    for (int j = 0; j < d; ++j) es += R[j] * wb[j];
    lep[i] = es + es;
  }
  __syncthreads();
  if (i < de) g0[i] = xn;
  __syncthreads();
  // cascade transfer: gam <- T_gamma gam + T_electron lep
  if (i < de) {
    const double* Tg = em.tg + ((int64_t)w * de + i) * de;
    const double* Te = em.te + ((int64_t)w * de + i) * de;
    double ag = 0.0, ae = 0.0;
    for (int m = 0; m < de; ++m) {
      ag += Tg[m] * g0[m];
      ae += Te[m] * lep[m];
    }
    em.gam[(int64_t)i * K + k] = ag + ae;
  }
}

// ---------------------------------------------------------------------------
// Dense rows (protons, He-4: a block from nearly every mother).  dense_pass
// walks the mothers once per (mother chunk, KT right-hand sides): each mother
// panel is loaded once and applied to every dense row that has a block from it
// (diagonal or band, shuffles along the bins), and the chunk's row sums go to
// dpart[chunk][row][rhs][bin].  The stage kernel adds the chunks in order in
// the dense rows' epilogues; their other blocks stay in the ordinary tasks.
// ---------------------------------------------------------------------------
constexpr int DENSE_MAX = 2;

template <int KT>
__global__ void __launch_bounds__(256) dense_pass(prh_operator op, prh_schedule sch, prh_state st,
                                                  int stage2, int step_local, int use_ctl) {
  int step = step_local;
  if (use_ctl) {
    step = st.step_ctl[0] + step_local;
    if (step >= st.step_ctl[1]) return;
  }
  const int lane = threadIdx.x & (WAVE - 1);
  const int K = op.K;
  const int d = op.d;
  const int ny = (K + KT - 1) / KT;
  const int item = blockIdx.x * 4 + wave_id();
  if (item >= op.ndchunk * ny) return;
  const int c = item / ny;
  const int k0 = (item % ny) * KT;
  const bool active = lane < d;
  const double* __restrict__ xin = stage2 ? st.a : st.x;
  const int w = sch.win[step];
  const double* __restrict__ mv = op.mvals + (int64_t)w * op.mvals_stride;
  const int nr = op.ndense;
  double acc[DENSE_MAX][KT];
#pragma unroll
  for (int r = 0; r < DENSE_MAX; ++r)
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) acc[r][kk] = 0.0;
  const int mo0 = c * op.dchunk;
  const int mo1 = min(mo0 + op.dchunk, op.nspec);
  for (int mo = mo0; mo < mo1; ++mo) {
    // every load of the mother (state panel and all diagonals of every dense
    // row's block) is issued before any use
    double y[KT];
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) {
      const int k = k0 + kk;
      y[kk] = (active && k < K) ? xin[((int64_t)mo * K + k) * d + lane] : 0.0;
    }
    double mr[DENSE_MAX][NSMAX];
    int olo[DENSE_MAX], ns[DENSE_MAX];
#pragma unroll
    for (int r = 0; r < DENSE_MAX; ++r) {
      const int* e = op.dense_tab + 3 * ((int64_t)(r < nr ? r : 0) * op.nspec + mo);
      olo[r] = e[0];
      ns[r] = r < nr ? e[1] : 0;
      const int64_t voff = e[2];
#pragma unroll
      for (int s2 = 0; s2 < NSMAX; ++s2)
        mr[r][s2] = (active && s2 < ns[r]) ? mv[voff + (int64_t)s2 * d + lane] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < DENSE_MAX; ++r) {
#pragma unroll
      for (int s2 = 0; s2 < NSMAX; ++s2) {
        if (s2 >= ns[r]) break;  // uniform
        const int src = lane + olo[r] + s2;
#pragma unroll
        for (int kk = 0; kk < KT; ++kk) acc[r][kk] += mr[r][s2] * shfl_or_zero(y[kk], src, d);
      }
    }
  }
  if (!active) return;
#pragma unroll
  for (int r = 0; r < DENSE_MAX; ++r) {
    if (r >= nr) break;
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) {
      const int k = k0 + kk;
      if (k < K) st.dpart[(((int64_t)c * nr + r) * K + k) * d + lane] = acc[r][kk];
    }
  }
}

template <int KT>
void launch_dense(const prh_operator& op, const prh_schedule& sch, const prh_state& st, int stage2,
                  int step_local, int use_ctl, hipStream_t s) {
  const int ny = (op.K + KT - 1) / KT;
  const int64_t items = (int64_t)op.ndchunk * ny;
  hipLaunchKernelGGL((dense_pass<KT>), dim3((unsigned)((items + 3) / 4)), dim3(256), 0, s, op, sch,
                     st, stage2, step_local, use_ctl);
}

int32_t validate(const prh_operator* op, const prh_schedule* sch, const prh_state* st,
                 const char* who) {
  if (!op || !sch || !st) return set_error(std::string(who) + ": null argument");
  if (op->d <= 0 || op->d > WAVE) return set_error(std::string(who) + ": d must be in [1, 64]");
  if (op->K <= 0) return set_error(std::string(who) + ": K must be >= 1");
  if (op->wg_waves != 0 && op->wg_waves != 8 && op->wg_waves != 16)
    return set_error(std::string(who) + ": wg_waves must be 0 (16), 8 or 16");
  if (op->ntasks % wg_waves_of(*op) != 0)
    return set_error(std::string(who) + ": ntasks must be a multiple of the workgroup waves");
  if (wg_waves_of(*op) != 16 && std::string(who).rfind("prh_adjoint", 0) == 0)
    return set_error(std::string(who) + ": the adjoint kernels run 16-wave workgroups");
  if (op->ntasks <= 0) return set_error(std::string(who) + ": empty operator");
  if (!(op->kt == 1 || op->kt == 2 || op->kt == 4 || op->kt == 8 || op->kt == 16))
    return set_error(std::string(who) + ": kt must be 1, 2, 4, 8 or 16");
  if (op->kt == 16 && (op->mvals_e || std::string(who).rfind("prh_adjoint", 0) == 0 || st->rec_x))
    return set_error(std::string(who) + ": kt = 16 is for one-field forward stages without recording");
  if ((int64_t)(op->ntasks / wg_waves_of(*op)) * ((op->K + op->kt - 1) / op->kt) > 0x7fffffff)
    return set_error(std::string(who) + ": grid too large");
  if (op->mvals_stride * 8 > 0x7fffffff || (int64_t)op->K * op->d * 8 > 0x7fffffff)
    return set_error(std::string(who) + ": a window's values or one species' [K][d] panel exceeds 2 GiB");
  if (op->nfin != 0) {
    if (op->nfin < 0 || op->nfin % wg_waves_of(*op) != 0 || !op->fin_tasks)
      return set_error(std::string(who) + ": nfin must be a multiple of the workgroup waves, with fin_tasks");
    if (op->ndense < 1)
      return set_error(std::string(who) + ": dense-row chunk tasks need ndense >= 1");
  }
  if (op->ndense > 0) {
    if (op->nfin == 0 && op->ndense > DENSE_MAX)
      return set_error(std::string(who) + ": at most 2 dense rows for the mother-major pass");
    if (op->mvals_e || op->kt > 8)
      return set_error(std::string(who) + ": dense rows are for the one-field stages (kt <= 8)");
    if (!op->dense_row || !st->dpart || op->ndchunk < 1)
      return set_error(std::string(who) + ": dense rows need dense_row, dpart and ndchunk >= 1");
    if (op->nfin == 0 && (!op->dense_tab || op->dchunk < 1 ||
                          (int64_t)op->dchunk * op->ndchunk < op->nspec))
      return set_error(std::string(who) + ": the mother-major pass needs dense_tab and chunks covering the species");
    if (std::string(who).rfind("prh_adjoint", 0) == 0)
      return set_error(std::string(who) + ": the adjoint operator has no dense rows");
  }
  if ((op->mvals_e == nullptr) != (op->escale == nullptr) ||
      (op->mvals_e != nullptr && op->rpair_e == nullptr))
    return set_error(std::string(who) + ": mvals_e, rpair_e and escale go together");
  if ((op->flags & PRH_F_ROW_PAIRS) &&
      (op->mvals_e || op->ndense > 0 || op->nfin != 0 || op->kt > 8 ||
       std::string(who).rfind("prh_adjoint", 0) == 0))
    return set_error(std::string(who) +
                     ": row-pair tasks run in the one-field forward stages without dense rows, kt <= 8");
  return validate_em(op, st, who);
}

template <bool S2, bool TWO, int WGW, int DM>
void launch_stage_dm(const prh_operator& op, const prh_schedule& sch, const prh_state& st,
                     int step_local, int use_ctl, hipStream_t s) {
  const int64_t nwg = (int64_t)(op.ntasks / WGW) * ((op.K + op.kt - 1) / op.kt);
  const dim3 grid((unsigned)nwg);
  constexpr int STH = WGW * WAVE;
  switch (op.kt) {
    case 1:
      hipLaunchKernelGGL((etd2_stage<S2, 1, TWO, WGW, DM>), grid, dim3(STH), 0, s, op, sch, st,
                         step_local, use_ctl);
      break;
    case 2:
      hipLaunchKernelGGL((etd2_stage<S2, 2, TWO, WGW, DM>), grid, dim3(STH), 0, s, op, sch, st,
                         step_local, use_ctl);
      break;
    case 4:
      hipLaunchKernelGGL((etd2_stage<S2, 4, TWO, WGW, DM>), grid, dim3(STH), 0, s, op, sch, st,
                         step_local, use_ctl);
      break;
    case 16:
      if constexpr (!TWO && DM == DM_NONE) {
        hipLaunchKernelGGL((etd2_stage<S2, 16, false, WGW, DM_NONE>), grid, dim3(STH), 0, s, op,
                           sch, st, step_local, use_ctl);
      }
      break;
    default:
      hipLaunchKernelGGL((etd2_stage<S2, 8, TWO, WGW, DM>), grid, dim3(STH), 0, s, op, sch, st,
                         step_local, use_ctl);
      break;
  }
}

// dense-row mode of a validated operator (dense rows are one-field, kt <= 8)
template <bool S2, bool TWO, int WGW>
void launch_stage_wg(const prh_operator& op, const prh_schedule& sch, const prh_state& st,
                     int step_local, int use_ctl, hipStream_t s) {
  if constexpr (!TWO) {
    if (op.ndense > 0) {
      if (op.nfin > 0)
        launch_stage_dm<S2, false, WGW, DM_CHUNKS>(op, sch, st, step_local, use_ctl, s);
      else
        launch_stage_dm<S2, false, WGW, DM_PASS>(op, sch, st, step_local, use_ctl, s);
      return;
    }
  }
  launch_stage_dm<S2, TWO, WGW, DM_NONE>(op, sch, st, step_local, use_ctl, s);
}

template <bool S2, int WGW>
void launch_finish_wg(const prh_operator& op, const prh_schedule& sch, const prh_state& st,
                      int step_local, int use_ctl, hipStream_t s) {
  const dim3 grid((unsigned)((int64_t)(op.nfin / WGW) * ((op.K + op.kt - 1) / op.kt)));
  constexpr int STH = WGW * WAVE;
  switch (op.kt) {
    case 1:
      hipLaunchKernelGGL((dense_finish<S2, 1, WGW>), grid, dim3(STH), 0, s, op, sch, st,
                         step_local, use_ctl);
      break;
    case 2:
      hipLaunchKernelGGL((dense_finish<S2, 2, WGW>), grid, dim3(STH), 0, s, op, sch, st,
                         step_local, use_ctl);
      break;
    case 4:
      hipLaunchKernelGGL((dense_finish<S2, 4, WGW>), grid, dim3(STH), 0, s, op, sch, st,
                         step_local, use_ctl);
      break;
    default:
      hipLaunchKernelGGL((dense_finish<S2, 8, WGW>), grid, dim3(STH), 0, s, op, sch, st,
                         step_local, use_ctl);
      break;
  }
}

template <bool S2, bool TWO>
void launch_stage_two(const prh_operator& op, const prh_schedule& sch, const prh_state& st,
                      int step_local, int use_ctl, hipStream_t s) {
  if (wg_waves_of(op) == 8)
    launch_stage_wg<S2, TWO, 8>(op, sch, st, step_local, use_ctl, s);
  else
    launch_stage_wg<S2, TWO, 16>(op, sch, st, step_local, use_ctl, s);
}

template <bool S2>
void launch_stage_kt(const prh_operator& op, const prh_schedule& sch, const prh_state& st,
                     int step_local, int use_ctl, hipStream_t s) {
  if (op.mvals_e) {
    launch_stage_two<S2, true>(op, sch, st, step_local, use_ctl, s);
    return;
  }
  if (op.ndense > 0 && op.nfin == 0) {
    switch (op.kt) {
      case 1: launch_dense<1>(op, sch, st, S2, step_local, use_ctl, s); break;
      case 2: launch_dense<2>(op, sch, st, S2, step_local, use_ctl, s); break;
      case 4: launch_dense<4>(op, sch, st, S2, step_local, use_ctl, s); break;
      default: launch_dense<8>(op, sch, st, S2, step_local, use_ctl, s); break;
    }
  }
  launch_stage_two<S2, false>(op, sch, st, step_local, use_ctl, s);
  if (op.nfin > 0) {
    if (wg_waves_of(op) == 8)
      launch_finish_wg<S2, 8>(op, sch, st, step_local, use_ctl, s);
    else
      launch_finish_wg<S2, 16>(op, sch, st, step_local, use_ctl, s);
  }
}

template <bool S2, bool TWO>
void launch_adjoint_two(const prh_operator& op, const prh_schedule& sch, const prh_adjoint& adj,
                        int step, hipStream_t s) {
  const int64_t nwg = (int64_t)(op.ntasks / SWG) * ((op.K + op.kt - 1) / op.kt);
  const dim3 grid((unsigned)nwg);
  switch (op.kt) {
    case 1:
      hipLaunchKernelGGL((etd2_adjoint<S2, 1, TWO>), grid, dim3(STHREADS), 0, s, op, sch, adj, step);
      break;
    case 2:
      hipLaunchKernelGGL((etd2_adjoint<S2, 2, TWO>), grid, dim3(STHREADS), 0, s, op, sch, adj, step);
      break;
    case 4:
      hipLaunchKernelGGL((etd2_adjoint<S2, 4, TWO>), grid, dim3(STHREADS), 0, s, op, sch, adj, step);
      break;
    default:
      hipLaunchKernelGGL((etd2_adjoint<S2, 8, TWO>), grid, dim3(STHREADS), 0, s, op, sch, adj, step);
      break;
  }
}

// ---------------------------------------------------------------------------
// EM cascade per-window operators built on the device (reference
// cascade/cascade.py:284-407 and 425-589, cascade/bethe_heitler.py:219-351).
//
// For every window w (redshift z_w, field scale s_w) the field-free tensors
// are contracted with the target-photon densities by prh_fold_gemm (MFMA);
// cascade_transfer_kernel then builds, per window in one 16-wave workgroup,
// the pair matrix P, the IC emission shape Sg, the degraded-electron map De,
// the IC cooling map Mic and the kinetic single-scatter iteration
//   esc += pesc G;  e += P ((G - pesc G) dE);
//   G_c = Mic (e_cold dE);  G_h = Sg (e_hot dE);  e = De (e_hot dE);
//   esc += pesc G_c;  G = G_h + G_c - pesc G_c
// until the energy left in G and e is below tol of each column's input; the
// outputs are T_gamma = esc and T_electron = esc (Mic dE).  The three
// products of a step share their right operand and run as one pass of
// v_mfma_f64_16x16x4 tiles over the 144 x 144 padded matrices held in a
// per-window workspace in HBM (L2-resident while the workgroup runs).
// ---------------------------------------------------------------------------
constexpr int CNP = 144;              // padded EM dimension (n <= 144)
constexpr int CTL = CNP / 16;         // 9 tiles per matrix side
constexpr int CWAVES = 8;
constexpr int CTHREADS = CWAVES * WAVE;
constexpr int CTILES = CTL * CTL;     // 81 output tiles
constexpr int CTPW = (CTILES + CWAVES - 1) / CWAVES;  // tiles per wave (11)
constexpr int64_t CMAT = (int64_t)CNP * CNP;
constexpr int CKU = 4;                // k-steps whose loads are issued together (36 = 9 x 4)
constexpr int CKU2 = 2;               // the same for the 3 x 3-tile products (register budget)
enum CascadeMat { CM_P = 0, CM_SG, CM_DE, CM_MIC, CM_G, CM_E, CM_E1, CM_ESC, CM_T, CM_COUNT };

// rows of the contraction output per window (prh_cascade layout)
__device__ __forceinline__ int64_t crow_emis(int) { return 0; }
__device__ __forceinline__ int64_t crow_pair(int n) { return (int64_t)n * n; }
__device__ __forceinline__ int64_t crow_power(int n) { return 2 * (int64_t)n * n; }
__device__ __forceinline__ int64_t crow_rate(int n) { return 2 * (int64_t)n * n + n; }
__device__ __forceinline__ int64_t crow_den(int n) { return 2 * (int64_t)n * n + 2 * n; }
__device__ __forceinline__ int64_t crow_num(int n) { return 2 * (int64_t)n * n + 2 * n + 1; }

// trapezoid of E * M[:, j] over E (column j, rows 0..n-1), sequential
__device__ __forceinline__ double column_energy(const double* M, int j, int n,
                                                const double* E) {
  double acc = 0.0;
  for (int i = 0; i + 1 < n; ++i) {
    const double dx = E[i + 1] - E[i];
    acc += dx * (E[i + 1] * M[(int64_t)(i + 1) * CNP + j] + E[i] * M[(int64_t)i * CNP + j]) / 2.0;
  }
  return acc;
}

__global__ void __launch_bounds__(CTHREADS) cascade_transfer_kernel(
    prh_cascade cs, const double* __restrict__ contr, int64_t cstride,
    const double* __restrict__ dtau, const double* __restrict__ dl, double* __restrict__ work,
    double* __restrict__ tg, double* __restrict__ te, int32_t* __restrict__ iters) {
  __shared__ double sE[CNP], sdE[CNP], sR[CNP], sloss[CNP], spesc[CNP], scold[CNP], sEin[CNP];
  __shared__ double sred[CNP];
  __shared__ int sdone;
  const int n = cs.n;
  const int w = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & (WAVE - 1);
  const int wv = wave_id();
  const double* cw = contr + (int64_t)w * cstride;
  double* M = work + (int64_t)w * CM_COUNT * CMAT;
  double* P = M + CM_P * CMAT;
  double* Sg = M + CM_SG * CMAT;
  double* De = M + CM_DE * CMAT;
  double* Mic = M + CM_MIC * CMAT;
  double* G = M + CM_G * CMAT;
  double* Ev = M + CM_E * CMAT;
  double* E1 = M + CM_E1 * CMAT;
  double* Esc = M + CM_ESC * CMAT;
  double* Tc = M + CM_T * CMAT;  // G_c of the iteration, until it is added into G

  for (int i = tid; i < CNP; i += CTHREADS) {
    sE[i] = i < n ? cs.E[i] : 0.0;
    sdE[i] = i < n ? cs.dE[i] : 0.0;
  }
  // emission into E1, pair spectrum into P, every other matrix zeroed: the
  // padded rows and columns (n <= i, j < 144) take part in the tile products
  // and must hold zeros, never stale workspace contents
  for (int64_t t = tid; t < CMAT; t += CTHREADS) {
    const int i = (int)(t / CNP), j = (int)(t % CNP);
    const bool in = i < n && j < n;
    E1[t] = in ? cs.pref_emis[j] * cw[crow_emis(n) + (int64_t)i * n + j] : 0.0;
    P[t] = in ? cs.pref_pair[j] * cw[crow_pair(n) + (int64_t)i * n + j] : 0.0;
    Sg[t] = 0.0;
    De[t] = 0.0;
    Mic[t] = 0.0;
    G[t] = (in && i == j) ? 1.0 : 0.0;
    Ev[t] = 0.0;
    Esc[t] = 0.0;
    Tc[t] = 0.0;
  }
  __syncthreads();
  // per column j: pair-matrix energy normalisation, IC emission rate R_j
  for (int j = tid; j < n; j += CTHREADS) {
    const double oe = column_energy(P, j, n, sE);
    if (oe > 0.0)
      for (int i = 0; i < n; ++i) P[(int64_t)i * CNP + j] = (P[(int64_t)i * CNP + j] / oe) * sE[j];
    double r = 0.0;
    for (int i = 0; i < n; ++i) r += E1[(int64_t)i * CNP + j] * sdE[i];
    sR[j] = r;
  }
  if (tid == 0) {
    const double den = cw[crow_den(n)];
    const double eps_mean = den > 0.0 ? cw[crow_num(n)] / den : 0.0;
    sred[0] = eps_mean;
  }
  __syncthreads();
  const double eps_mean = sred[0];
  for (int j = tid; j < n; j += CTHREADS) {
    const double pw = cs.pref_emis[j] * cw[crow_power(n) + j];
    const double rt = cs.pref_emis[j] * cw[crow_rate(n) + j];
    sloss[j] = pw - rt * eps_mean;
  }
  __syncthreads();
  // Sg, cold flag and the degraded-electron map (column j by one thread, rows
  // in order as the host's scatter-add)
  for (int j = tid; j < n; j += CTHREADS) {
    const double r = sR[j];
    double eg = 0.0;
    for (int i = 0; i < n; ++i) {
      const double s = r > 0.0 ? E1[(int64_t)i * CNP + j] / r : 0.0;
      Sg[(int64_t)i * CNP + j] = s;
      eg += (sE[i] * s) * sdE[i];
    }
    scold[j] = ((eg < cs.cool_floor * sE[j]) || (r <= 0.0)) ? 1.0 : 0.0;
    if (r > 0.0) {
      for (int k = 0; k < n; ++k) {
        const int b = cs.cic_b[(int64_t)k * n + j];
        const double wk = Sg[(int64_t)k * CNP + j] * sdE[k];
        if (b < 0 || !(wk > 0.0)) continue;
        const double f = cs.cic_f[(int64_t)k * n + j];
        De[(int64_t)b * CNP + j] += wk * f / sdE[b];
        De[(int64_t)(b + 1) * CNP + j] += wk * (1.0 - f) / sdE[b + 1];
      }
    }
  }
  // IC cooling map: row i cumulates emis/loss along j, relative to column i
  for (int i = tid; i < n; i += CTHREADS) {
    double cum = 0.0, prev = 0.0, at_i = 0.0;
    for (int j = 0; j < n; ++j) {
      const double l = sloss[j];
      const double g = l > 0.0 ? E1[(int64_t)i * CNP + j] / l : 0.0;
      if (j > 0) cum = cum + (sE[j] - sE[j - 1]) * (g + prev) / 2.0;
      prev = g;
      if (j == i) at_i = cum;
      Mic[(int64_t)i * CNP + j] = j >= i ? cum - at_i : 0.0;
    }
    spesc[i] = exp(-dtau[(int64_t)w * n + i] * dl[w]);
    sEin[i] = sE[i] * sdE[i];
  }
  for (int i = n + tid; i < CNP; i += CTHREADS) {
    spesc[i] = 0.0;
    sEin[i] = 0.0;
    scold[i] = 0.0;
  }
  __syncthreads();
  for (int j = tid; j < n; j += CTHREADS) {
    const double oe = column_energy(Mic, j, n, sE);
    if (oe > 0.0)
      for (int i = 0; i < n; ++i) Mic[(int64_t)i * CNP + j] = (Mic[(int64_t)i * CNP + j] / oe) * sE[j];
  }
  __syncthreads();
  // the four fixed left operands of the iteration stored transposed (k-major):
  // an operand fragment (16 rows x 4 k) is then four 128-byte row segments
  double* fixed[4] = {P, Sg, De, Mic};
  for (int m = 0; m < 4; ++m) {
    double* X = fixed[m];
    for (int64_t t = tid; t < CMAT; t += CTHREADS) Tc[t] = X[t];
    __syncthreads();
    for (int64_t t = tid; t < CMAT; t += CTHREADS) {
      const int i = (int)(t / CNP), j = (int)(t % CNP);
      X[t] = Tc[(int64_t)j * CNP + i];
    }
    __syncthreads();
  }

  const int col = lane & 15;
  const int kq = lane >> 4;
  int it = 0;
  for (it = 1; it <= cs.max_scatter; ++it) {
    // e1 = e + P ((G - pesc G) dE);  esc += pesc G
    // unit = 3 row tiles x 1 column tile (27 units): per k-step 3 + 1 operand
    // loads for 3 products, the loads of CKU k-steps issued together
#pragma unroll 1
    for (int u = wv; u < 3 * CTL; u += CWAVES) {
      const int r0 = (u / CTL) * 48, c0 = (u % CTL) * 16;
      double4_t acc[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) acc[r] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
      for (int k0 = 0; k0 < CNP; k0 += 4 * CKU) {
        double a[CKU][3], g[CKU];
#pragma unroll
        for (int v = 0; v < CKU; ++v) {
          const int k = k0 + 4 * v + kq;
#pragma unroll
          for (int r = 0; r < 3; ++r) a[v][r] = P[(int64_t)k * CNP + r0 + 16 * r + col];
          g[v] = G[(int64_t)k * CNP + c0 + col];
        }
#pragma unroll
        for (int v = 0; v < CKU; ++v) {
          const int k = k0 + 4 * v + kq;
          const double b = (g[v] - spesc[k] * g[v]) * sdE[k];
#pragma unroll
          for (int r = 0; r < 3; ++r)
            acc[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[v][r], b, acc[r], 0, 0, 0);
        }
      }
#pragma unroll
      for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = r0 + 16 * r + kq + 4 * v, j = c0 + col;
          const int64_t o = (int64_t)i * CNP + j;
          E1[o] = Ev[o] + acc[r][v];
          Esc[o] += spesc[i] * G[o];
        }
      }
    }
    __syncthreads();
    // G_c = Mic (e1 cold dE) -> Tc, G_h = Sg (e1 hot dE) -> G, e = De (e1 hot dE)
    // -> Ev; unit = one product on 3 x 3 tiles (27 units): per k-step 3 + 3
    // operand loads for 9 products
#pragma unroll 1
    for (int u = wv; u < 27; u += CWAVES) {
      const int pr = u / 9, st = u % 9;
      const int r0 = (st / 3) * 48, c0 = (st % 3) * 48;
      const double* A = pr == 0 ? Mic : (pr == 1 ? Sg : De);
      const bool cold = pr == 0;
      double4_t acc[3][3];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[r][c] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
      for (int k0 = 0; k0 < CNP; k0 += 4 * CKU2) {
        double a[CKU2][3], e[CKU2][3];
#pragma unroll
        for (int v = 0; v < CKU2; ++v) {
          const int k = k0 + 4 * v + kq;
#pragma unroll
          for (int r = 0; r < 3; ++r) a[v][r] = A[(int64_t)k * CNP + r0 + 16 * r + col];
#pragma unroll
          for (int c = 0; c < 3; ++c) e[v][c] = E1[(int64_t)k * CNP + c0 + 16 * c + col];
        }
#pragma unroll
        for (int v = 0; v < CKU2; ++v) {
          const int k = k0 + 4 * v + kq;
          const bool take = (scold[k] != 0.0) == cold;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const double y = take ? e[v][c] * sdE[k] : 0.0;
#pragma unroll
            for (int r = 0; r < 3; ++r)
              acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[v][r], y, acc[r][c], 0, 0, 0);
          }
        }
      }
      double* out = pr == 0 ? Tc : (pr == 1 ? G : Ev);
#pragma unroll
      for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int i = r0 + 16 * r + kq + 4 * v, j = c0 + 16 * c + col;
            const int64_t o = (int64_t)i * CNP + j;
            out[o] = acc[r][c][v];
            if (cold) Esc[o] += spesc[i] * acc[r][c][v];
          }
        }
      }
    }
    __syncthreads();
    // G = (G_h + G_c) - pesc G_c
    for (int64_t t = tid; t < CMAT; t += CTHREADS) {
      const int i = (int)(t / CNP);
      G[t] = (G[t] + Tc[t]) - spesc[i] * Tc[t];
    }
    __syncthreads();
    // energy still in G and e per input column, relative to the column's input
    double worst = -1.0;
    for (int j = tid; j < n; j += CTHREADS) {
      double eg = 0.0, ee = 0.0;
      for (int i = 0; i < n; ++i) {
        eg += sEin[i] * G[(int64_t)i * CNP + j];
        ee += sEin[i] * Ev[(int64_t)i * CNP + j];
      }
      worst = fmax(worst, (eg + ee) / sEin[j]);
    }
    if (tid < CNP) sred[tid] = worst;
    __syncthreads();
    if (tid == 0) {
      double m = sred[0];
      for (int j = 1; j < CNP; ++j) m = fmax(m, sred[j]);
      sdone = m < cs.tol ? 1 : 0;
    }
    __syncthreads();
    if (sdone) {
      for (int64_t t = tid; t < CMAT; t += CTHREADS) {
        const int i = (int)(t / CNP);
        Esc[t] += spesc[i] * G[t];
      }
      __syncthreads();
      break;
    }
  }
  if (tid == 0) iters[w] = it > cs.max_scatter ? cs.max_scatter : it;
  // outputs: T_gamma = esc, T_electron = esc (Mic dE)
  double* tgw = tg + (int64_t)w * n * n;
  double* tew = te + (int64_t)w * n * n;
  for (int64_t t = tid; t < (int64_t)n * n; t += CTHREADS) {
    const int i = (int)(t / n), j = (int)(t % n);
    tgw[t] = Esc[(int64_t)i * CNP + j];
  }
#pragma unroll 1
  for (int q = 0; q < CTPW; ++q) {
    const int t = wv + CWAVES * q;
    if (t >= CTILES) break;
    const int r0 = (t / CTL) * 16, c0 = (t % CTL) * 16;
    double4_t acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < CNP; k0 += 4) {
      const int k = k0 + kq;
      const int j = c0 + col;
      const double a = Esc[(int64_t)(r0 + col) * CNP + k];
      const double b = Mic[(int64_t)j * CNP + k] * sdE[j];  // Mic is stored transposed
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = r0 + kq + 4 * v, j = c0 + col;
      if (i < n && j < n) tew[(int64_t)i * n + j] = acc[v];
    }
  }
}

// Bethe-Heitler pair shapes R[w][i][j] (lepton bin i of the EM grid per
// nucleon bin j): K_bh contracted over ln eps with the window's target
// densities by the trapezoid over the points where n > 0, restricted to the
// band E_i <= Ep_j, E_i > 1e-6 Ep_j, Ep_j >= bh_emin, and normalised per
// column so that 2 int E R dE = Ep_j.  One workgroup per window; the column
// pass is one thread per nucleon bin.
__global__ void __launch_bounds__(256) cascade_bh_kernel(prh_cascade cs,
                                                         const double* __restrict__ nb_dens,
                                                         double* __restrict__ rbh) {
  __shared__ double wts[PRH_BH_MAX];
  const int n = cs.n, d = cs.d, nb = cs.nb;
  const int w = blockIdx.x;
  const int tid = threadIdx.x;
  const double* nw = nb_dens + (int64_t)w * nb;
  if (tid == 0) {
    // trapezoid weights on ln eps over the used points, times n eps
    for (int u = 0; u < nb; ++u) wts[u] = 0.0;
    int prev = -1;
    for (int u = 0; u < nb; ++u) {
      if (!(nw[u] > 0.0)) continue;
      if (prev >= 0) {
        const double h = 0.5 * (log(cs.eps_b[u]) - log(cs.eps_b[prev]));
        wts[prev] += h;
        wts[u] += h;
      }
      prev = u;
    }
    for (int u = 0; u < nb; ++u) wts[u] *= nw[u] > 0.0 ? nw[u] * cs.eps_b[u] : 0.0;
  }
  __syncthreads();
  double* R = rbh + (int64_t)w * n * d;
  for (int64_t t = tid; t < (int64_t)n * d; t += blockDim.x) {
    const int i = (int)(t / d), j = (int)(t % d);
    const double e = cs.E[i], ep = cs.Ecr[j];
    double v = 0.0;
    if (e <= ep && e > 1e-6 * ep && ep >= cs.bh_emin) {
      const double* kr = cs.kern_b + t * nb;
      for (int u = 0; u < nb; ++u) v += kr[u] * wts[u];
    }
    R[t] = v;
  }
  __syncthreads();
  for (int j = tid; j < d; j += blockDim.x) {
    double carried = 0.0;
    for (int i = 0; i + 1 < n; ++i)
      carried += (cs.E[i + 1] - cs.E[i]) *
                 (cs.E[i + 1] * R[(int64_t)(i + 1) * d + j] + cs.E[i] * R[(int64_t)i * d + j]) / 2.0;
    carried *= 2.0;
    const double f = carried > 0.0 ? cs.Ecr[j] / carried : 0.0;
    for (int i = 0; i < n; ++i) R[(int64_t)i * d + j] *= f;
  }
}

}  // namespace

// Static device scratch of prh_etd2_steps_noalloc: [256 B step counter]
// [a: nspec K d][F(x): nspec K d][dense-row chunk partials].  Every GPU holds
// its own copy at its own address; host code uses only the looked-up address.
static __device__ __align__(256) double scratch[PRH_SCRATCH_BYTES / 8];

namespace {

constexpr int PRH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[PRH_MAX_DEVICES] = {};

// Device address of `scratch` on the current device (looked up once per device).
int32_t scratch_on_current_device(double** out) {
  int dev = -1;
  if (int32_t e = check_hip(hipGetDevice(&dev), "prh_etd2_steps_noalloc: hipGetDevice")) return e;
  if (dev < 0 || dev >= PRH_MAX_DEVICES)
    return set_error("prh_etd2_steps_noalloc: device index outside [0, 64)");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    if (int32_t e = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                              "prh_etd2_steps_noalloc: hipGetSymbolAddress(scratch)"))
      return e;
    if (!p) return set_error("prh_etd2_steps_noalloc: null scratch address");
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[dev];
  return 0;
}

}  // namespace

struct prh_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int32_t baked_first = -1;  // >= 0: steps baked in at capture (uneven chunks)
  int32_t chunk = 0;
  int32_t nsteps = 0;
  int32_t* step_ctl = nullptr;
};

extern "C" {

int32_t prh_version(void) { return 5; }

const char* prh_last_error(void) { return g_last_error.c_str(); }

int32_t prh_fold_gemm(const double* kern, int64_t nrows, int32_t dph, const double* phot,
                      int32_t nwin, double* out, int64_t out_stride, void* stream) {
  if (nrows < 0 || nwin < 0) return set_error("prh_fold_gemm: negative size");
  if (dph <= 0 || dph > 4096) return set_error("prh_fold_gemm: dph must be in [1, 4096]");
  if (out_stride < nrows) return set_error("prh_fold_gemm: out_stride < nrows");
  if (nrows == 0 || nwin == 0) return 0;
  if (!kern || !phot || !out) return set_error("prh_fold_gemm: null pointer");
  const int64_t waves = (nrows + 15) / 16;
  const int64_t blocks = (waves + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
  if (blocks > 0x7fffffff) return set_error("prh_fold_gemm: too many rows");
  if (dph <= 4 * FOLD_MAX_KSTEPS)
    hipLaunchKernelGGL(fold_gemm_mfma, dim3((unsigned)((waves + FOLD_WAVES - 1) / FOLD_WAVES)),
                       dim3(FOLD_THREADS), 0, (hipStream_t)stream, kern, nrows, (int)dph, phot,
                       (int)nwin, out, out_stride);
  else
    hipLaunchKernelGGL(contract_rows_mfma, dim3((unsigned)blocks), dim3(THREADS), 0,
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
  if (int32_t e = validate(op, sch, st, "prh_etd2_steps")) return e;
  if (k_begin < 0 || k_end > sch->nsteps || k_begin > k_end)
    return set_error("prh_etd2_steps: step range outside the schedule");
  if (k_begin == k_end) return 0;
  hipStream_t s = (hipStream_t)stream;
  for (int k = k_begin; k < k_end; ++k) {
    launch_stage_kt<false>(*op, *sch, *st, k, 0, s);
    launch_stage_kt<true>(*op, *sch, *st, k, 0, s);
    if (st->em) launch_em(*op, *sch, *st, k, 0, s);
  }
  return check_hip(hipGetLastError(), "prh_etd2_steps launch");
}

int32_t prh_etd2_steps_noalloc(const prh_operator* op, const prh_schedule* sch, double* x,
                               int32_t k_begin, int32_t k_end, void* stream) {
  if (!op || !sch || !x) return set_error("prh_etd2_steps_noalloc: null argument");
  if (op->nspec <= 0 || op->K <= 0 || op->d <= 0)
    return set_error("prh_etd2_steps_noalloc: empty operator");
  const int64_t panel = (int64_t)op->nspec * op->K * op->d;
  const int64_t dpart = op->ndense > 0 ? (int64_t)op->ndchunk * op->ndense * op->K * op->d : 0;
  if (dpart < 0) return set_error("prh_etd2_steps_noalloc: negative dense-row sizes");
  const int64_t need = 256 + 8 * (2 * panel + dpart);
  if (need > PRH_SCRATCH_BYTES)
    return set_error("prh_etd2_steps_noalloc: needs " + std::to_string(need) +
                     " bytes of scratch, the static array holds " +
                     std::to_string((long long)PRH_SCRATCH_BYTES));
  double* base = nullptr;
  if (int32_t e = scratch_on_current_device(&base)) return e;
  prh_state st{};
  st.x = x;
  st.step_ctl = reinterpret_cast<int32_t*>(base);
  st.a = base + 32;
  st.fphi = st.a + panel;
  st.dpart = dpart > 0 ? st.fphi + panel : nullptr;
  // every stage-1 task writes a and F(x) of its row before stage 2 reads
  // them, chunk tasks write their partials before the finishing launch, and
  // direct launches do not read the step counter: nothing needs clearing
  return prh_etd2_steps(op, sch, &st, k_begin, k_end, stream);
}

int32_t prh_etd2_stage(const prh_operator* op, const prh_schedule* sch, const prh_state* st,
                       int32_t step, int32_t stage, void* stream) {
  if (int32_t e = validate(op, sch, st, "prh_etd2_stage")) return e;
  if (step < 0 || step >= sch->nsteps) return set_error("prh_etd2_stage: step outside the schedule");
  if (stage != 1 && stage != 2) return set_error("prh_etd2_stage: stage must be 1 or 2");
  hipStream_t s = (hipStream_t)stream;
  if (stage == 1) {
    launch_stage_kt<false>(*op, *sch, *st, step, 0, s);
  } else {
    launch_stage_kt<true>(*op, *sch, *st, step, 0, s);
    if (st->em) launch_em(*op, *sch, *st, step, 0, s);
  }
  return check_hip(hipGetLastError(), "prh_etd2_stage launch");
}

int32_t prh_em_step(const prh_operator* op, const prh_schedule* sch, const prh_state* st,
                    int32_t step, void* stream) {
  if (int32_t e = validate(op, sch, st, "prh_em_step")) return e;
  if (!st->em) return set_error("prh_em_step: the state has no EM sector");
  if (step < 0 || step >= sch->nsteps) return set_error("prh_em_step: step outside the schedule");
  launch_em(*op, *sch, *st, step, 0, (hipStream_t)stream);
  return check_hip(hipGetLastError(), "prh_em_step launch");
}

int32_t prh_adjoint_seed(const prh_operator* opT, const prh_schedule* sch,
                         const prh_adjoint* adj, int32_t step, void* stream) {
  if (!adj) return set_error("prh_adjoint_seed: null adjoint");
  prh_state dummy{};
  if (int32_t e = validate(opT, sch, &dummy, "prh_adjoint_seed")) return e;
  if (step < 0 || step >= sch->nsteps) return set_error("prh_adjoint_seed: step outside the schedule");
  if (!adj->lam || !adj->u) return set_error("prh_adjoint_seed: lam and u are required");
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid((unsigned)(opT->ntasks / SWG));
  if (opT->mvals_e)
    hipLaunchKernelGGL((adjoint_seed<true>), grid, dim3(STHREADS), 0, s, *opT, *sch, *adj, step);
  else
    hipLaunchKernelGGL((adjoint_seed<false>), grid, dim3(STHREADS), 0, s, *opT, *sch, *adj, step);
  return check_hip(hipGetLastError(), "prh_adjoint_seed launch");
}

int32_t prh_adjoint_step(const prh_operator* opT, const prh_schedule* sch, const prh_adjoint* adj,
                         int32_t step, void* stream) {
  if (!adj) return set_error("prh_adjoint_step: null adjoint");
  prh_state dummy{};
  if (int32_t e = validate(opT, sch, &dummy, "prh_adjoint_step")) return e;
  if (step < 0 || step >= sch->nsteps) return set_error("prh_adjoint_step: step outside the schedule");
  if (!adj->lam || !adj->lam_a || !adj->u || !adj->v)
    return set_error("prh_adjoint_step: lam, lam_a, u and v are required");
  const bool want_s = opT->mvals_e && adj->gdot;
  if (want_s) {
    if (!adj->rec_x || !adj->rec_a || !adj->rec_fx || !adj->rec_fa)
      return set_error("prh_adjoint_step: the EBL-scale gradient needs the recorded forward step");
    if (step - adj->rec_first < 0 || step - adj->rec_first >= adj->rec_count)
      return set_error("prh_adjoint_step: step outside the recorded chunk");
  }
  if (adj->gsrc0 && (!adj->gsrc1 || !adj->lnz1))
    return set_error("prh_adjoint_step: source accumulators need gsrc1 and lnz1");
  hipStream_t s = (hipStream_t)stream;
  if (opT->mvals_e) {
    launch_adjoint_two<false, true>(*opT, *sch, *adj, step, s);
    launch_adjoint_two<true, true>(*opT, *sch, *adj, step, s);
  } else {
    launch_adjoint_two<false, false>(*opT, *sch, *adj, step, s);
    launch_adjoint_two<true, false>(*opT, *sch, *adj, step, s);
  }
  return check_hip(hipGetLastError(), "prh_adjoint_step launch");
}

int32_t prh_adjoint_steps(const prh_operator* opT, const prh_schedule* sch,
                          const prh_adjoint* adj, int32_t k_begin, int32_t k_end, void* stream) {
  if (k_begin < 0 || k_end > (sch ? sch->nsteps : 0) || k_begin > k_end)
    return set_error("prh_adjoint_steps: step range outside the schedule");
  for (int32_t k = k_end - 1; k >= k_begin; --k) {
    if (int32_t e = prh_adjoint_step(opT, sch, adj, k, stream)) return e;
  }
  return 0;
}

int32_t prh_graph_create(const prh_operator* op, const prh_schedule* sch, const prh_state* st,
                         int32_t chunk, prh_graph** out) {
  if (int32_t e = validate(op, sch, st, "prh_graph_create")) return e;
  if (!out) return set_error("prh_graph_create: null output");
  if (!st->step_ctl) return set_error("prh_graph_create: state has no step_ctl buffer");
  if (chunk < 1 || chunk > 4096) return set_error("prh_graph_create: chunk must be in [1, 4096]");
  *out = nullptr;
  hipStream_t cap = nullptr;
  if (int32_t e = check_hip(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking),
                            "prh_graph_create stream"))
    return e;
  prh_graph* g = new prh_graph();
  g->chunk = chunk;
  g->nsteps = sch->nsteps;
  g->step_ctl = st->step_ctl;
  hipError_t err = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (err == hipSuccess) {
    for (int j = 0; j < chunk; ++j) {
      launch_stage_kt<false>(*op, *sch, *st, j, 1, cap);
      launch_stage_kt<true>(*op, *sch, *st, j, 1, cap);
      if (st->em) launch_em(*op, *sch, *st, j, 1, cap);
    }
    hipLaunchKernelGGL(step_ctl_advance, dim3(1), dim3(1), 0, cap, st->step_ctl, chunk);
    err = hipStreamEndCapture(cap, &g->graph);
  }
  if (err == hipSuccess) err = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  hipStreamDestroy(cap);
  if (err != hipSuccess) {
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    delete g;
    return set_error(std::string("prh_graph_create: ") + hipGetErrorString(err));
  }
  *out = g;
  return 0;
}

int32_t prh_graph_run(prh_graph* g, int32_t k_begin, int32_t k_end, void* stream) {
  if (!g || !g->exec) return set_error("prh_graph_run: invalid graph");
  if (g->baked_first >= 0) {
    if (k_begin != g->baked_first || k_end != g->baked_first + g->chunk)
      return set_error("prh_graph_run: an uneven chunk graph runs exactly its captured steps");
    hipError_t err = hipGraphLaunch(g->exec, (hipStream_t)stream);
    if (err != hipSuccess) return set_error(std::string("prh_graph_run: ") + hipGetErrorString(err));
    return check_hip(hipGetLastError(), "prh_graph_run");
  }
  if (k_begin < 0 || k_end < k_begin || k_end > g->nsteps)
    return set_error("prh_graph_run: step range outside the schedule");
  if (k_begin == k_end) return 0;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(step_ctl_set, dim3(1), dim3(1), 0, s, g->step_ctl, k_begin, k_end);
  const int32_t nchunks = (k_end - k_begin + g->chunk - 1) / g->chunk;
  for (int32_t c = 0; c < nchunks; ++c) {
    hipError_t err = hipGraphLaunch(g->exec, s);
    if (err != hipSuccess) return set_error(std::string("prh_graph_run: ") + hipGetErrorString(err));
  }
  return check_hip(hipGetLastError(), "prh_graph_run");
}

int32_t prh_graph_destroy(prh_graph* g) {
  if (!g) return 0;
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  delete g;
  return 0;
}


int32_t prh_em_transfer_stack(const double* tg, const double* te, const double* rbh,
                              int32_t nwin, int32_t d_em, int32_t d, double* out, void* stream) {
  if (!tg || !te || !rbh || !out) return set_error("prh_em_transfer_stack: null argument");
  if (d_em < 1 || d_em > PRH_EM_MAX || d < 1 || d > WAVE || nwin < 0)
    return set_error("prh_em_transfer_stack: bad sizes");
  const int64_t total = (int64_t)nwin * d_em * (d_em + d);
  if (total == 0) return 0;
  hipLaunchKernelGGL(em_stack_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, tg, te, rbh, (int)nwin, (int)d_em, (int)d, out);
  return check_hip(hipGetLastError(), "prh_em_transfer_stack launch");
}

int32_t prh_em_step_transfer(const prh_operator* op, const prh_schedule* sch, const prh_em* em,
                             double* qstep, void* stream) {
  if (!op || !sch || !em || !qstep) return set_error("prh_em_step_transfer: null argument");
  if (em->d_em < 1 || em->d_em > PRH_EM_MAX)
    return set_error("prh_em_step_transfer: d_em must be in [1, 144]");
  if (!em->egrid || !em->dcoef || !em->ddiag || !em->tg)
    return set_error("prh_em_step_transfer: EM sector with a null buffer");
  if (!sch->h || !sch->dldz || !sch->hadia || !sch->win || sch->nsteps < 0)
    return set_error("prh_em_step_transfer: null schedule buffer");
  if (sch->nsteps == 0) return 0;
  hipLaunchKernelGGL(em_qstep_kernel, dim3((unsigned)sch->nsteps), dim3(256), 0,
                     (hipStream_t)stream, *op, *sch, *em, qstep);
  return check_hip(hipGetLastError(), "prh_em_step_transfer launch");
}

int32_t prh_toeplitz_rows(const prh_toeplitz* tz, const int64_t* sel, int64_t nsel, double* out,
                          void* stream) {
  if (!tz || !sel || !out) return set_error("prh_toeplitz_rows: null argument");
  if (tz->d < 1 || tz->dph < 1 || tz->n1 < 1 || tz->n2 < 1 || tz->nrows < 0)
    return set_error("prh_toeplitz_rows: empty factor tables");
  if (!tz->desc || !tz->tab1d || !tz->tab2d || !tz->fmo || !tz->fdiff)
    return set_error("prh_toeplitz_rows: null factor table");
  if (nsel < 0) return set_error("prh_toeplitz_rows: negative row count");
  if (nsel == 0) return 0;
  const int64_t total = nsel * tz->dph;
  const int64_t blocks = (total + 255) / 256;
  if (blocks > 0x7fffffff) return set_error("prh_toeplitz_rows: too many rows");
  hipLaunchKernelGGL(toeplitz_rows_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     (hipStream_t)stream, *tz, sel, nsel, out);
  return check_hip(hipGetLastError(), "prh_toeplitz_rows launch");
}

int32_t prh_cascade_transfer(const prh_cascade* cs, const double* contr, int64_t cstride,
                             const double* dtau, const double* dl, int32_t nwin, double* work,
                             int64_t work_bytes, double* tg, double* te, int32_t* iters,
                             void* stream) {
  if (!cs || !contr || !dtau || !dl || !work || !tg || !te || !iters)
    return set_error("prh_cascade_transfer: null argument");
  if (cs->n < 2 || cs->n > CNP) return set_error("prh_cascade_transfer: n must be in [2, 144]");
  if (!cs->E || !cs->dE || !cs->pref_emis || !cs->pref_pair || !cs->cic_b || !cs->cic_f)
    return set_error("prh_cascade_transfer: null field-free table");
  if (cs->max_scatter < 1) return set_error("prh_cascade_transfer: max_scatter must be >= 1");
  if (nwin < 0) return set_error("prh_cascade_transfer: negative window count");
  if (nwin == 0) return 0;
  const int64_t rows = 2 * (int64_t)cs->n * cs->n + 2 * cs->n + 2;
  if (cstride < rows) return set_error("prh_cascade_transfer: cstride below the contraction rows");
  if (work_bytes < (int64_t)nwin * CM_COUNT * CMAT * 8)
    return set_error("prh_cascade_transfer: workspace below 9 * 144 * 144 doubles per window");
  hipLaunchKernelGGL(cascade_transfer_kernel, dim3((unsigned)nwin), dim3(CTHREADS), 0,
                     (hipStream_t)stream, *cs, contr, cstride, dtau, dl, work, tg, te, iters);
  return check_hip(hipGetLastError(), "prh_cascade_transfer launch");
}

int32_t prh_cascade_bh(const prh_cascade* cs, const double* nb_dens, int32_t nwin, double* rbh,
                       void* stream) {
  if (!cs || !nb_dens || !rbh) return set_error("prh_cascade_bh: null argument");
  if (cs->n < 2 || cs->n > CNP || cs->d < 1 || cs->d > WAVE)
    return set_error("prh_cascade_bh: n must be in [2, 144] and d in [1, 64]");
  if (cs->nb < 2 || cs->nb > PRH_BH_MAX) return set_error("prh_cascade_bh: nb must be in [2, 64]");
  if (!cs->E || !cs->Ecr || !cs->eps_b || !cs->kern_b)
    return set_error("prh_cascade_bh: null field-free table");
  if (nwin < 0) return set_error("prh_cascade_bh: negative window count");
  if (nwin == 0) return 0;
  hipLaunchKernelGGL(cascade_bh_kernel, dim3((unsigned)nwin), dim3(256), 0, (hipStream_t)stream,
                     *cs, nb_dens, rbh);
  return check_hip(hipGetLastError(), "prh_cascade_bh launch");
}

int32_t prh_field_tables(const double* zs, const double* ss, int32_t nw, const double* P,
                         const double* W, const double* E0, int32_t ni, int32_t nm,
                         double planck_pref, double e_cmb, double zpow, double* out,
                         void* stream) {
  if (nw < 0 || ni < 0 || nm < 0) return set_error("prh_field_tables: negative size");
  if ((int64_t)nw * ni == 0) return 0;
  if (!zs || !ss || !P || !W || !E0 || !out) return set_error("prh_field_tables: null pointer");
  const int64_t n = (int64_t)nw * ni;
  hipLaunchKernelGGL(field_tables_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, zs, ss, (int)nw, P, W, E0, (int)ni, (int)nm,
                     planck_pref, e_cmb, zpow, out);
  return check_hip(hipGetLastError(), "prh_field_tables launch");
}

int32_t prh_fold_rows(const double* kern, int64_t nrows, int32_t dph, const double* phot,
                      const int32_t* in_rows, const int32_t* out_rows, int32_t ncol, double* out,
                      int64_t out_stride, void* stream) {
  if (nrows < 0 || ncol < 0) return set_error("prh_fold_rows: negative size");
  if (dph <= 0 || dph > 4 * FOLD_MAX_KSTEPS) return set_error("prh_fold_rows: dph must be in [1, 128]");
  if (out_stride < nrows) return set_error("prh_fold_rows: out_stride < nrows");
  if (nrows == 0 || ncol == 0) return 0;
  if (!kern || !phot || !in_rows || !out_rows || !out) return set_error("prh_fold_rows: null pointer");
  const int64_t waves = (nrows + 15) / 16;
  const int64_t blocks = (waves + FOLD_WAVES - 1) / FOLD_WAVES;
  hipLaunchKernelGGL(fold_rows_mfma, dim3((unsigned)blocks), dim3(FOLD_THREADS), 0,
                     (hipStream_t)stream, kern, nrows, (int)dph, phot, in_rows, out_rows, (int)ncol,
                     out, out_stride);
  return check_hip(hipGetLastError(), "prh_fold_rows launch");
}

int32_t prh_uneven_graph_create(const prh_operator* op, const prh_uneven* un, const prh_state* st,
                                const prh_refold* rf, int32_t k_begin, int32_t k_end,
                                prh_graph** out) {
  if (!out) return set_error("prh_uneven_graph_create: null output");
  *out = nullptr;
  if (!rf || !rf->ref_off) return set_error("prh_uneven_graph_create: null re-fold lists");
  if (k_begin < 0 || k_end <= k_begin || k_end > (un ? un->nsteps : 0))
    return set_error("prh_uneven_graph_create: step range outside the schedule");
  if (int32_t e = check_uneven(op, un, st, "prh_uneven_graph_create")) return e;
  for (int32_t t = k_begin; t < k_end; ++t)
    if (rf->ref_off[t + 1] < rf->ref_off[t])
      return set_error("prh_uneven_graph_create: re-fold offsets must not decrease");
  hipStream_t cap = nullptr;
  if (int32_t e = check_hip(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking),
                            "prh_uneven_graph_create stream"))
    return e;
  prh_graph* g = new prh_graph();
  g->baked_first = k_begin;
  g->chunk = k_end - k_begin;
  g->nsteps = un->nsteps;
  int32_t status = 0;
  hipError_t err = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (err == hipSuccess) {
    for (int32_t t = k_begin; t < k_end && status == 0; ++t) {
      const int32_t n = rf->ref_off[t + 1] - rf->ref_off[t];
      if (rf->enabled && n > 0)
        status = prh_fold_rows(rf->kern, rf->nrows, rf->dph, rf->phot, rf->in_rows + rf->ref_off[t],
                               rf->out_rows + rf->ref_off[t], n, const_cast<double*>(un->mrhs),
                               op->mvals_stride, cap);
      if (status == 0) status = prh_uneven_step(op, un, st, t, cap);
    }
    err = hipStreamEndCapture(cap, &g->graph);
  }
  if (err == hipSuccess && status == 0) err = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  hipStreamDestroy(cap);
  if (err != hipSuccess || status != 0) {
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    delete g;
    if (status != 0) return status;
    return set_error(std::string("prh_uneven_graph_create: ") + hipGetErrorString(err));
  }
  *out = g;
  return 0;
}

int32_t prh_uneven_step(const prh_operator* op, const prh_uneven* un, const prh_state* st,
                        int32_t step, void* stream) {
  if (int32_t e = check_uneven(op, un, st, "prh_uneven_step")) return e;
  if (step < 0 || step >= un->nsteps) return set_error("prh_uneven_step: step outside the schedule");
  if (st->em) {
    const prh_em* em = st->em;
    if (em->d_em < 1 || em->d_em > PRH_EM_MAX)
      return set_error("prh_uneven_step: d_em must be in [1, 144]");
    if (!em->egrid || !em->dcoef || !em->ddiag || !em->cbh || !em->bhw || !em->rbh || !em->tg ||
        !em->te || !em->gam)
      return set_error("prh_uneven_step: EM sector with a null buffer");
    if (em->nwin != un->nwin)
      return set_error("prh_uneven_step: the EM tables need one entry per column window");
  }
  const int32_t live = un->live ? un->live[step] : op->K;
  if (live < 0 || live > op->K) return set_error("prh_uneven_step: live column count outside [0, K]");
  if (live == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  launch_uneven<false>(*op, *un, *st, step, live, s);
  launch_uneven<true>(*op, *un, *st, step, live, s);
  if (st->em)
    hipLaunchKernelGGL(em_uneven_kernel, dim3((unsigned)live), dim3(256), 0, s, *op, *un, *st,
                       *st->em, step);
  return check_hip(hipGetLastError(), "prh_uneven_step launch");
}

}  // extern "C"
