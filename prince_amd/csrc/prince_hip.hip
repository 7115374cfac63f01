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
constexpr int SWG = PRH_WG_WAVES;     // waves per stage workgroup
constexpr int STHREADS = SWG * WAVE;  // 1024

__device__ __forceinline__ double shfl_or_zero(double v, int src, int d) {
  const double r = __shfl(v, src < 0 ? 0 : (src > 63 ? 63 : src), WAVE);
  return (src >= 0 && src < d) ? r : 0.0;
}

// step_local: step index (use_ctl = 0) or offset from step_ctl[0] (use_ctl = 1,
// graph replay; steps at or beyond step_ctl[1] do nothing).
// Blocks whose loads are issued together before any of them is used: runs
// of diagonal blocks and of band blocks with at most NSMAX diagonals.
constexpr int NSMAX = 8;
template <int KT>
struct Batch {
  static constexpr int DIAG = KT <= 2 ? 8 : (KT == 4 ? 4 : 2);
  static constexpr int BAND = KT == 1 ? 4 : (KT == 2 ? 2 : 1);
};

enum PartKind { PART_DIAG = 0, PART_BAND = 1, PART_GENERIC = 2 };

// step_local: step index (use_ctl = 0) or offset from step_ctl[0] (use_ctl = 1,
// graph replay; steps at or beyond step_ctl[1] do nothing).
template <bool STAGE2, int KT>
__global__ void __launch_bounds__(STHREADS) etd2_stage(prh_operator op, prh_schedule sch,
                                                       prh_state st, int step_local,
                                                       int use_ctl) {
  __shared__ double red[SWG][KT][WAVE];
  int step = step_local;
  if (use_ctl) {
    step = st.step_ctl[0] + step_local;
    if (step >= st.step_ctl[1]) return;  // uniform over the grid
  }
  const int lane = threadIdx.x & (WAVE - 1);
  const int wv = threadIdx.x / WAVE;
  const int K = op.K;
  const int d = op.d;
  // XCD-aware placement: workgroup L runs on XCD L % 8; when the number of
  // right-hand-side groups is a multiple of 8, every XCD owns whole groups
  // so their state panels stay in that XCD's L2.
  const int nx = op.ntasks / SWG;
  const int ny = (K + KT - 1) / KT;
  const int L = blockIdx.x;
  int gx, gy;
  if (ny % 8 == 0) {
    const int j = L >> 3;
    gy = (L & 7) + 8 * (j / nx);
    gx = j % nx;
  } else {
    gx = L % nx;
    gy = L / nx;
  }
  const int k0 = gy * KT;
  const int* tk = op.tasks + 6 * ((int64_t)gx * SWG + wv);
  const int da = tk[0];
  const int b_begin = tk[1];
  const int b_end = tk[2];
  const int lead = tk[3];
  const int nparts = tk[4];
  const int kind = tk[5];
  const bool active = lane < d;

  const double* __restrict__ xin = STAGE2 ? st.a : st.x;
  const int w = sch.win[step];
  const double* __restrict__ mv = op.mvals + (int64_t)w * op.mvals_stride;

  // ---- off-diagonal coupling over this part's blocks, in column order
  double acc[KT];
#pragma unroll
  for (int kk = 0; kk < KT; ++kk) acc[kk] = 0.0;
  if (da >= 0 && kind == PART_DIAG) {
    constexpr int UB = Batch<KT>::DIAG;
    for (int b = b_begin; b < b_end; b += UB) {
      double m[UB];
      double xv[UB][KT];
#pragma unroll
      for (int u = 0; u < UB; ++u) {
        const int bb = (b + u < b_end) ? b + u : b_end - 1;
        const int* blk = op.blocks + 4 * bb;
        const int mo = blk[0];
        const int64_t voff = blk[3];
        m[u] = (active && b + u < b_end) ? mv[voff + lane] : 0.0;
#pragma unroll
        for (int kk = 0; kk < KT; ++kk) {
          const int k = k0 + kk;
          xv[u][kk] = (active && k < K) ? xin[((int64_t)mo * K + k) * d + lane] : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < UB; ++u) {
#pragma unroll
        for (int kk = 0; kk < KT; ++kk) acc[kk] += m[u] * xv[u][kk];
      }
    }
  } else if (da >= 0 && kind == PART_BAND) {
    constexpr int UB = Batch<KT>::BAND;
    for (int b = b_begin; b < b_end; b += UB) {
      double m[UB][NSMAX];
      double xv[UB][KT];
      int olo[UB], ns[UB];
#pragma unroll
      for (int u = 0; u < UB; ++u) {
        const bool valid = b + u < b_end;
        const int bb = valid ? b + u : b_end - 1;
        const int* blk = op.blocks + 4 * bb;
        const int mo = blk[0];
        olo[u] = blk[1];
        ns[u] = valid ? blk[2] : 0;
        const int64_t voff = blk[3];
        const int nsl = blk[2];
#pragma unroll
        for (int s2 = 0; s2 < NSMAX; ++s2) {
          const int sc = s2 < nsl ? s2 : nsl - 1;
          m[u][s2] = active ? mv[voff + (int64_t)sc * d + lane] : 0.0;
        }
#pragma unroll
        for (int kk = 0; kk < KT; ++kk) {
          const int k = k0 + kk;
          xv[u][kk] = (active && k < K) ? xin[((int64_t)mo * K + k) * d + lane] : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < UB; ++u) {
#pragma unroll
        for (int s2 = 0; s2 < NSMAX; ++s2) {
          if (s2 < ns[u]) {
            const int src = lane + olo[u] + s2;
#pragma unroll
            for (int kk = 0; kk < KT; ++kk) acc[kk] += m[u][s2] * shfl_or_zero(xv[u][kk], src, d);
          }
        }
      }
    }
  } else if (da >= 0) {
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
        xv[kk] = (active && k < K) ? xin[((int64_t)mo * K + k) * d + lane] : 0.0;
      }
      for (int s2 = 0; s2 < nslot; ++s2) {
        const int src = lane + olo + s2;
        const double m = active ? mv[voff + (int64_t)s2 * d + lane] : 0.0;
#pragma unroll
        for (int kk = 0; kk < KT; ++kk) acc[kk] += m * shfl_or_zero(xv[kk], src, d);
      }
    }
  }

  // ---- combine the parts of split rows through LDS (in part order)
  if (nparts > 1) {
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) red[wv][kk][lane] = acc[kk];
  }
  __syncthreads();
  if (da < 0 || wv != lead) return;
  if (nparts > 1) {
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) {
      double sum = red[lead][kk][lane];
      for (int p = 1; p < nparts; ++p) sum += red[lead + p][kk][lane];
      acc[kk] = sum;
    }
  }

  // ---- epilogue: row block da, bins = lanes, KT right-hand sides
  const double dl = sch.dldz[step];
  const double hh = sch.h[step];
  const bool pdiff = (op.flags & PRH_F_PARTIAL_DIFF) != 0;
  const bool inj = (op.flags & PRH_F_INJECTION) != 0 && op.src_slot[da] >= 0;
  double kap = 0.0;
  double dval = 0.0;
  if (active) {
    kap = sch.hadia[step] * op.egrid[lane];
    kap = kap + op.pscale[da] * op.rpair[(int64_t)w * d + lane];
    dval = mv[op.zslots + (int64_t)da * d + lane] * dl;
    if (pdiff) {
      double kxd = kap * op.ddiag[lane];
      kxd = kxd * dl;
      dval = dval + kxd;
    }
  }
  // diagonal phi factors (solvers/etd2.py:329-390)
  const double hd = dval * hh;
  const double e = exp(hd);
  const double ahd = fabs(hd);
  double phi;
  if (!STAGE2) {
    phi = (ahd > PHI1_SMALL) ? (e - 1.0) / hd : (hd * (1.0 / 6.0) + 0.5) * hd + 1.0;
  } else {
    phi = (ahd > PHI2_SMALL) ? ((e - 1.0) - hd) / (hd * hd)
                             : (hd * (1.0 / 24.0) + (1.0 / 6.0)) * hd + 0.5;
  }
  double dcf[7];
#pragma unroll
  for (int o = 0; o < 7; ++o) dcf[o] = (active && pdiff) ? op.dcoef[lane * 7 + o] : 0.0;
  const double finj = sch.finj[step];
  const int slot = inj ? op.src_slot[da] : 0;
  const int J = op.nsrc;

#pragma unroll
  for (int kk = 0; kk < KT; ++kk) {
    const int k = k0 + kk;
    if (k >= K) break;  // uniform over the wave
    const int64_t own = ((int64_t)da * K + k) * d + lane;
    const double xown = active ? xin[own] : 0.0;
    double out = acc[kk];
    if (pdiff) {
      const double kx = kap * xown;
      double dx = 0.0;
#pragma unroll
      for (int o = -3; o <= 3; ++o) {
        if (o == 0) continue;
        const double kxo = shfl_or_zero(kx, lane + o, d);
        if (dcf[o + 3] != 0.0) dx += dcf[o + 3] * kxo;
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
    if (!STAGE2) {
      double a = e * xown;
      double sc = phi * out;
      sc = sc * hh;
      a = a + sc;
      st.a[own] = a;
      st.fphi[own] = out;
    } else {
      double sc = out - st.fphi[own];
      sc = sc * hh;
      sc = sc * phi;
      st.x[own] = xown + sc;
    }
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

int32_t validate(const prh_operator* op, const prh_schedule* sch, const prh_state* st,
                 const char* who) {
  if (!op || !sch || !st) return set_error(std::string(who) + ": null argument");
  if (op->d <= 0 || op->d > WAVE) return set_error(std::string(who) + ": d must be in [1, 64]");
  if (op->K <= 0) return set_error(std::string(who) + ": K must be >= 1");
  if (op->ntasks % SWG != 0)
    return set_error(std::string(who) + ": ntasks must be a multiple of 16");
  if (!(op->kt == 1 || op->kt == 2 || op->kt == 4 || op->kt == 8))
    return set_error(std::string(who) + ": kt must be 1, 2, 4 or 8");
  if ((int64_t)(op->ntasks / SWG) * ((op->K + op->kt - 1) / op->kt) > 0x7fffffff)
    return set_error(std::string(who) + ": grid too large");
  return 0;
}

template <bool S2>
void launch_stage_kt(const prh_operator& op, const prh_schedule& sch, const prh_state& st,
                     int step_local, int use_ctl, hipStream_t s) {
  const int64_t nwg = (int64_t)(op.ntasks / SWG) * ((op.K + op.kt - 1) / op.kt);
  const dim3 grid((unsigned)nwg);
  switch (op.kt) {
    case 1:
      hipLaunchKernelGGL((etd2_stage<S2, 1>), grid, dim3(STHREADS), 0, s, op, sch, st, step_local,
                         use_ctl);
      break;
    case 2:
      hipLaunchKernelGGL((etd2_stage<S2, 2>), grid, dim3(STHREADS), 0, s, op, sch, st, step_local,
                         use_ctl);
      break;
    case 4:
      hipLaunchKernelGGL((etd2_stage<S2, 4>), grid, dim3(STHREADS), 0, s, op, sch, st, step_local,
                         use_ctl);
      break;
    default:
      hipLaunchKernelGGL((etd2_stage<S2, 8>), grid, dim3(STHREADS), 0, s, op, sch, st, step_local,
                         use_ctl);
      break;
  }
}

}  // namespace

struct prh_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int32_t chunk = 0;
  int32_t nsteps = 0;
  int32_t* step_ctl = nullptr;
};

extern "C" {

int32_t prh_version(void) { return 2; }

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
  if (int32_t e = validate(op, sch, st, "prh_etd2_steps")) return e;
  if (k_begin < 0 || k_end > sch->nsteps || k_begin > k_end)
    return set_error("prh_etd2_steps: step range outside the schedule");
  if (op->ntasks <= 0 || k_begin == k_end) return 0;
  hipStream_t s = (hipStream_t)stream;
  for (int k = k_begin; k < k_end; ++k) {
    launch_stage_kt<false>(*op, *sch, *st, k, 0, s);
    launch_stage_kt<true>(*op, *sch, *st, k, 0, s);
  }
  return check_hip(hipGetLastError(), "prh_etd2_steps launch");
}

int32_t prh_etd2_stage(const prh_operator* op, const prh_schedule* sch, const prh_state* st,
                       int32_t step, int32_t stage, void* stream) {
  if (int32_t e = validate(op, sch, st, "prh_etd2_stage")) return e;
  if (step < 0 || step >= sch->nsteps) return set_error("prh_etd2_stage: step outside the schedule");
  if (stage != 1 && stage != 2) return set_error("prh_etd2_stage: stage must be 1 or 2");
  if (op->ntasks <= 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  if (stage == 1)
    launch_stage_kt<false>(*op, *sch, *st, step, 0, s);
  else
    launch_stage_kt<true>(*op, *sch, *st, step, 0, s);
  return check_hip(hipGetLastError(), "prh_etd2_stage launch");
}

int32_t prh_graph_create(const prh_operator* op, const prh_schedule* sch, const prh_state* st,
                         int32_t chunk, prh_graph** out) {
  if (int32_t e = validate(op, sch, st, "prh_graph_create")) return e;
  if (!out) return set_error("prh_graph_create: null output");
  if (!st->step_ctl) return set_error("prh_graph_create: state has no step_ctl buffer");
  if (chunk < 1 || chunk > 4096) return set_error("prh_graph_create: chunk must be in [1, 4096]");
  if (op->ntasks <= 0) return set_error("prh_graph_create: empty operator");
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

}  // extern "C"
