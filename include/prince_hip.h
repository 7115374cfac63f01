/*
 * prince_hip.h -- C ABI of the MI355X (gfx950) ETD2 propagation library.
 *
 * Every entry point returns 0 on success and a non-zero status on failure;
 * prh_last_error() then describes the failure (thread-local string).  No
 * C++ exception crosses this boundary.  All pointers to device memory are
 * owned by the caller (the Python host holds them as torch tensors); the
 * library never allocates device memory and keeps no pointer beyond the
 * call, except inside a captured step graph (prh_graph_*), which holds the
 * pointers of the operator/schedule/state it was built from until it is
 * destroyed.  This mirrors the caller-owned-buffer rule of the reference's
 * native sparse handle (src/prince_cr/mkl_sparse.py:12-14, 106-108).  Work
 * is enqueued on the given stream and is asynchronous.
 *
 * Data layout in device memory (one GPU, FP64 throughout):
 *   state  x, a, fphi : [nspec][K][d]  (species block, right-hand side,
 *                       energy bin); for K = 1 this is the reference
 *                       state vector order princeidx*d + i.
 *   mvals             : [nwin][mvals_stride]; per rate-cache window the
 *                       diagonal of M (n doubles at diag_off) and the
 *                       off-diagonal values in block-band layout (at the
 *                       block offsets voff).  With two photon-field
 *                       components, mvals_e holds the second component's
 *                       fold and row k of the operator uses
 *                       mvals + escale[k] * mvals_e.
 *   blocks            : [nblocks][4] = {mother species, first offset
 *                       o_lo = j - i, number of offsets nslot, value
 *                       offset voff}; block value (s, i) at voff + s*d + i
 *                       couples row (da, i) to column (mo, i + o_lo + s).
 *   tasks             : [ntasks][6] = {da, blk_begin, blk_end, lead,
 *                       nparts, kind}; task t runs as wave t % 16 of a
 *                       workgroup of row group t / 16.  The nparts parts of
 *                       one species row sit in waves lead .. lead+nparts-1
 *                       of one workgroup and are summed, in part order,
 *                       through LDS.  kind: 0 = only diagonal blocks, 1 =
 *                       only blocks of at most 8 diagonals, 2 = any.
 *                       da < 0 marks an empty wave.
 */
#ifndef PRINCE_HIP_H
#define PRINCE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRH_WG_WAVES 16 /* waves per workgroup of the stage kernels */

typedef struct prh_operator {
  int32_t nspec;   /* species blocks */
  int32_t d;       /* energy bins per block (<= 64) */
  int32_t K;       /* right-hand sides */
  int32_t ntasks;  /* multiple of the workgroup waves (wg_waves, default PRH_WG_WAVES) */
  const int32_t* tasks;
  const int32_t* blocks;
  int64_t mvals_stride; /* doubles per window in mvals */
  int64_t diag_off;     /* offset of the diagonal inside a window */
  const double* mvals;
  const double* egrid;  /* [d] bin-centre energies (adiabatic loss) */
  const double* dcoef;  /* [d][7] FD coefficients for offsets -3..3 */
  const double* ddiag;  /* [d] FD diagonal */
  const double* pscale; /* [nspec] pair-production scale per species */
  const double* rpair;  /* [nwin][d] pair-production rate per window */
  const int32_t* src_slot; /* [nspec] source slot or -1 */
  int32_t nsrc;            /* sources per right-hand side (J) */
  int32_t flags;           /* PRH_F_* */
  const double* src_grid;  /* [nslot][J][K][d] injection grids */
  int32_t kt;              /* right-hand sides per wave: 1, 2, 4 or 8 */
  int32_t pad_;
  const double* mvals_e;   /* [nwin][mvals_stride] second component or NULL */
  const double* rpair_e;   /* [nwin][d] second component's pair rate or NULL */
  const double* escale;    /* [K] second component's scale per rhs or NULL */
  /* Waves per workgroup of the stage kernels: 0 or 16 (default), or 8
   * (better at large K).  tasks are packed for it (ntasks a multiple); the
   * adjoint and uneven kernels take 16 only. */
  int32_t wg_waves;
  /* Dense rows (one-field stages only; ndense = 0: off).  Their blocks
   * from the listed mothers are applied by a mother-major pass over chunks of
   * dchunk species (ndchunk chunks): dense_tab [ndense][nspec][3] = {o_lo,
   * nslot (0: no block), value offset}; dense_row [nspec] = dense index or -1;
   * the chunk partials go to prh_state.dpart and are added in chunk order in
   * those rows' epilogues.  The rows' tasks carry their remaining blocks. */
  int32_t ndense;
  int32_t dchunk;
  int32_t ndchunk;
  const int32_t* dense_row;
  const int32_t* dense_tab;
  /* Dense rows cut into mother chunks inside the stage kernel (nfin > 0, in
   * place of the mother-major pass; dense_tab and dchunk unused): a task
   * whose kind carries PRH_TASK_TO_DPART sums the blocks [b_begin, b_end) of
   * row da and stores them as chunk tasks[3] at dpart[chunk][dense_row[da]];
   * a second launch over fin_tasks [nfin][6] (nfin a multiple of wg_waves)
   * runs those rows' epilogues: each row is nparts waves with an empty block
   * range, wave p adding chunks [p ndchunk / nparts, (p+1) ndchunk / nparts)
   * in order before the parts are summed in part order. */
  int32_t nfin;
  int32_t pad3_;
  const int32_t* fin_tasks;
} prh_operator;

/* Stage task = {da, b_begin, b_end, lead wave, nparts, kind}; kind's low
 * bits: 0 diagonal blocks only, 1 blocks of <= 8 diagonals, 2 general. */
#define PRH_TASK_KIND_MASK 3
#define PRH_TASK_TO_DPART 4 /* dense-row chunk task (see nfin) */
/* Row-pair task (one-field stage kernels without dense-row modes): rows da
 * and da+1, one part each, diagonal blocks only; blocks [b_begin, b_begin +
 * (kind >> PRH_TASK_PAIR_SHIFT)) belong to row da, the rest up to b_end to
 * row da+1.  The wave walks both rows' mothers merged, loading a mother panel
 * both rows read once.  lead is the wave itself, nparts 1. */
#define PRH_TASK_PAIR 8
#define PRH_TASK_PAIR_SHIFT 8

#define PRH_F_PARTIAL_DIFF 1 /* continuous-loss stencil term on */
#define PRH_F_INJECTION 2    /* source term b(z) on */
#define PRH_F_ROW_PAIRS 4    /* tasks hold row-pair tasks (PRH_TASK_PAIR): one-field
                                forward stages only, no dense rows, kt <= 8 */

typedef struct prh_schedule {
  int32_t nsteps;
  int32_t pad_;
  const double* h;        /* [nsteps] z step (negative) */
  const double* dldz;     /* [nsteps] dl/dz at the step start */
  const double* hadia;    /* [nsteps] H(z) * cm2sec (adiabatic kappa / E) */
  const double* finj;     /* [nsteps] dl/dz * cm2sec (injection factor) */
  const int32_t* win;     /* [nsteps] rate-cache window of the step */
  const double* src_coef; /* [nsteps][J][K] norm * evolution(z) per source */
} prh_schedule;

/* Co-evolved EM cascade (reference propagation.py:555-571, 620-665, 836-878):
 * photons on their own grid of d_em <= PRH_EM_MAX bins ride the ETD2 step with
 * the adiabatic FD operator; after the step the Bethe-Heitler pairs of the
 * charged nuclei, w[j][k] = bhw[win][j] * sum_s cbh[s] x[s][j][k], are spread
 * by R_BH into e+ and e- and the window's transfer acts:
 * gam <- T_gamma gam + T_electron (e- + e+), e+- <- 0 (the e+- blocks are
 * zero at every step boundary and are not stored). */
#define PRH_EM_MAX 144
typedef struct prh_em {
  int32_t d_em;         /* EM bins */
  int32_t nwin;         /* windows of bhw, rbh, tg, te */
  const double* egrid;  /* [d_em] EM bin centres */
  const double* dcoef;  /* [d_em][7] FD off-diagonal coefficients (offsets -3..3) */
  const double* ddiag;  /* [d_em] FD diagonal */
  const double* cbh;    /* [nspec] A_s * pair-loss scale of species s */
  const double* bhw;    /* [nwin][d] dl_step * (dE/E) * pair rate of the window */
  const double* rbh;    /* [nwin][d_em][d] Bethe-Heitler e+- shape (one sign) */
  const double* tg;     /* [nwin][d_em][d_em] photon transfer T_gamma */
  const double* te;     /* [nwin][d_em][d_em] e+- transfer T_electron */
  /* [d_em][K] photon state.  Shared-schedule steps alternate: step t reads
   * gam and writes gam2 for even t, the reverse for odd t (after n steps the
   * photons are in gam2 when n is odd); the uneven batch updates gam in place. */
  double* gam;
  double* wbh;          /* [d][K] scratch: Bethe-Heitler weights of the step */
  double* gam2;         /* [d_em][K] the other photon buffer (shared schedule) */
  /* [nwin][d_em][d_em + d] per window [T_gamma | 2 T_electron R_BH]
   * (prh_em_transfer_stack); shared-schedule steps read its Bethe-Heitler
   * columns, the uneven batch uses tg, te and rbh directly. */
  const double* tstack;
  /* [nsteps][d_em][d_em] per step T_gamma[win(t)] P(t), P(t) the photon ETD2
   * step as a matrix (prh_em_step_transfer): the shared-schedule EM step is
   * one product [qstep[t] | 2 T_electron R_BH] [photons; Bethe-Heitler
   * weights].  Not used by the uneven batch. */
  const double* qstep;
} prh_em;

/* Per window w: out[w] = [T_gamma[w] | 2 T_electron[w] R_BH[w]] ([d_em][d_em + d]),
 * the stacked transfer of the shared-schedule EM step: e+ = e- = R_BH w_BH
 * feed T_electron (e- + e+), so T_electron 2 R_BH acts on the Bethe-Heitler
 * weights directly (the reference's _inject_bh_pairs followed by
 * _apply_em_cascade, propagation.py:555-571, 620-665, as one product). */
int32_t prh_em_transfer_stack(const double* tg, const double* te, const double* rbh,
                              int32_t nwin, int32_t d_em, int32_t d, double* out, void* stream);

/* qstep[t] = T_gamma[win(t)] P(t) for every step t of the schedule
 * ([nsteps][d_em][d_em], caller-allocated): P(t) is the photon ETD2 step of
 * the reference's EM block (its two stages with the adiabatic FD operator,
 * propagation.py:867-878 with etd2.py:412-439) written as the matrix it is,
 * so that T_gamma applied after the step (_apply_em_cascade,
 * propagation.py:555-571) and the step itself become one product. */
int32_t prh_em_step_transfer(const prh_operator* op, const prh_schedule* sch, const prh_em* em,
                             double* qstep, void* stream);

/* Opaque handle of a captured step graph (prh_graph_create, prh_uneven_graph_create). */
typedef struct prh_graph prh_graph;

/* Uneven batch: every right-hand side k has its own z grid (its own z_max),
 * refresh windows and coupling values.  Per-step tables are [nsteps][K];
 * win < 0 marks a finished column (its state is left as it is).  mrhs holds
 * each column's window values in the window layout of prh_operator
 * ([nslot][K][mvals_stride], diagonal at diag_off), written by prh_fold_rows:
 * with nslot = 1 a column's only slot is re-folded at the step its window
 * starts; with nslot = 2 window w (a row of rpair) lives in slot w & 1, so
 * the next window of a column can be folded while its current one is in use
 * (windows of one column must then have consecutive ids).  live, a HOST
 * array [nsteps] or NULL (all K), is the number of columns with step t: the
 * columns with win >= 0 at step t must be 0 .. live[t] - 1 (columns ordered
 * by decreasing step count), and only those are launched -- finished columns
 * retire from the grid. */
typedef struct prh_uneven {
  int32_t nsteps;        /* steps of the longest column */
  int32_t nwin;          /* windows over all columns (rows of rpair) */
  const double* h;       /* [nsteps][K] z step of column k (negative) */
  const double* dldz;    /* [nsteps][K] */
  const double* hadia;   /* [nsteps][K] H(z) * cm2sec (0 when adiabatic losses are off) */
  const double* finj;    /* [nsteps][K] dl/dz * cm2sec */
  const int32_t* win;    /* [nsteps][K] window id (row of rpair) or -1 */
  const double* src_coef;/* [nsteps][J][K] norm * evolution(z_k) per source */
  const double* rpair;   /* [nwin][d] pair-production rate of each column window */
  const double* mrhs;    /* [nslot][K][mvals_stride] window values of each column */
  int32_t nslot;         /* 1 or 2 value slots per column */
  int32_t pad_;
  const int32_t* live;   /* HOST [nsteps]: live columns of each step, or NULL */
} prh_uneven;

/* Per-step re-fold lists of the uneven batch for graph capture: before step t
 * the value rows out_rows[ref_off[t] .. ref_off[t+1]) of mrhs (slot * K +
 * column) are folded from the photon rows in_rows[...] (device arrays);
 * ref_off is a HOST array of nsteps + 1.  With two slots the windows starting
 * in a group of steps can all be folded before its first step. */
typedef struct prh_refold {
  const double* kern;      /* [nrows][dph] window-layout fold kernel */
  int64_t nrows;
  int32_t dph;
  int32_t enabled;         /* 0: photo-hadronic coupling off, no re-folds */
  const double* phot;      /* [nwin][dph] photon vectors of all column windows */
  const int32_t* in_rows;  /* device */
  const int32_t* out_rows; /* device */
  const int32_t* ref_off;  /* host, [nsteps + 1] */
} prh_refold;

typedef struct prh_state {
  double* x;         /* [nspec][K][d] solution, updated in place */
  double* a;         /* [nspec][K][d] stage-1 vector */
  double* fphi;      /* [nspec][K][d] F(x) of stage 1 */
  int32_t* step_ctl; /* [2] device step window {first, end} used by graphs */
  /* Optional record of each step of a chunk (NULL: off): x, a, F(x), F(a)
   * at [step - rec_first][nspec][K][d] (graph replay: [step - step_ctl[0]]). */
  double* rec_x;
  double* rec_a;
  double* rec_fx;
  double* rec_fa;
  int32_t rec_first;
  int32_t rec_count;
  const prh_em* em;  /* host pointer: EM sector stepped after stage 2, or NULL */
  double* dpart;     /* [ndchunk][ndense][K][d] dense-row chunk partials, or NULL */
} prh_state;

/* Adjoint sweep buffers (all [nspec][K][d] unless noted). */
typedef struct prh_adjoint {
  double* lam;   /* dJ/dx at the end of the step on entry, at its start on exit */
  double* lam_a; /* dJ/da of the stage-1 vector */
  double* u;     /* h phi2 lam' (written by the previous adjoint step or the seed) */
  double* v;     /* h phi1 lam_a - u */
  const double* rec_x; /* recorded forward chunk, see prh_state */
  const double* rec_a;
  const double* rec_fx;
  const double* rec_fa;
  int32_t rec_first;
  int32_t rec_count;
  double* gsrc0;       /* [nslot][J][K][d] += dl/dz*cm2sec * coef * h phi1 lam_a, or NULL */
  double* gsrc1;       /* same weighted with ln(1+z) */
  const double* lnz1;  /* [nsteps] ln(1 + z_step) */
  double* gdot;        /* += per-element terms of dJ/d(escale), or NULL */
} prh_adjoint;

/* Library version and last error (thread-local, never NULL). */
int32_t prh_version(void);
const char* prh_last_error(void);

/* Rate fold for many photon vectors at once (FP64 MFMA GEMM):
 *   out[w*out_stride + r] = sum_j phot[w*dph + j] * kern[r*dph + j]
 * for w < nwin, r < nrows.  Replaces the per-window dense fold
 * M.data = B @ n_gamma of src/prince_cr/interaction_rates.py:1051-1061
 * (host) and :930-964 (GPU), and the split fold of :824-847, evaluated for
 * every refresh window of a solve in one launch.  dph <= 4096 (above 128 the
 * kernel rows are re-read per window tile instead of held in registers: the
 * cascade's tensor contractions, cascade/cascade.py:78-188). */
int32_t prh_fold_gemm(const double* kern, int64_t nrows, int32_t dph, const double* phot,
                      int32_t nwin, double* out, int64_t out_stride, void* stream);

/* Row gather dst[r] = sel[r] >= 0 ? src[sel[r]] : 0 for rows of ncols
 * doubles.  Builds the split fold kernel of
 * src/prince_cr/interaction_rates.py:780-822 (rows of B that feed M_off and
 * the diagonal) directly in block-band order. */
int32_t prh_gather_rows(const double* src, const int64_t* sel, int64_t nrows, int32_t ncols,
                        double* dst, void* stream);

/* ETD2 steps k_begin <= k < k_end on the device (two fused stage kernels
 * per step, launched one by one).  Replaces the step loop of
 * src/prince_cr/solvers/etd2.py:478-517 with the step body of
 * etd2.py:412-439 / 442-475 and the operator of
 * src/prince_cr/solvers/propagation.py:1360-1508 (apply_F, diagonal d,
 * kappa = kappa_adia + kappa_pair, injection b). */
int32_t prh_etd2_steps(const prh_operator* op, const prh_schedule* sch, const prh_state* st,
                       int32_t k_begin, int32_t k_end, void* stream);

/* Allocation-free variant of prh_etd2_steps.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * hipGraph stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The caller passes the operator, the schedule and the state x
 * ([nspec][K][d], updated in place); the stage vectors a and F(x) and the
 * dense-row chunk partials come from one static device array per GPU of
 * PRH_SCRATCH_BYTES (held until the process ends), whose address is looked up
 * once per device on the first call (the device current at the call is the
 * stream's device).  Needs 256 + 16 * nspec * K * d bytes plus
 * 8 * ndchunk * ndense * K * d with dense-row chunk tasks, else the call
 * fails with an error naming both sizes.  Narrow layout, no EM sector, no
 * step recording.  It allocates nothing and keeps nothing between calls;
 * calls that overlap on one GPU (other streams, other threads) share the
 * array and race. */
#define PRH_SCRATCH_BYTES (64LL << 20)
int32_t prh_etd2_steps_noalloc(const prh_operator* op, const prh_schedule* sch, double* x,
                               int32_t k_begin, int32_t k_end, void* stream);

/* One stage (1 or 2) of ETD2 step `step`: the same kernels as
 * prh_etd2_steps, launched individually (per-kernel timing, debugging). */
int32_t prh_etd2_stage(const prh_operator* op, const prh_schedule* sch, const prh_state* st,
                       int32_t step, int32_t stage, void* stream);

/* Adjoint of ETD2 step `step` (two transposed stage kernels).  opT is the
 * operator in transposed block layout with the transposed FD stencil.  No
 * counterpart in the reference (which has no adjoint); it is the reverse of
 * prh_etd2_steps for one step. */
int32_t prh_adjoint_step(const prh_operator* opT, const prh_schedule* sch, const prh_adjoint* adj,
                         int32_t step, void* stream);
/* Adjoint steps k_end-1 down to k_begin (prh_adjoint_step in a loop). */
int32_t prh_adjoint_steps(const prh_operator* opT, const prh_schedule* sch, const prh_adjoint* adj,
                          int32_t k_begin, int32_t k_end, void* stream);
/* One EM-sector step (photon ETD2 step, Bethe-Heitler injection, cascade
 * transfer) after the nuclear step `step`; `st` supplies the post-step nuclear
 * state, the photons move between em->gam and em->gam2 by the parity of
 * `step` (see prh_em).  prh_etd2_steps, prh_etd2_stage (stage 2) and graphs
 * run it by themselves when st->em is set.  Replaces the step hook of the reference's
 * solve loop (propagation.py:867-878): _inject_bh_pairs (620-665) followed by
 * _apply_em_cascade (555-571), with the photon blocks of the ETD2 step
 * (etd2.py:412-439). */
int32_t prh_em_step(const prh_operator* op, const prh_schedule* sch, const prh_state* st,
                    int32_t step, void* stream);

/* Photon-field tables for nw (z, s) pairs: out[w][i] = sum_m W[i][m] *
 * (n_CMB(P[i][m], z_w) + s_w (1 + z_w)^zpow E0[i][m]), the CMB Planck density
 * planck_pref E^2 / (exp(E / (e_cmb (1+z))) - 1) (zero where the ratio >= 709)
 * plus a z-separable component tabulated at z = 0.  Gives the per-window photon
 * vectors (interaction_rates.py:105-116, W = 1, nm = 1) and pair-loss rates
 * (interaction_rates.py:1277-1312, W = trapezoid weight * phi(xi)/xi^2) of the
 * uneven batch, whose columns each see CMB + s_k EBL at their own windows. */
int32_t prh_field_tables(const double* zs, const double* ss, int32_t nw, const double* P,
                         const double* W, const double* E0, int32_t ni, int32_t nm,
                         double planck_pref, double e_cmb, double zpow, double* out,
                         void* stream);

/* Rate fold of listed columns: out[out_rows[c]][r] = sum_j kern[r][j] phot[in_rows[c]][j]
 * -- the columns whose window changes at a step of the uneven batch
 * (interaction_rates.py:824-847 per column). */
int32_t prh_fold_rows(const double* kern, int64_t nrows, int32_t dph, const double* phot,
                      const int32_t* in_rows, const int32_t* out_rows, int32_t ncol, double* out,
                      int64_t out_stride, void* stream);

/* Both ETD2 stages of step `step` of the uneven batch (kt 1/2/4).
 * Each column follows its own reference solve (propagation.py:813-890 with
 * etd2.py:412-439); only the un->live[step] leading columns are launched.  With st->em set, the EM
 * step of every live column follows (its own window's T_gamma, T_electron,
 * R_BH and Bethe-Heitler weights: tables indexed by the column windows of
 * un->win, em->nwin == un->nwin; the photon state em->gam [d_em][K] is
 * updated in place, em->gam2 and em->wbh are not used): the reference's step
 * hook (propagation.py:867-878) per column. */
int32_t prh_uneven_step(const prh_operator* op, const prh_uneven* un, const prh_state* st,
                        int32_t step, void* stream);

/* Capture steps [k_begin, k_end) of the uneven batch -- each step's column
 * re-fold (prh_fold_rows) and both stages (prh_uneven_step), steps baked in --
 * as one graph; run it with prh_graph_run(g, k_begin, k_end, stream) and free
 * it with prh_graph_destroy.  The 64-step chunks of the uneven solve. */
int32_t prh_uneven_graph_create(const prh_operator* op, const prh_uneven* un, const prh_state* st,
                                const prh_refold* rf, int32_t k_begin, int32_t k_end,
                                prh_graph** out);

/* u = h phi2 lam for step `step`, before the first prh_adjoint_step of a sweep. */
int32_t prh_adjoint_seed(const prh_operator* opT, const prh_schedule* sch, const prh_adjoint* adj,
                         int32_t step, void* stream);

/* Step graphs: a hipGraph holding `chunk` steps (2*chunk stage kernels and
 * one step-window advance) whose kernels read their step from
 * st->step_ctl, so one instantiated graph replays every chunk of a solve.
 * Replaces the per-window CUDA-graph capture/replay of
 * src/prince_cr/solvers/propagation.py:2186-2337. */
int32_t prh_graph_create(const prh_operator* op, const prh_schedule* sch, const prh_state* st,
                         int32_t chunk, prh_graph** out);
int32_t prh_graph_run(prh_graph* g, int32_t k_begin, int32_t k_end, void* stream);
int32_t prh_graph_destroy(prh_graph* g);

/* ---------------------------------------------------------------------------
 * Rate kernel rebuilt on the device from its Toeplitz factors.
 * ------------------------------------------------------------------------- */
typedef struct prh_toeplitz {
  int32_t d;            /* cosmic-ray bins */
  int32_t dph;          /* photon bins */
  int32_t n1;           /* 1-D tables (inclusive and non-elastic response differences) */
  int32_t n2;           /* 2-D tables (redistributed channels) */
  int64_t nrows;        /* kernel rows (descriptors), in the CSR order of the coupling matrix */
  const int32_t* desc;  /* [nrows][5]: kind (0 boost-conserving, 1 redistributed), t_gain, t_loss, i, k */
  const double* tab1d;  /* [n1][d + dph - 1] */
  const double* tab2d;  /* [n2][d + dph - 1][2 d - 1] */
  const double* fmo;    /* [d] m_p / E_i */
  const double* fdiff;  /* [d][d] (m_p / E_i) dE_i / dE_k */
} prh_toeplitz;

/* Kernel rows sel[r] (-1: a zero row) expanded into out [nsel][dph]: row s
 * with descriptor (kind, tg, tl, i, k) is, at photon bin j (w = i + j),
 *   kind 0: fmo[i] tab1d[tg][w] - fmo[i] tab1d[tl][w]
 *   kind 1: max(fdiff[i][k] tab2d[tg][w][k - i + d - 1], 0) - [k == i] fmo[i] tab1d[tl][w]
 * (a term with table index -1 absent) -- the tile construction of
 * interaction_rates.py:488-595 for boost-conserving and redistributed
 * channels, evaluated on the device from the per-channel tables instead of
 * uploading the dense kernel (the reference's _batch_matrix). */
int32_t prh_toeplitz_rows(const prh_toeplitz* tz, const int64_t* sel, int64_t nsel, double* out,
                          void* stream);

/* ---------------------------------------------------------------------------
 * EM-cascade window operators built on the device.
 * ------------------------------------------------------------------------- */
#define PRH_BH_MAX 64 /* target-photon points of the Bethe-Heitler kernel */

/* Field-free cascade data of one EM grid (device pointers). */
typedef struct prh_cascade {
  int32_t n;           /* EM bins, <= 144 */
  int32_t d;           /* nucleon bins of the Bethe-Heitler source, <= 64 */
  int32_t nb;          /* target-photon points of the Bethe-Heitler kernel, <= PRH_BH_MAX */
  int32_t max_scatter; /* iteration cap of the kinetic cascade */
  double tol;          /* energy left in G and e, relative to each column's input, at convergence */
  double cool_floor;   /* a column whose IC emission carries less than this fraction of E cools */
  double bh_emin;      /* lowest nucleon energy with a Bethe-Heitler source */
  const double* E;         /* [n] EM grid centres */
  const double* dE;        /* [n] bin widths (gradient of E) */
  const double* pref_emis; /* [n] IC emission prefactor (3/4) sigma_T c / gamma^2 */
  const double* pref_pair; /* [n] pair-injection prefactor */
  const int32_t* cic_b;    /* [n][n] (k, j): lower bin of E_j - E_k, -1 off the grid */
  const double* cic_f;     /* [n][n] weight of that lower bin */
  const double* Ecr;       /* [d] nucleon grid */
  const double* eps_b;     /* [nb] Bethe-Heitler target-photon grid */
  const double* kern_b;    /* [n][d][nb] Bethe-Heitler kernel */
} prh_cascade;

/* Kinetic single-scatter cascade transfer of nwin windows (one 16-wave
 * workgroup each).  contr [nwin][cstride]: the window's target densities
 * contracted with the field-free tensors (prh_fold_gemm): IC emission
 * [n][n], pair injection [n][n], IC power [n], IC rate [n], sum w n,
 * sum w eps n; dtau [nwin][n] the opacity per unit length, dl [nwin] the
 * path length of one step.  work: 9 * 144 * 144 doubles per window.
 * Outputs T_gamma, T_electron [nwin][n][n] and the iteration counts.
 * Replaces the per-window host builds of the reference's cascade
 * (cascade/cascade.py:331-407, batched form 444-589). */
int32_t prh_cascade_transfer(const prh_cascade* cs, const double* contr, int64_t cstride,
                             const double* dtau, const double* dl, int32_t nwin, double* work,
                             int64_t work_bytes, double* tg, double* te, int32_t* iters,
                             void* stream);

/* Bethe-Heitler e+- shapes R [nwin][n][d] from the target densities at the
 * kernel's photon grid [nwin][nb] (cascade/bethe_heitler.py:219-265,
 * batched 319-351). */
int32_t prh_cascade_bh(const prh_cascade* cs, const double* nb_dens, int32_t nwin, double* rbh,
                       void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PRINCE_HIP_H */
