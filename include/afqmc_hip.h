/*
 * afqmc_hip.h -- C-ABI of the MI355X walker step (plain pointers and sizes).
 *
 * The reference has no plugin or FFI seam: its driver calls C++ templates
 * directly.  Each entry point below replaces one of those calls and keeps the
 * reference's argument meaning and in-memory layouts (row-major, complex as
 * interleaved double pairs, walker index LAST for G/vbias/X/vHS and FIRST for
 * the walker set), so a caller holding the reference's arrays can switch by
 * swapping the call.  All functions return an afqmc_status code; none throws.
 * Each context owns one HIP stream; calls are ordered on it, and only calls
 * that return host data synchronise.
 *
 * Reference paths are relative to the miniAFQMC source root.
 */
#ifndef AFQMC_HIP_H
#define AFQMC_HIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  AFQMC_OK = 0,
  AFQMC_ERR_ARG = 1,         /* null pointer / bad enum */
  AFQMC_ERR_SHAPE = 2,       /* inconsistent dimensions (reference: assert) */
  AFQMC_ERR_ALLOC = 3,       /* host or device allocation failed */
  AFQMC_ERR_HIP = 4,         /* HIP runtime error (message in afqmc_ctx_error) */
  AFQMC_ERR_SINGULAR = 5,    /* exact zero pivot in an overlap matrix (reference: LAPACK info>0) */
  AFQMC_ERR_NONFINITE = 6,   /* NaN/Inf in an input field buffer */
  AFQMC_ERR_STATE = 7,       /* call out of order (e.g. no system uploaded) */
  AFQMC_ERR_UNSUPPORTED = 8  /* valid request this build does not implement */
} afqmc_status;

/* Arithmetic of the Cholesky contractions and the propagator.  FP64 is kept
 * for the overlap matrices, their LU/inverse, the log-overlap ratios, the
 * orthogonalisation and every energy accumulation in all modes. */
typedef enum {
  AFQMC_FP64_CSR = 0,   /* reference algorithm: CSR csrmm for vbias/vHS/Vakbl, FP64 everywhere */
  AFQMC_FP64_DENSE = 1, /* dense half-rotated / full Cholesky GEMMs on FP64 MFMA */
  AFQMC_MIXED_FP32 = 2, /* complex FP32 MFMA contractions and propagator, FP32 accumulation */
  AFQMC_MIXED_BF16 = 3  /* split-BF16 (hi+lo, 3 products) MFMA contractions and propagator, FP32 accumulation */
} afqmc_mode;

/* CSR in the layout of SparseMatrix<T> (src/Matrix/SparseMatrix.hpp:36-494):
 * 0-based rowptr[nrows+1], col[nnz], val[2*nnz] (complex interleaved),
 * rows sorted by (row, col). */
typedef struct {
  int nrows, ncols;
  long long nnz;
  const int* rowptr;
  const int* col;
  const double* val;
} afqmc_csr_view;

typedef struct afqmc_ctx afqmc_ctx;

/* Buffers addressable with afqmc_copy(), in the reference's layouts
 * (src/miniapps/miniafqmc_base.cpp:213-224). */
typedef enum {
  AFQMC_BUF_W = 0,       /* walkers        [nwalk][2][M][N] */
  AFQMC_BUF_WDATA = 1,   /* walker data    [nwalk][8] */
  AFQMC_BUF_G = 2,       /* compact G      [2*N*M][nwalk] */
  AFQMC_BUF_VBIAS = 3,   /* force bias     [K][nwalk] */
  AFQMC_BUF_X = 4,       /* shifted fields [K][nwalk] */
  AFQMC_BUF_VHS = 5,     /* HS potential   [M*M][nwalk] */
  AFQMC_BUF_HYBRIDW = 6, /* hybrid weight  [nwalk] */
  /* terms of the last local energy, [nwalk][3] = (one-body, Coulomb,
   * exchange), E_L = their sum; read-only.  With Vakbl the whole two-body
   * energy sum G.(Vakbl G)/2 (energy.hpp:93-95) is the second term and the
   * third is zero. */
  AFQMC_BUF_EPARTS = 7,
  AFQMC_BUF_GFULL = 8    /* full G [2*M*M][nwalk] (non-transposed bias only) */
} afqmc_buffer;

/* ---- context ------------------------------------------------------------ */
/* Replaces base::afqmc_sys::setup (src/AFQMC/afqmc_sys.hpp:61-93): sizes all
 * device workspaces for M orbitals, N electrons per spin, K Cholesky vectors
 * and nwalk walkers. */
int afqmc_ctx_create(int nmo, int naea, int nchol, int nwalk, int mode, afqmc_ctx** out);
int afqmc_ctx_destroy(afqmc_ctx* ctx);
/* message of the last failed call on this context ("" if none) */
const char* afqmc_ctx_error(const afqmc_ctx* ctx);
int afqmc_sync(afqmc_ctx* ctx);

/* Options (set before afqmc_upload_system). */
typedef enum {
  /* 1 (default): keep the same-spin exchange block of Vakbl dense in device
   * memory, [2][NM][NM] in the contraction precision, when it fits, so the
   * two-body energy is one GEMM per spin; 0: always use the Cholesky form. */
  AFQMC_OPT_DENSE_EXCHANGE = 0,
  /* mixed modes only: precision level of the local energy.
   *   0: the whole energy in the contraction precision;
   *   1 (default): FP64 density matrix (rebuilt from the FP64 inverses),
   *      one-body and Coulomb terms in FP64 (an FP64 copy of the half-rotated
   *      vectors, 2*K*N*M complex doubles, is kept), exchange in the
   *      contraction precision;
   *   2: as 1 with the exchange in FP32;
   *   3: as 1 with the exchange in FP64 (dense exchange block in FP64). */
  AFQMC_OPT_ENERGY_PRECISION = 1,
  /* 0 (default): orthogonalise through two Cholesky QR passes (FP64 MFMA
   * GEMMs) and the sign reconstruction, which reproduce the LAPACK LQ factor;
   * matrices too ill-conditioned for that route (and N > 128) use the blocked
   * Householder kernel.  1: always the blocked Householder kernel.  May be
   * changed at any time. */
  AFQMC_OPT_ORTHO_HOUSEHOLDER = 2,
  /* the driver's -t flag (src/miniapps/miniafqmc_base.cpp:256-276).  1
   * (default): compact G and the half-rotated SpvnT for the bias.  0: full G
   * and T(Spvn) (AFQMC_FP64_CSR only; set before afqmc_upload_system, which
   * then also builds the transposed CSR of Spvn). */
  AFQMC_OPT_TRANSPOSED = 3,
  /* dense exchange in FP32 or split-BF16 (mixed levels 0-2) only: the
   * exchange GEMM sums the columns of the exchange block in chunks of this
   * many (a positive multiple of 256; default 2048), each chunk accumulated
   * in FP32 on the matrix cores and the chunks' contributions to the energy
   * added in FP64, which bounds the FP32 accumulation error.  0: one pass
   * over all columns.  May be changed at any time.  No counterpart in the
   * reference (its energy is FP64 throughout). */
  AFQMC_OPT_EXCHANGE_CHUNK = 4,
  /* AFQMC_MIXED_BF16 with M a multiple of 64 up to 512 and N a multiple of
   * 16: 1 runs the whole propagator (P W, the six Taylor terms, P U) as one
   * kernel per walker and 32 columns, U kept in LDS; 0 (default) runs it as
   * eight GEMM launches, which measured 2 % faster at 512 orbitals.  Both
   * give the same walkers bit for bit.  May be changed at any time. */
  AFQMC_OPT_FUSED_PROPAGATOR = 5,
  /* AFQMC_MIXED_BF16: 2 (default) runs the GEMMs whose operands are both
   * split-BF16 with k contiguous (vHS, the bias product, the Taylor steps
   * after the first, P U) with both operands staged from global memory
   * straight into a ring of LDS stages by 16-byte direct loads, and the
   * others as 1; 1 stages the split-BF16 left operand (and a right one that
   * allows it) by direct loads into two LDS stages; 0 passes every operand
   * through registers.  All give the same results bit for bit.  May be
   * changed at any time.  No counterpart in the reference. */
  AFQMC_OPT_DIRECT_LDS = 6,
  /* 0 (default): each GEMM call site runs the tile shape the autotuner
   * (afqmc_tune) or the size heuristic picked; k > 0 forces candidate k of
   * the call site's kind on every GEMM (tests cover every shape this way;
   * results are the same bit for bit).  No counterpart in the reference. */
  AFQMC_OPT_TILE = 7,
  /* AFQMC_MIXED_BF16, set before afqmc_upload_system: a mask of the
   * contractions run in three-piece split-BF16 (each operand x = h + m + l,
   * three bf16 pieces, 24 significant bits, six products accumulated in
   * FP32) instead of two pieces: 1 the bias product (G from an FP32 product),
   * 2 the HS potential, 4 the propagator (P W, the Taylor terms, P U).
   * Needs M, K and N*M multiples of 16 for the bits that contract over them
   * (AFQMC_ERR_UNSUPPORTED otherwise).  Default 0.  No counterpart in the
   * reference. */
  AFQMC_OPT_THREE_PIECE = 8,
  /* with bit 4 of AFQMC_OPT_THREE_PIECE: a mask of the propagator's eight
   * products that take three pieces, in launch order (bit 0 P W, bits 1-6
   * the Horner steps U_6 .. U_1, bit 7 P U); the others take two pieces of
   * the same operands.  Default 255 (all).  May be changed at any time.  No
   * counterpart in the reference. */
  AFQMC_OPT_THREE_PIECE_STEPS = 9,
  /* with bit 4 of AFQMC_OPT_THREE_PIECE: a mask of the Horner steps (bits
   * 1-6, as above; bits 0 and 7 are refused) whose products take one piece,
   * the hi bf16 words of the split operands (8 significant bits, one MFMA
   * product), unless AFQMC_OPT_THREE_PIECE_STEPS gives them three.  The
   * innermost steps are scaled by V^k/k!, so their rounding barely reaches
   * the walkers.  Default 0.  May be changed at any time.  No counterpart in
   * the reference. */
  AFQMC_OPT_ONE_PIECE_STEPS = 10,
  /* AFQMC_FP64_CSR with the transposed bias, set before afqmc_upload_system:
   * 1 (default) keeps a dense copy of the half-rotated SpvnT when at least
   * half of its entries are stored (those under the cutoff are zero) and
   * runs the bias product on FP64 MFMA; 0 always runs the CSR product.  No
   * counterpart in the reference (its csrmm is the only form). */
  AFQMC_OPT_CSR_DENSE = 11,
  /* nonzero (AFQMC_MIXED_BF16 at energy levels 1-2, set before the system
   * upload): the local energy's Coulomb term contracts the Cholesky vectors
   * with G in three-piece split-BF16 (24-bit operands, FP32 accumulation
   * over chunks summed in FP64) instead of FP64; G itself stays FP64 for the
   * one-body term.  0 (the default) keeps the FP64 contraction.  No
   * counterpart in the reference (its energy is FP64 throughout,
   * src/AFQMC/energy.hpp:63-102). */
  AFQMC_OPT_COULOMB_THREE_PIECE = 12
} afqmc_option;
int afqmc_set_option(afqmc_ctx* ctx, int option, double value);

/* Replaces the data hand-off of afqmc::Initialize
 * (src/Matrix/initialize_serial.hpp:35-228) followed by
 * base::halfrotate_cholesky (src/AFQMC/rotate.hpp:65-169, called from
 * src/miniapps/miniafqmc_base.cpp:181-185).  Inputs are what Initialize
 * returns: trialwfn_alpha/beta = conj(Psi) [M][N], Propg1 [M][M], haj [2N][M],
 * Spvn CSR [M*M][K] already scaled by sqrt(dt), and optionally Vakbl CSR
 * [2NM][2NM] (NULL: the two-body energy uses the Cholesky form built from
 * Spvn, which equals Vakbl's when Vakbl is built from the same vectors).
 * The half rotation runs on the device; `cutoff` is rotate.hpp's cutoff for
 * the CSR SpvnT used by AFQMC_FP64_CSR (dense modes keep every entry). */
int afqmc_upload_system(afqmc_ctx* ctx, const double* trial_alpha, const double* trial_beta,
                        const double* propg1, const double* haj, const afqmc_csr_view* spvn,
                        const afqmc_csr_view* vakbl, double cutoff, double dt);
/* The half-rotated, transposed Cholesky matrix SpvnT [K][2NM] built at
 * upload (rotate.hpp:65-169 output).  Call with rowptr/col/val NULL to get
 * nnz; AFQMC_FP64_CSR mode only. */
int afqmc_get_spvnT(afqmc_ctx* ctx, long long* nnz, int* rowptr, int* col, double* val);

/* Walker set: W = conj(trialwfn) per spin, W_data[n][1] = 1, rest 0
 * (src/miniapps/miniafqmc_base.cpp:225-236). */
int afqmc_init_walkers(afqmc_ctx* ctx);
/* host <-> device copy of one buffer in the reference layout.  to_device=1
 * uploads host -> device.  Buffers held in another layout or precision on
 * the device are converted. */
int afqmc_copy(afqmc_ctx* ctx, int buffer, int to_device, double* host);

/* ---- phases of one substep (src/miniapps/miniafqmc_base.cpp:249-339) ---- */
/* afqmc_sys::calculate_mixed_density_matrix(W, W_data, Gc, compact=true)
 * (src/AFQMC/afqmc_sys.hpp:95-116): G and W_data[:,2..3] (overlaps). */
int afqmc_mixed_density_matrix(afqmc_ctx* ctx);
/* base::get_vbias(SpvnT, Gc, vbias, transposed=true) (src/AFQMC/vbias.hpp:45-83) */
int afqmc_vbias(afqmc_ctx* ctx);

/* No counterpart in the reference: an allocation-free variant of
 * afqmc_mixed_density_matrix for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It needs no context, allocates nothing and reads nothing back
 * to the host; it runs the same three kernels (overlap product, batched
 * inverse + determinant, Gc product) on `stream` (a hipStream_t, NULL for the
 * null stream) in FP64.  All pointers are DEVICE memory:
 *   trialwfn [2][M][N]  conj(Psi) per spin, as afqmc_upload_system's
 *                       trial_alpha, trial_beta;
 *   W        [nwalk][2][M][N] walkers;
 *   Gc       [2*N*M][nwalk]   output, AFQMC_BUF_G's layout;
 *   ovlp     [nwalk][2]       output, det(W_s^T trialwfn_s) (W_data[w][2+s]).
 * The overlap matrices (nwalk*2*N*N complex doubles) live in a static device
 * array of AFQMC_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * exits; larger requests return AFQMC_ERR_SHAPE (N = 128: 128 walkers per
 * call), N > 128 returns AFQMC_ERR_UNSUPPORTED.  Every call on one GPU
 * shares that array, so calls that overlap in time on one GPU (other
 * streams or threads) race.  An exactly singular overlap matrix gives det 0
 * and zero Gc columns (there is no status: nothing is read back). */
#define AFQMC_NOALLOC_SCRATCH_BYTES (64u << 20)
int afqmc_mixed_density_matrix_noalloc(int nmo, int naea, int nwalk, const double* trialwfn, const double* W,
                                       double* Gc, double* ovlp, void* stream);
/* Replaces afqmc_sys::calculate_mixed_density_matrix(W, W_data, G, false)
 * (src/AFQMC/afqmc_sys.hpp:95-116 with mixed_density_matrix.hpp:93-101):
 * full G [2*M*M][nwalk] = conj(A) T1^-1 W^T per spin, and the overlaps.
 * Requires AFQMC_OPT_TRANSPOSED = 0. */
int afqmc_mixed_density_matrix_full(afqmc_ctx* ctx);
/* Replaces base::get_vbias(Spvn, G, vbias, false) (src/AFQMC/vbias.hpp:65-82):
 * vbias = T(Spvn) G_alpha + T(Spvn) G_beta from the full G. */
int afqmc_vbias_full(afqmc_ctx* ctx);
/* X/hybrid-weight loop (src/miniapps/miniafqmc_base.cpp:281-289): x holds the
 * real N(0,1) draws [K][nwalk] (NULL: use the fields staged for substep
 * `substep` by afqmc_stage_fields).  X = x + i vbias; hybridW. */
int afqmc_fields(afqmc_ctx* ctx, const double* x, long substep);
/* base::get_vHS(Spvn, X, vHS) (src/AFQMC/vHS.hpp:43-56) */
int afqmc_vhs(afqmc_ctx* ctx);
/* afqmc_sys::propagate(W, Propg1, vHS) with base::apply_expM order 6
 * (src/AFQMC/afqmc_sys.hpp:146-173, src/AFQMC/vHS.hpp:61-91) */
int afqmc_propagate(afqmc_ctx* ctx);
/* W_data[:,5] = W_data[:,4]; W_data[:,6..7] = W_data[:,2..3]
 * (src/miniapps/miniafqmc_base.cpp:303-309) */
int afqmc_save_overlaps(afqmc_ctx* ctx);
/* afqmc_sys::calculate_overlaps (src/AFQMC/afqmc_sys.hpp:135-144) */
int afqmc_calculate_overlaps(afqmc_ctx* ctx);
/* weight / hybrid-energy / Eshift update (src/miniapps/miniafqmc_base.cpp:317-327) */
int afqmc_weights(afqmc_ctx* ctx, double dt, double* eshift_inout);
/* Population control by the comb method, for the place after a step's local
 * energy where the reference's driver marks that branching would happen
 * (src/miniapps/miniafqmc_base.cpp:346; the reference implements none).
 * With w_i = max(0, Re W_data[i][1]) and W = sum_i w_i, the nwalk teeth
 * t_j = (u + j) * W / nwalk, j = 0..nwalk-1, pick the walker whose
 * cumulative-weight interval holds them, so walker i is copied floor or ceil
 * of nwalk * w_i / W times, copies in index order; walkers with zero weight
 * are dropped.  Walkers and their W_data rows move together and every new
 * weight is W / nwalk (the population and the total weight are kept).  u is
 * the caller's uniform draw in [0, 1); parent_out (may be NULL) receives the
 * parent index of each new walker [nwalk].  AFQMC_ERR_NONFINITE when W is
 * not positive or a weight is not finite (walkers untouched). */
int afqmc_branch(afqmc_ctx* ctx, double u, int* parent_out);
/* afqmc_sys::orthogonalize (src/AFQMC/afqmc_sys.hpp:175-207): Householder LQ
 * with LAPACK's conventions (zgelqf + zunglq) on each W[w][s]. */
int afqmc_orthogonalize(afqmc_ctx* ctx);
/* afqmc_sys::calculate_energy (src/AFQMC/afqmc_sys.hpp:118-133,
 * src/AFQMC/energy.hpp:63-102) on the current G: W_data[:,0] = E_L and the
 * weighted average.  Call afqmc_mixed_density_matrix first, as the driver
 * does (miniafqmc_base.cpp:342-343). */
int afqmc_energy(afqmc_ctx* ctx, double* eav_out);

/* ---- numerics seam ----------------------------------------------------------- */
/* Complex GEMM on the GPU with ma::product's row-major meaning
 * (src/Numerics/ma_operations.hpp:94-107 -> BLAS::gemm,
 * src/Numerics/OhmmsBlas.h:270-278): C[m][n] = alpha*op(A)*op(B) + beta*C,
 * op(A) in {'N','T'}, op(B) in {'N','T','C' (B^H),'R' (conj(B))}, host arrays.  precision: 0 FP64, 1 FP32, 2 split-BF16
 * (the arithmetic of the modes above); bits 8..15 of `precision` may force
 * one tile candidate (1-based) instead of the size heuristic. */
int afqmc_zgemm(int precision, char opA, char opB, int m, int n, int k, const double* alpha, const double* A,
                int lda, const double* B, int ldb, const double* beta, double* C, int ldc);

/* ---- whole-loop entry points -------------------------------------------- */
/* Device copy of the real N(0,1) draws of `nsub` substeps, x_all
 * [nsub][K][nwalk]; afqmc_fields(ctx, NULL, s) and afqmc_run use them. */
int afqmc_stage_fields(afqmc_ctx* ctx, const double* x_all, long nsub);
/* Draw the auxiliary fields on the device, inside each substep, as the
 * reference draws them inside its substep loop (Random.generate_normal,
 * src/miniapps/miniafqmc_base.cpp:282, Box-Muller of
 * src/Utilities/RandomGenerator.h:50-62).  enable != 0: afqmc_fields(ctx,
 * NULL, s), afqmc_substep, afqmc_run and afqmc_tune generate the draws of
 * substep s for walker w from (seed, s, w), bit for bit the numbers
 * afqmc_synth_fields(seed, K, nwalk, 0, nwalk, s, x) returns on the host (so
 * the same run can be replayed on the CPU), and need no staged block.
 * enable == 0: the staged fields are used again.  The stream itself differs
 * from the reference's Mersenne-Twister stream (a counter-based generator, so
 * a captured substep can draw in parallel). */
int afqmc_device_fields(afqmc_ctx* ctx, uint64_t seed, int enable);
/* One substep (miniafqmc_base.cpp:251-338): every phase above, with the
 * orthogonalisation and overlap refresh when do_ortho != 0.  Draws come from
 * the staged fields of `substep`.  Eshift is kept on the device; read it with
 * afqmc_get_eshift. */
int afqmc_substep(afqmc_ctx* ctx, long substep, int do_ortho, double dt);
int afqmc_get_eshift(afqmc_ctx* ctx, double* eshift);
int afqmc_set_eshift(afqmc_ctx* ctx, double eshift);
/* The driver's step loop (miniafqmc_base.cpp:248-350) from substep index
 * `first_substep` of the staged fields: nsteps blocks of nsubsteps substeps,
 * orthogonalisation when the running substep counter (starting at
 * first_substep) is a positive multiple of northo, and energy at the end of
 * each block into eav_out[nsteps] (may be NULL to skip the energy). */
int afqmc_run(afqmc_ctx* ctx, long first_substep, int nsteps, int nsubsteps, int northo, double dt,
              double* eav_out);
/* Tile autotuning: runs one substep (fields of staged substep 0) and one
 * energy evaluation, timing every tile candidate of each GEMM call site and
 * keeping the fastest for this context.  The walker set, W_data and Eshift
 * are undefined afterwards: call afqmc_init_walkers (or afqmc_copy) next. */
int afqmc_tune(afqmc_ctx* ctx, double dt);
/* Capture one substep (with and without orthogonalisation) into HIP graphs
 * and replay them from then on in afqmc_substep/afqmc_run (enable != 0). */
int afqmc_use_graphs(afqmc_ctx* ctx, int enable);

/* Several independent systems (contexts of any sizes and modes on the same
 * GPU) advanced together.  No counterpart in the reference, which runs one
 * system; for batches of uneven systems.  afqmc_group_substep captures one
 * substep of every member into a single HIP graph on first use, each member
 * on its own stream branch, and replays it: one graph launch per substep
 * advances every system, their kernels running side by side.  Each member
 * needs its fields staged (afqmc_stage_fields) for the substep; a member's
 * results equal those of afqmc_substep on that member alone.  The graph is
 * captured again when dt changes or a member's staged fields or options
 * change.  Members stay usable on their own (their streams wait for the
 * group's substeps); afqmc_group_sync waits for everything. */
typedef struct afqmc_group afqmc_group;
int afqmc_group_create(afqmc_ctx* const* ctxs, int n, afqmc_group** out);
int afqmc_group_destroy(afqmc_group* group);
int afqmc_group_substep(afqmc_group* group, long substep, int do_ortho, double dt);
int afqmc_group_sync(afqmc_group* group);
const char* afqmc_group_error(const afqmc_group* group);

/* ---- measurement -------------------------------------------------------- */
/* Per-kernel timing with HIP events on the context's stream.  While enabled,
 * every launch is bracketed by events; afqmc_kernel_stats fills up to `cap`
 * entries (name, launches, total ms) and returns the count in *count. */
int afqmc_profile(afqmc_ctx* ctx, int enable);
int afqmc_kernel_stats(afqmc_ctx* ctx, int cap, char (*names)[64], long long* launches, double* total_ms,
                       int* count);

/* ---- synthetic input (the reference's example afqmc.h5 is not shipped) --- */
typedef struct {
  int nmo, naea, nchol;
  double density;   /* kept fraction of Cholesky entries (1 = dense) */
  double sigma_L;   /* std of Re and Im of an unthresholded Cholesky entry */
  double sigma_h;   /* std of the one-body Gaussian before symmetrisation */
  double dt;        /* time step: Propg1 = expm(-dt/2 h), Spvn *= sqrt(dt) */
  uint64_t seed;
  int with_vakbl;   /* also build the compacted two-body matrix (small systems) */
} afqmc_synth_params;

typedef struct {
  int nmo, naea, nchol;
  const double* trial_alpha; /* [M][N] conj(Psi_alpha) */
  const double* trial_beta;  /* [M][N] conj(Psi_beta) */
  const double* haj;         /* [2N][M] */
  const double* propg1;      /* [M][M] */
  const double* h1;          /* [M][M] one-body Hamiltonian */
  afqmc_csr_view spvn;       /* [M*M][K], sqrt(dt)-scaled */
  afqmc_csr_view vakbl;      /* [2NM][2NM] or nnz = 0 */
} afqmc_synth_view;

typedef struct afqmc_synth afqmc_synth;
int afqmc_synth_create(const afqmc_synth_params* p, afqmc_synth** out);
int afqmc_synth_destroy(afqmc_synth* s);
int afqmc_synth_get(const afqmc_synth* s, afqmc_synth_view* v);
/* real N(0,1) draws for walkers [w0, w0+nwalk) of a run with wtot walkers,
 * one substep, written as [K][nwalk] */
int afqmc_synth_fields(uint64_t seed, int nchol, int nwalk, int w0, int wtot, long substep, double* x);

#ifdef __cplusplus
}
#endif
#endif
