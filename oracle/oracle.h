/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of the miniAFQMC walker step, used as the checker for
 * the HIP implementation.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never links it.
 *
 * Complex arrays are interleaved doubles (re, im), row-major, in the
 * reference's in-memory layouts:
 *   trial (trialwfn = conj(Psi))   [M][N]         per spin
 *   W     walkers                  [nw][2][M][N]
 *   Wd    walker data              [nw][8]   (0 eloc, 1 weight, 2/3 ovlp a/b,
 *                                             4 E_hyb, 5 old E_hyb, 6/7 old ovlp)
 *   Gc    compact Green function   [2*N*M][nw]
 *   vbias, X                       [K][nw]
 *   vHS                            [M*M][nw]
 *   Spvn  CSR [M*M][K] (sqrt(dt) scaled), SpvnT CSR [K][2*N*M], Vakbl CSR [2NM][2NM]
 *
 * Parity status of every function is listed in oracle.cpp's header.
 */
#ifndef AFQMC_ORACLE_H
#define AFQMC_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* --- numerics seam ------------------------------------------------------- */
int orc_csrmm(char transa, int M, int N, int K, const double* alpha,
              const double* val, const int* indx, const int* pntrb, const int* pntre,
              const double* B, int ldb, const double* beta, double* C, int ldc);
/* row-major C[m][n] = alpha*op(A)*op(B) + beta*C, op in {'N','T','H'} */
void orc_product(char opA, char opB, int m, int n, int k, const double* alpha,
                 const double* A, int lda, const double* B, int ldb,
                 const double* beta, double* C, int ldc);
/* LAPACK-convention LU / inverse / det on the column-major view of a
 * row-major array (ma_lapack.hpp); piv is 1-based. */
int orc_getrf(int rows, int cols, double* A, int lda, int* piv);
int orc_getri(int n, double* A, int lda, const int* piv);
int orc_invert(int n, double* A, int* piv, double* det_out);
int orc_determinant(int n, double* A, int* piv, double* det_out);
int orc_gelqf(int rows, int cols, double* A, int lda, double* tau);
int orc_glq(int rows, int cols, double* A, int lda, const double* tau);

/* --- AFQMC kernels -------------------------------------------------------- */
int orc_mixed_density_matrix(int M, int N, const double* conjA, const double* B,
                             double* C, int compact, double* ovlp_out);
int orc_overlap(int M, int N, const double* conjA, const double* B, double* ovlp_out);
int orc_calculate_mixed_density_matrix(int M, int N, int nw, const double* trial_a,
                                       const double* trial_b, const double* W,
                                       double* Wd, double* G, int compact);
int orc_calculate_overlaps(int M, int N, int nw, const double* trial_a,
                           const double* trial_b, const double* W, double* Wd);
int orc_get_vbias(int rows, int cols, const int* rowptr, const int* col, const double* val,
                  int grows, int nw, const double* G, double* v, int transposed);
int orc_get_vhs(int rows, int cols, const int* rowptr, const int* col, const double* val,
                int nw, const double* X, double* v);
int orc_apply_expM(int M, int N, const double* V, double* S, int order);
int orc_propagate(int M, int N, int nw, const double* Propg, const double* vHS, double* W);
int orc_orthogonalize(int M, int N, int nw, double* W);
int orc_calculate_energy(int M, int N, int nw, double* Wd, const double* Gc,
                         const double* haj, int vrows, const int* vrowptr,
                         const int* vcol, const double* vval, double* eav_out);
int orc_energy_cholesky(int M, int N, int K, int nw, double* Wd, const double* Gc,
                        const double* haj, const double* Lt, double* eav_out);
int orc_fields(int K, int nw, const double* vbias, const double* x, double* X, double* hybridW);
int orc_weights(int nw, double dt, const double* hybridW, double* Wd, double* eshift_inout);
/* comb population control (the driver marks the place, miniafqmc_base.cpp:346) */
int orc_branch(int M, int N, int nw, double u, double* W, double* Wd, int* parent);

/* half rotation: two-pass count + fill (rotate.hpp:65-169).  Call with
 * out arrays NULL to get nnz; then with arrays of that size. */
long orc_halfrotate_cholesky(int M, int N, int K, const double* trial_a, const double* trial_b,
                             const int* rowptr, const int* col, const double* val,
                             double cutoff, int* out_rowptr, int* out_col, double* out_val);
/* the same in one computation: begin returns a handle and nnz, take fills
 * arrays of that size (or only releases the handle when out_rowptr is NULL) */
void* orc_halfrotate_begin(int M, int N, int K, const double* trial_a, const double* trial_b,
                           const int* rowptr, const int* col, const double* val, double cutoff,
                           long* nnz_out);
void orc_halfrotate_take(void* handle, int* out_rowptr, int* out_col, double* out_val);
/* COO -> CSR with (row, col) ordering (SparseMatrix.hpp:295-319) */
int orc_compress(int nrows, long nnz, int* rows, int* cols, double* vals, int* rowptr);

/* orc_run status for a non-finite auxiliary-field draw (the library's
 * AFQMC_ERR_NONFINITE); a positive status is LAPACK's info of the first
 * singular overlap matrix (the library's AFQMC_ERR_SINGULAR). */
#define ORC_NONFINITE (-6)
/* OpenMP threads of the walker-parallel loops (n <= 0: all processors);
 * results do not depend on the count. */
void orc_set_threads(int n);
int orc_get_threads(void);

/* full driver loop (miniafqmc_base.cpp:213-349), aux fields from x_all
 * [nsteps*nsubsteps][K][nw]; spt_rp == NULL selects the non-transposed
 * branch (full G and T(Spvn) for the bias, miniafqmc_base.cpp:267-276).
 * energy_mode 0: Vakbl CSR, 1: Cholesky form from
 * Lt (dense [K][2NM], unscaled), 2: no energy (timing samples).  eav_out has nsteps+1 entries (initial first). */
int orc_run(int M, int N, int K, int nw, int nsteps, int nsubsteps, int northo, double dt,
            const double* trial_a, const double* trial_b, const double* Propg, const double* haj,
            const int* sp_rowptr, const int* sp_col, const double* sp_val,
            const int* spt_rowptr, const int* spt_col, const double* spt_val,
            int energy_mode, const int* v_rowptr, const int* v_col, const double* v_val,
            const double* Lt, const double* x_all, double* W, double* Wd, double* eav_out,
            double* eshift_out, double* phase_seconds);

#ifdef __cplusplus
}
#endif
#endif
