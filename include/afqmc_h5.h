/* afqmc_h5.h -- the miniapp's HDF5 input (afqmc.h5) read and written in the
 * reference's format, for hosts that have the HDF5 C library.
 *
 * Built as libafqmc_h5.so (host code only, linked against libhdf5) next to
 * libafqmc_hip.so when the HDF5 headers are found at build time; the GPU
 * library does not depend on it.  The reader hands over exactly what
 * afqmc::Initialize (src/Matrix/initialize_serial.hpp:35-228) hands the
 * driver, in the layouts afqmc_upload_system takes.
 *
 * File layout (complex values are double datasets with a trailing
 * dimension of 2, indices are int unless noted):
 *   /Wavefunctions/PureSingleDeterminant/
 *     dims             int[8]  = {0, nnz(Vakbl), 2M^2, 2M^2, M, NAEA, NAEB, 0}
 *     hij_indx, hij    int[n], complex[n]: one-body entries at i*M + j,
 *                      i in [0, NAEA) (alpha) or [M, M + NAEB) (beta)
 *     Wavefun          complex[2*M*N]: Psi in [spin][i][a] order
 *     SpHijkl_vals, SpHijkl_cols, SpHijkl_rowIndex
 *                      CSR over (i*M + j) indices, 2M^2 x 2M^2
 *   /Propagators/phaseless_ImpSamp_ForceBias/
 *     Spvn_dims        long[5] = {nnz, M^2, K, M, nblk}
 *     Spvn_propg1      complex[M^2]
 *     Spvn_block_sizes int[nblk]
 *     Spvn_index_<b>   int[2*count]: (row = i*M + k, col = n) pairs
 *     Spvn_vals_<b>    complex[count]: Cholesky vectors without sqrt(dt)
 */
#ifndef AFQMC_H5_H
#define AFQMC_H5_H

#include "afqmc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct afqmc_h5_system afqmc_h5_system;

/* Replaces afqmc::Initialize(hdf_archive&, dt, afqmc_sys&, Propg1, Spvn,
 * haj, Vakbl) (src/Matrix/initialize_serial.hpp:35-228).  Reads `path` and
 * builds: trial_alpha/trial_beta = conj(Wavefun) per spin ([M][N], :113-130),
 * haj [2N][M] with the one-body rows compacted (a = i for alpha rows,
 * i - M + NAEA for beta rows, :91-108), Vakbl as CSR [2NM][2NM] with rows
 * and columns compacted the same way and sorted by (row, column)
 * (:134-164), Propg1 [M][M] (:186-194) and Spvn as CSR [M^2][K] sorted by
 * (row, column) and scaled by sqrt(dt) (:196-221).  Where the reference
 * only asserts (rows outside the occupied ones, indices out of range,
 * inconsistent sizes) the reader fails with AFQMC_ERR_SHAPE; a missing
 * group or dataset, or a file that is not HDF5, gives AFQMC_ERR_ARG.
 * The view's h1 is NULL (the file holds only the half-rotated one-body
 * terms) and vakbl.nnz is 0 when the file holds no two-body entries. */
int afqmc_h5_read(const char* path, double dt, afqmc_h5_system** out);
int afqmc_h5_get(const afqmc_h5_system* s, afqmc_synth_view* v);
int afqmc_h5_destroy(afqmc_h5_system* s);

/* Writes `v` in the format above (the inverse of afqmc_h5_read): Wavefun =
 * conj(trial), the compacted one-body rows and Vakbl indices expanded back
 * to spin-orbital indices, every haj entry written.  v->spvn must hold the
 * Cholesky vectors WITHOUT the sqrt(dt) factor, as the file does.
 * spvn_block: Cholesky entries per Spvn_index_<b>/Spvn_vals_<b> block
 * (0: one block).  The reference has no writer (its afqmc.h5 comes from
 * QMCPACK's converters); this one makes files that its reader accepts. */
int afqmc_h5_write(const char* path, const afqmc_synth_view* v, long spvn_block);

/* Message of the calling thread's last failure ("" if none). */
const char* afqmc_h5_error(void);

#ifdef __cplusplus
}
#endif
#endif
