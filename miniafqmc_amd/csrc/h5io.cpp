// HDF5 input/output in the reference's afqmc.h5 format
// (src/Matrix/initialize_serial.hpp:35-228, read side; io/hdf_stl.h:25-49 for
// how std::vector<T> datasets are shaped: one dimension, plus a trailing 2
// for complex values, io/hdf_dataspace.h:57-78).  Host code only.
#include "afqmc_h5.h"

#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <numeric>
#include <string>
#include <vector>

using cplx = std::complex<double>;

struct afqmc_h5_system {
  int M = 0, N = 0, K = 0;
  std::vector<cplx> ta, tb, haj, propg1;
  std::vector<int> sp_rp, sp_col, v_rp, v_col;
  std::vector<cplx> sp_val, v_val;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
  g_err = msg;
  return code;
}

// Closes an HDF5 handle when it leaves scope.
struct Hid {
  hid_t id;
  herr_t (*close)(hid_t);
  Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~Hid()
  {
    if (id >= 0) close(id);
  }
  Hid(const Hid&) = delete;
  Hid& operator=(const Hid&) = delete;
};

// Dataset shape: the extent of each dimension.
bool shape_of(hid_t grp, const char* name, std::vector<hsize_t>& dims)
{
  if (H5Lexists(grp, name, H5P_DEFAULT) <= 0) return false;
  Hid d(H5Dopen2(grp, name, H5P_DEFAULT), H5Dclose);
  if (d.id < 0) return false;
  Hid sp(H5Dget_space(d.id), H5Sclose);
  const int r = H5Sget_simple_extent_ndims(sp.id);
  if (r < 0) return false;
  dims.assign(r, 0);
  if (r > 0) H5Sget_simple_extent_dims(sp.id, dims.data(), nullptr);
  return true;
}

// Reads a one-dimensional dataset of n elements of `mtype`, n from the file
// (std::vector reads resize to the dataset, hdf_stl.h:38-43).
template <class T> int read_vec(hid_t grp, const char* name, hid_t mtype, std::vector<T>& out)
{
  std::vector<hsize_t> dims;
  if (!shape_of(grp, name, dims)) return fail(AFQMC_ERR_ARG, std::string("missing dataset ") + name);
  if (dims.size() != 1) return fail(AFQMC_ERR_SHAPE, std::string(name) + ": expected a one-dimensional dataset");
  out.assign(dims[0], T());
  if (dims[0] == 0) return AFQMC_OK;
  Hid d(H5Dopen2(grp, name, H5P_DEFAULT), H5Dclose);
  if (H5Dread(d.id, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) < 0)
    return fail(AFQMC_ERR_ARG, std::string("cannot read ") + name);
  return AFQMC_OK;
}

// Complex vector: double dataset [n][2].
int read_cvec(hid_t grp, const char* name, std::vector<cplx>& out)
{
  std::vector<hsize_t> dims;
  if (!shape_of(grp, name, dims)) return fail(AFQMC_ERR_ARG, std::string("missing dataset ") + name);
  if (dims.size() != 2 || dims[1] != 2)
    return fail(AFQMC_ERR_SHAPE, std::string(name) + ": expected a complex dataset [n][2]");
  out.assign(dims[0], cplx());
  if (dims[0] == 0) return AFQMC_OK;
  Hid d(H5Dopen2(grp, name, H5P_DEFAULT), H5Dclose);
  if (H5Dread(d.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) < 0)
    return fail(AFQMC_ERR_ARG, std::string("cannot read ") + name);
  return AFQMC_OK;
}

int write_ds(hid_t grp, const char* name, hid_t mtype, int rank, const hsize_t* dims, const void* data)
{
  Hid sp(H5Screate_simple(rank, dims, nullptr), H5Sclose);
  if (sp.id < 0) return fail(AFQMC_ERR_ARG, std::string("cannot make the dataspace of ") + name);
  Hid d(H5Dcreate2(grp, name, mtype, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  if (d.id < 0) return fail(AFQMC_ERR_ARG, std::string("cannot create dataset ") + name);
  bool empty = false;
  for (int i = 0; i < rank; i++) empty |= dims[i] == 0;
  if (!empty && H5Dwrite(d.id, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    return fail(AFQMC_ERR_ARG, std::string("cannot write ") + name);
  return AFQMC_OK;
}

template <class T> int write_vec(hid_t grp, const char* name, hid_t mtype, const std::vector<T>& v)
{
  const hsize_t dims[1] = {(hsize_t)v.size()};
  return write_ds(grp, name, mtype, 1, dims, v.data());
}

int write_cvec(hid_t grp, const char* name, const cplx* v, size_t n)
{
  const hsize_t dims[2] = {(hsize_t)n, 2};
  return write_ds(grp, name, H5T_NATIVE_DOUBLE, 2, dims, v);
}

#define RET(x)                 \
  do {                         \
    int r_ = (x);              \
    if (r_ != AFQMC_OK) return r_; \
  } while (0)

// Spin-orbital row i in [0, 2M) -> compacted occupied row a in [0, 2N), or -1
// (initialize_serial.hpp:97-104): alpha rows i < N, beta rows M <= i < M + N.
inline int compact_row(int i, int M, int N)
{
  if (i >= 0 && i < N) return i;
  if (i >= M && i < M + N) return i - M + N;
  return -1;
}
inline int expand_row(int a, int M, int N) { return a < N ? a : a - N + M; }

// CSR from (row, col, value) triplets, sorted by (row, col) with ties kept in
// input order (SparseMatrix::compress, SparseMatrix.hpp:295-319).
void to_csr(int nrows, std::vector<int>& rows, std::vector<int>& cols, std::vector<cplx>& vals, std::vector<int>& rp,
            std::vector<int>& ci, std::vector<cplx>& cv)
{
  const size_t n = vals.size();
  std::vector<size_t> ord(n);
  std::iota(ord.begin(), ord.end(), size_t(0));
  std::stable_sort(ord.begin(), ord.end(), [&](size_t x, size_t y) {
    return rows[x] < rows[y] || (rows[x] == rows[y] && cols[x] < cols[y]);
  });
  rp.assign(nrows + 1, 0);
  ci.resize(n);
  cv.resize(n);
  for (size_t t = 0; t < n; t++) {
    ci[t] = cols[ord[t]];
    cv[t] = vals[ord[t]];
    rp[rows[ord[t]] + 1]++;
  }
  for (int r = 0; r < nrows; r++) rp[r + 1] += rp[r];
}

int read_file(hid_t f, double dt, afqmc_h5_system& s)
{
  if (H5Lexists(f, "Wavefunctions", H5P_DEFAULT) <= 0 ||
      H5Lexists(f, "Wavefunctions/PureSingleDeterminant", H5P_DEFAULT) <= 0)
    return fail(AFQMC_ERR_ARG, "H5Group /Wavefunctions/PureSingleDeterminant does not exist");
  if (H5Lexists(f, "Propagators", H5P_DEFAULT) <= 0 ||
      H5Lexists(f, "Propagators/phaseless_ImpSamp_ForceBias", H5P_DEFAULT) <= 0)
    return fail(AFQMC_ERR_ARG, "H5Group /Propagators/phaseless_ImpSamp_ForceBias does not exist");
  {
    Hid g(H5Gopen2(f, "Wavefunctions/PureSingleDeterminant", H5P_DEFAULT), H5Gclose);
    std::vector<int> dims;
    RET(read_vec(g.id, "dims", H5T_NATIVE_INT, dims));
    if (dims.size() != 8) return fail(AFQMC_ERR_SHAPE, "dims must hold 8 entries");
    if (dims[5] != dims[6]) return fail(AFQMC_ERR_SHAPE, "expecting NAEA == NAEB");
    if (dims[7] != 0) return fail(AFQMC_ERR_SHAPE, "found spin-unrestricted integrals (dims[7] != 0)");
    const int M = dims[4], N = dims[5];
    if (M <= 0 || N <= 0 || N > M) return fail(AFQMC_ERR_SHAPE, "invalid NMO / NAEA");
    const long M2 = 2L * M * M;
    if (dims[2] != M2 || dims[3] != M2) return fail(AFQMC_ERR_SHAPE, "incorrect dimensions of the 2-body Hamiltonian");
    s.M = M;
    s.N = N;
    const long NM = (long)N * M;

    // one-body: haj[a][j] = h(i, j)
    std::vector<int> hidx;
    std::vector<cplx> hval;
    RET(read_vec(g.id, "hij_indx", H5T_NATIVE_INT, hidx));
    RET(read_cvec(g.id, "hij", hval));
    if (hidx.size() != hval.size()) return fail(AFQMC_ERR_SHAPE, "hij_indx and hij differ in length");
    s.haj.assign(2 * NM, cplx());
    for (size_t n = 0; n < hidx.size(); n++) {
      const int i = hidx[n] / M, j = hidx[n] % M;
      const int a = hidx[n] >= 0 ? compact_row(i, M, N) : -1;
      if (a < 0) return fail(AFQMC_ERR_SHAPE, "one-body entry outside the occupied rows");
      s.haj[(size_t)a * M + j] = hval[n];
    }

    // trial: conj of Wavefun per spin
    std::vector<cplx> wf;
    RET(read_cvec(g.id, "Wavefun", wf));
    if ((long)wf.size() != 2 * NM) return fail(AFQMC_ERR_SHAPE, "inconsistent dimensions in Wavefun");
    s.ta.resize(NM);
    s.tb.resize(NM);
    for (long t = 0; t < NM; t++) {
      s.ta[t] = std::conj(wf[t]);
      s.tb[t] = std::conj(wf[NM + t]);
    }

    // two-body: CSR over spin-orbital pair indices, compacted to (a*M + j)
    std::vector<cplx> vv;
    std::vector<int> vc, vr;
    RET(read_cvec(g.id, "SpHijkl_vals", vv));
    RET(read_vec(g.id, "SpHijkl_cols", H5T_NATIVE_INT, vc));
    RET(read_vec(g.id, "SpHijkl_rowIndex", H5T_NATIVE_INT, vr));
    if (vv.size() != (size_t)dims[1] || vc.size() != vv.size())
      return fail(AFQMC_ERR_SHAPE, "SpHijkl sizes disagree with dims[1]");
    if (vr.size() != (size_t)M2 + 1 || vr[0] < 0 || vr[M2] != (long)vv.size())
      return fail(AFQMC_ERR_SHAPE, "SpHijkl_rowIndex must hold 2*NMO*NMO+1 offsets ending at nnz");
    std::vector<int> rows(vv.size());
    for (long r = 0; r < M2; r++) {
      if (vr[r + 1] < vr[r]) return fail(AFQMC_ERR_SHAPE, "SpHijkl_rowIndex is not monotone");
      for (int t = vr[r]; t < vr[r + 1]; t++) rows[t] = (int)r;
    }
    auto compact_pair = [&](int x) -> int {
      if (x < 0 || x >= M2) return -1;
      const int a = compact_row(x / M, M, N);
      return a < 0 ? -1 : a * M + x % M;
    };
    for (size_t t = 0; t < vv.size(); t++) {
      rows[t] = compact_pair(rows[t]);
      vc[t] = compact_pair(vc[t]);
      if (rows[t] < 0 || vc[t] < 0) return fail(AFQMC_ERR_SHAPE, "two-body entry outside the occupied rows");
    }
    to_csr((int)(2 * NM), rows, vc, vv, s.v_rp, s.v_col, s.v_val);
  }
  {
    Hid g(H5Gopen2(f, "Propagators/phaseless_ImpSamp_ForceBias", H5P_DEFAULT), H5Gclose);
    const int M = s.M;
    std::vector<long> ld;
    RET(read_vec(g.id, "Spvn_dims", H5T_NATIVE_LONG, ld));
    if (ld.size() != 5) return fail(AFQMC_ERR_SHAPE, "Spvn_dims must hold 5 entries");
    const long ntot = ld[0], nrows = ld[1], K = ld[2], nblk = ld[4];
    if (nrows != (long)M * M || ld[3] != M) return fail(AFQMC_ERR_SHAPE, "Spvn_dims disagree with NMO");
    if (K <= 0 || K > 0x7fffffffL || nblk < 0 || ntot < 0) return fail(AFQMC_ERR_SHAPE, "invalid Spvn_dims");
    s.K = (int)K;
    std::vector<cplx> p1;
    RET(read_cvec(g.id, "Spvn_propg1", p1));
    if ((long)p1.size() != (long)M * M) return fail(AFQMC_ERR_SHAPE, "incorrect dimensions of the 1-body propagator");
    s.propg1 = std::move(p1);
    std::vector<int> counts;
    RET(read_vec(g.id, "Spvn_block_sizes", H5T_NATIVE_INT, counts));
    if ((long)counts.size() != nblk) return fail(AFQMC_ERR_SHAPE, "Spvn_block_sizes disagrees with Spvn_dims[4]");
    std::vector<int> rows, cols, idx;
    std::vector<cplx> vals, bv;
    rows.reserve(ntot);
    cols.reserve(ntot);
    vals.reserve(ntot);
    for (long b = 0; b < nblk; b++) {
      const std::string in = "Spvn_index_" + std::to_string(b), vn = "Spvn_vals_" + std::to_string(b);
      RET(read_vec(g.id, in.c_str(), H5T_NATIVE_INT, idx));
      RET(read_cvec(g.id, vn.c_str(), bv));
      if ((long)bv.size() != counts[b] || (long)idx.size() != 2L * counts[b])
        return fail(AFQMC_ERR_SHAPE, "Spvn block " + std::to_string(b) + " disagrees with Spvn_block_sizes");
      for (long n = 0; n < counts[b]; n++) {
        const int r = idx[2 * n], c = idx[2 * n + 1];
        if (r < 0 || r >= nrows || c < 0 || c >= K) return fail(AFQMC_ERR_SHAPE, "Spvn index out of range");
        rows.push_back(r);
        cols.push_back(c);
        vals.push_back(bv[n]);
      }
    }
    if ((long)vals.size() != ntot) return fail(AFQMC_ERR_SHAPE, "Spvn block sizes do not add up to Spvn_dims[0]");
    to_csr((int)nrows, rows, cols, vals, s.sp_rp, s.sp_col, s.sp_val);
    const double sq = std::sqrt(dt);
    for (auto& v : s.sp_val) v *= sq;
  }
  return AFQMC_OK;
}

} // namespace

extern "C" {

const char* afqmc_h5_error(void) { return g_err.c_str(); }

int afqmc_h5_read(const char* path, double dt, afqmc_h5_system** out)
{
  g_err.clear();
  if (!path || !out || !(dt > 0)) return fail(AFQMC_ERR_ARG, "null argument or dt <= 0");
  *out = nullptr;
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  if (H5Fis_hdf5(path) <= 0) return fail(AFQMC_ERR_ARG, std::string("not an HDF5 file: ") + path);
  Hid f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (f.id < 0) return fail(AFQMC_ERR_ARG, std::string("cannot open ") + path);
  afqmc_h5_system* s = new (std::nothrow) afqmc_h5_system;
  if (!s) return fail(AFQMC_ERR_ALLOC, "out of host memory");
  int r;
  try {
    r = read_file(f.id, dt, *s);
  } catch (const std::bad_alloc&) {
    r = fail(AFQMC_ERR_ALLOC, "out of host memory");
  }
  if (r != AFQMC_OK) {
    delete s;
    return r;
  }
  *out = s;
  return AFQMC_OK;
}

int afqmc_h5_get(const afqmc_h5_system* s, afqmc_synth_view* v)
{
  if (!s || !v) return fail(AFQMC_ERR_ARG, "null argument");
  v->nmo = s->M;
  v->naea = s->N;
  v->nchol = s->K;
  v->trial_alpha = reinterpret_cast<const double*>(s->ta.data());
  v->trial_beta = reinterpret_cast<const double*>(s->tb.data());
  v->haj = reinterpret_cast<const double*>(s->haj.data());
  v->propg1 = reinterpret_cast<const double*>(s->propg1.data());
  v->h1 = nullptr;
  v->spvn = afqmc_csr_view{s->M * s->M, s->K, (long long)s->sp_val.size(), s->sp_rp.data(), s->sp_col.data(),
                           reinterpret_cast<const double*>(s->sp_val.data())};
  const int nak = 2 * s->N * s->M;
  v->vakbl = afqmc_csr_view{nak, nak, (long long)s->v_val.size(), s->v_rp.data(), s->v_col.data(),
                            reinterpret_cast<const double*>(s->v_val.data())};
  return AFQMC_OK;
}

int afqmc_h5_destroy(afqmc_h5_system* s)
{
  delete s;
  return AFQMC_OK;
}

int afqmc_h5_write(const char* path, const afqmc_synth_view* v, long spvn_block)
{
  g_err.clear();
  if (!path || !v || !v->trial_alpha || !v->trial_beta || !v->haj || !v->propg1 || spvn_block < 0)
    return fail(AFQMC_ERR_ARG, "null argument or negative block size");
  const int M = v->nmo, N = v->naea, K = v->nchol;
  if (M <= 0 || N <= 0 || N > M || K <= 0) return fail(AFQMC_ERR_SHAPE, "invalid dimensions");
  const long NM = (long)N * M, M2 = 2L * M * M;
  const afqmc_csr_view& sp = v->spvn;
  if (sp.nrows != M * M || sp.ncols != K || sp.nnz < 0 || !sp.rowptr || (sp.nnz > 0 && (!sp.col || !sp.val)))
    return fail(AFQMC_ERR_SHAPE, "Spvn must be CSR [M*M][K]");
  const afqmc_csr_view& vk = v->vakbl;
  const bool have_v = vk.nnz > 0;
  if (have_v && (vk.nrows != 2 * NM || vk.ncols != 2 * NM || !vk.rowptr || !vk.col || !vk.val))
    return fail(AFQMC_ERR_SHAPE, "Vakbl must be CSR [2NM][2NM]");
  if (M2 >= 0x7fffffffL) return fail(AFQMC_ERR_SHAPE, "2*NMO*NMO exceeds the file's int indices");
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  try {
    Hid f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (f.id < 0) return fail(AFQMC_ERR_ARG, std::string("cannot create ") + path);
    {
      Hid g0(H5Gcreate2(f.id, "Wavefunctions", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
      Hid g(H5Gcreate2(g0.id, "PureSingleDeterminant", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
      if (g.id < 0) return fail(AFQMC_ERR_ARG, "cannot create /Wavefunctions/PureSingleDeterminant");
      const std::vector<int> dims = {0, (int)(have_v ? vk.nnz : 0), (int)M2, (int)M2, M, N, N, 0};
      RET(write_vec(g.id, "dims", H5T_NATIVE_INT, dims));
      std::vector<int> hidx(2 * NM);
      for (int a = 0; a < 2 * N; a++)
        for (int j = 0; j < M; j++) hidx[(size_t)a * M + j] = expand_row(a, M, N) * M + j;
      RET(write_vec(g.id, "hij_indx", H5T_NATIVE_INT, hidx));
      RET(write_cvec(g.id, "hij", reinterpret_cast<const cplx*>(v->haj), 2 * NM));
      std::vector<cplx> wf(2 * NM);
      const cplx* ta = reinterpret_cast<const cplx*>(v->trial_alpha);
      const cplx* tb = reinterpret_cast<const cplx*>(v->trial_beta);
      for (long t = 0; t < NM; t++) {
        wf[t] = std::conj(ta[t]);
        wf[NM + t] = std::conj(tb[t]);
      }
      RET(write_cvec(g.id, "Wavefun", wf.data(), wf.size()));
      // two-body: rows and columns expanded to spin-orbital pair indices;
      // compacted order is preserved, so the expanded rows stay sorted
      std::vector<int> rix(M2 + 1, 0), cols;
      std::vector<cplx> vals;
      if (have_v) {
        const long long base = vk.rowptr[0];
        cols.resize(vk.nnz);
        vals.assign(reinterpret_cast<const cplx*>(vk.val) + base, reinterpret_cast<const cplx*>(vk.val) + base + vk.nnz);
        auto expand_pair = [&](int x) { return expand_row(x / M, M, N) * M + x % M; };
        for (long r = 0; r < 2 * NM; r++) {
          const int R = expand_pair((int)r);
          rix[R + 1] = vk.rowptr[r + 1] - vk.rowptr[r];
          for (int t = vk.rowptr[r]; t < vk.rowptr[r + 1]; t++) {
            if (vk.col[t] < 0 || vk.col[t] >= 2 * NM) return fail(AFQMC_ERR_SHAPE, "Vakbl column out of range");
            cols[t - base] = expand_pair(vk.col[t]);
          }
        }
        for (long R = 0; R < M2; R++) rix[R + 1] += rix[R];
      }
      RET(write_cvec(g.id, "SpHijkl_vals", vals.data(), vals.size()));
      RET(write_vec(g.id, "SpHijkl_cols", H5T_NATIVE_INT, cols));
      RET(write_vec(g.id, "SpHijkl_rowIndex", H5T_NATIVE_INT, rix));
    }
    {
      Hid g0(H5Gcreate2(f.id, "Propagators", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
      Hid g(H5Gcreate2(g0.id, "phaseless_ImpSamp_ForceBias", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
      if (g.id < 0) return fail(AFQMC_ERR_ARG, "cannot create /Propagators/phaseless_ImpSamp_ForceBias");
      const long nnz = sp.nnz;
      const long blk = spvn_block > 0 ? spvn_block : std::max(nnz, 1L);
      const long nblk = nnz > 0 ? (nnz + blk - 1) / blk : 0;
      const std::vector<long> ld = {nnz, (long)M * M, K, M, nblk};
      RET(write_vec(g.id, "Spvn_dims", H5T_NATIVE_LONG, ld));
      RET(write_cvec(g.id, "Spvn_propg1", reinterpret_cast<const cplx*>(v->propg1), (size_t)M * M));
      std::vector<int> counts(nblk);
      for (long b = 0; b < nblk; b++) counts[b] = (int)std::min(blk, nnz - b * blk);
      RET(write_vec(g.id, "Spvn_block_sizes", H5T_NATIVE_INT, counts));
      const long long base = sp.rowptr[0];
      if (sp.rowptr[sp.nrows] - base != nnz) return fail(AFQMC_ERR_SHAPE, "Spvn rowptr disagrees with nnz");
      std::vector<int> rowof(nnz);
      for (int r = 0; r < sp.nrows; r++) {
        if (sp.rowptr[r + 1] < sp.rowptr[r]) return fail(AFQMC_ERR_SHAPE, "Spvn rowptr is not monotone");
        for (long long t = sp.rowptr[r]; t < sp.rowptr[r + 1]; t++) rowof[t - base] = r;
      }
      const int* scol = sp.col + base;
      const cplx* sv = reinterpret_cast<const cplx*>(sp.val) + base;
      for (long b = 0; b < nblk; b++) {
        const long t0 = b * blk, cnt = counts[b];
        std::vector<int> idx(2 * cnt);
        for (long n = 0; n < cnt; n++) {
          idx[2 * n] = rowof[t0 + n];
          idx[2 * n + 1] = scol[t0 + n];
        }
        RET(write_vec(g.id, ("Spvn_index_" + std::to_string(b)).c_str(), H5T_NATIVE_INT, idx));
        RET(write_cvec(g.id, ("Spvn_vals_" + std::to_string(b)).c_str(), sv + t0, cnt));
      }
    }
  } catch (const std::bad_alloc&) {
    return fail(AFQMC_ERR_ALLOC, "out of host memory");
  }
  return AFQMC_OK;
}

} // extern "C"
