// Batched Lanczos state and vector kernels (declarations).
#pragma once

#include "kernels.hpp"

namespace tme {

constexpr int kDotChunk = 2048;  // rows per partial-dot block (8 per thread)
constexpr int kTsChunk = 2048;   // rows per tall-skinny A^H B block
constexpr int kTsMaxRows = 128;  // coefficient rows of k_tsmm (basis vectors per problem)

// Device-side state of a batch of Lanczos runs.
struct LzState {
  double2* V;        // B x (ncvmax+1) x n
  double2* MV;       // same layout
  double2* W;        // B x n work vector
  double2* MW;       // B x n
  double2* h;        // B x kmax coefficients
  double2* part;     // B x kmax x nchunk partial dots
  double* alpha;     // B x ncvmax
  double* beta;      // B x ncvmax
  double* bnew;      // B
  int* jcur;         // B current step
  int* ncv;          // B subspace size
  int* live;         // B still iterating
  int* frozen;       // B breakdown at jcur
  int* act;          // B derived each step
  long long ldb;     // (ncvmax+1) * n
  int ncvmax;
  int nchunk;
};

__global__ void k_update_act(LzState st, int B);
__global__ void k_copy_mv(LzState st, long long n);
__global__ void k_mdot(LzState st, long long n, int mode, int kmax);
__global__ void k_mdot_reduce(LzState st, int mode, int kmax);
__global__ void k_three_term(LzState st, long long n);
__global__ void k_axpy_multi(LzState st, long long n, int kmax);
__global__ void k_spmv_bsr(int nf, const int* rowptr, const int* colidx, const double2* const* blocks,
                           const double2* X, long long ldx, double2* Y, long long ldy, const int* act);
__global__ void k_finish_beta(LzState st, int B, int kmax);
__global__ void k_finish_beta_start(LzState st, int B, int kmax);
__global__ void k_normalize_next(LzState st, long long n);
__global__ void k_advance(LzState st, int B);
__global__ void k_combine(const double2* V, long long ldb, long long n, const double* C, int ldc,
                          long long cstride, const int* nk, const int* nc, double2* out, long long ldo);
__global__ void k_build_phase_blocks(long long nslot, const int* slot_row, const int* colidx, const int* tslot,
                                     const unsigned char* pinned, const double* Xc, const double* Xp,
                                     const double* dXc, const double* dXp, double s, double pre, double pim,
                                     double2* out, int skew, double2 scale);
__global__ void k_spmv_multi(int nf, const int* rowptr, const int* colidx, const double2* const* blocks,
                             const double2* X, long long ldb, long long n, const int* ncol, double2* Y);
// per column c: ||K x - lam M x||^2, ||K x||^2, Re(x^H M x) over n entries (chunk partials), columns cs apart
__global__ void k_residual_partial(const double2* X, const double2* KX, const double2* MX, long long ldb, long long n,
                                   long long cs, const double* lam, int lam_ld, const int* ncol, double* part,
                                   int nchunk);
__global__ void k_expand_modes(int nn, const int* free_of_full, const int* node_class, double cr,
                               double sr, int c1, int c2, int cax, const double2* X, long long n,
                               const int* src, const int* conjq, int nm, double pre, double pim,
                               const double* scale, double2* out);
__global__ void k_real_to_complex(const double* x, double2* y, long long n);
__global__ void k_gram_partial(const double2* Y, const double2* Z, long long ldb, long long n, const int* ncol,
                               int nmax, double2* part, int nchunk);
__global__ void k_combine_c(const double2* V, long long ldb, long long n, const double2* Zc, int ldc,
                            long long cstride, const int* nk, double2* out, long long ldo);
__global__ void k_copy_col(const double2* src, long long lds, long long n, int q, const int* ncol, double2* dst,
                           long long ldd);
__global__ void k_flag_cols(const int* ncol, int q, int B, int* act);

// (A_b^H B_b) partial sums over row chunks of n rows; columns cs apart.  NH: eight-column
// parts of B (kb <= 8 NH); a block's four waves are WA column groups of NAB 16-column
// blocks of A; gridDim.z covers ka in steps of 16 WA NAB columns
template <bool TWO, int NH, int NAB, int WA>
__global__ void k_tsgemm_mfma(const double2* A, long long lda_b, const double2* Bm, const double2* Bm2,
                              long long ldb_b, long long n, long long cs, const int* ka_b, int kb, const int* act,
                              int kamax, double2* part, double2* part2, int nchunk);
__global__ void k_tsgemm_reduce(const double2* part, const int* ka_b, int kb, const int* act, int kamax, int nchunk,
                                double2* C);
__global__ void k_tsmm(const double2* A, long long lda_b, long long n, long long cs, const double2* C, const int* ka_b, int kc,
                       const int* act, int kamax, double alpha, double beta, double2* Y, long long ldy_b,
                       const int* yoff, int r0);
__global__ void k_mixed_residual(const double2* Y, long long ldy, const int* s0v, double2* KW, const double2* MW,
                                 long long ldw, const ProbParam* params, int ncol, long long n, const int* act);
__global__ void k_colnorm_partial(const double2* R, long long ldr, const double2* Y, long long ldy, const int* s0v,
                                  int ncol, long long n, const int* act, double* part, int nchunk);
__global__ void k_add_cols(double2* W, const double2* D, long long ld, int ncol, long long n, const int* act);
__global__ void k_copy_cols(const double2* src, long long lds, const int* s0v, double2* dst, long long ldd,
                            const int* d0v, int ncol, long long n, const int* act);
template <int CPL>
__global__ void k_spmv_pair(int nf, int B, const int* rowptr, const int* colidx, const double2* const* blocks1,
                            const double2* const* blocks2, const double2* X, long long ldx, long long n,
                            const int* ncol, int cmul, int c0, double2* Y1, double2* Y2, long long ldy, int acc1,
                            int acc2);
// Per-problem operands of k_spmv_km: the problem's operator's real slot arrays
// (K base and its omega^2 derivative, M), its real-slot flags, and omega^2
struct SplitKM {
  const double2* pack;  // per slot 14 double2: (Kc, dKc) of its 9 entries, then Mc and a zero
  const unsigned char* real;
  double s;
};
__global__ void k_pack_real_slots(long long nslot, const double* Kc, const double* dKc, const double* Mc,
                                  double2* pack);
__global__ void k_real_slots(long long nslot, const int* slot_row, const int* colidx, const int* tslot,
                             const unsigned char* pinned, const double* Kp, const double* dKp, const double* Mp,
                             unsigned char* real);
template <int CPL>
__global__ void k_spmv_km(int nf, int B, const int* rowptr, const int* colidx, const SplitKM* sk,
                          const double2* const* kblocks, const double2* const* mblocks, const double2* X,
                          long long ldx, long long n, const int* ncol, int c0, double2* Y1, double2* Y2,
                          long long ldy);
__global__ void k_qep_swap_halves(const double2* BX, double2* AX, long long ld, long long n, long long cs,
                                  const int* ncol);
__global__ void k_axpy_half(const double2* src, long long lds, const int* s0v, long long soff, double2* dst,
                            long long ldd, const int* d0v, long long doff, int ncol, long long len, long long stride,
                            double alpha, double beta, const int* act);
__global__ void k_head_dot(const double2* X, const double2* Y, long long ldb, long long stride, long long len,
                           const int* ncol, int ldc, double* part, int nchunk);
__global__ void k_spmv_abs(int nf, const int* rowptr, const int* colidx, const double2* const* blocks1,
                           const double2* const* blocks2, const double2* const* blocks3, const double2* X,
                           long long ldx, long long cs, const int* ncol, double2* Y1, double2* Y2, double2* Y3,
                           long long ldy, const SplitKM* sk);
__global__ void k_floor_partial(const double2* AK, const double2* AD, const double2* AM, long long ldb, long long n,
                                long long cs, const double* lam,
                                int lam_ld, const int* ncol, double* part, int nchunk);
// Host: H x = lambda G x, G Hermitian positive definite (Cholesky reduction + herm_eig);
// false when G is not numerically positive definite.  X[k*m+q] (re, im) = component k of vector q.
bool gen_eig_chol(int m, const std::vector<double>& H, const std::vector<double>& G, std::vector<double>& lam,
                  std::vector<double>& X);

// Host: Hermitian eigenproblem (Householder + QL), theta ascending, S[k*m+q] (re, im)
void herm_eig(int m, const std::vector<double>& A, std::vector<double>& theta, std::vector<double>& S);

// Host: generalised Hermitian eigenproblem A z = mu B z (n <= a few dozen),
// B positive definite; A, B row-major n x n complex (re, im interleaved).
// mu ascending; Z column q (Z[k * n + q]) is B-orthonormal.
void herm_gen_eig(int n, const std::vector<double>& A, const std::vector<double>& B, std::vector<double>& mu,
                  std::vector<double>& Z);

// Host-side symmetric tridiagonal eigensolver (implicit QL, Wilkinson shifts):
// d (n) diagonal, e (n) off-diagonal e[i] = T(i, i+1); on exit d holds the
// eigenvalues in ascending order and z (n x n, column-major) the eigenvectors.
void tridiag_eigen(int n, std::vector<double>& d, std::vector<double> e, std::vector<double>& z);

}  // namespace tme
