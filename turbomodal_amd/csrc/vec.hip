// Batched vector kernels of the shift-invert Lanczos (gfx950).
//
// Every kernel works on a batch of problems (grid.y = problem) whose Lanczos
// step index j_b may differ; per-problem state lives on the device (jcur,
// act, frozen) so a whole step is launched without a host round trip.
// Reductions are two-phase with a fixed chunking and a fixed combine order,
// so a problem's arithmetic does not depend on what else is in the batch.
#include "vec.hpp"

namespace tme {

namespace {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconjmul(double2 a, double2 b) {  // conj(a) b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
// acc += a b and acc += conj(a) b as four FMAs each
__device__ __forceinline__ void cmac(double2& acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}
__device__ __forceinline__ void cmacc(double2& acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(-a.y, b.x, acc.y);
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < (int)(blockDim.x / 64); ++q) t += sh[q];
  return t;  // valid in thread 0
}

}  // namespace

__global__ void k_update_act(LzState st, int B) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  st.act[b] = (st.live[b] && !st.frozen[b] && st.jcur[b] < st.ncv[b]) ? 1 : 0;
}

// W_b = MV_b[:, j_b]
__global__ void k_copy_mv(LzState st, long long n) {
  const int b = blockIdx.y;
  if (!st.act[b]) return;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2* src = st.MV + b * st.ldb + (long long)st.jcur[b] * n;
  st.W[b * n + i] = src[i];
}

// partial dots: part[b][k][chunk] = sum_chunk conj(A_k[i]) x[i]
// mode 0: k in [0, j_b] over MV (reorthogonalisation coefficients)
// mode 1: k = j_b only over MV (alpha)
// mode 2: single dot conj(W) MW (beta^2)
__global__ __launch_bounds__(256) void k_mdot(LzState st, long long n, int mode, int kmax) {
  const int b = blockIdx.y;
  if (!st.act[b]) return;
  const int chunk = blockIdx.x;
  const long long i0 = (long long)chunk * kDotChunk;
  const long long i1 = i0 + kDotChunk < n ? i0 + kDotChunk : n;
  __shared__ double sh[8];
  const double2* x = (mode == 2 ? st.MW : st.W) + b * n;
  const int j = st.jcur[b];
  int k0, k1;
  if (mode == 0) { k0 = 0; k1 = j + 1; }
  else if (mode == 1) { k0 = j; k1 = j + 1; }
  else { k0 = 0; k1 = 1; }
  // cache this thread's x values
  double2 xv[kDotChunk / 256];
  for (int q = 0; q < kDotChunk / 256; ++q) {
    long long i = i0 + threadIdx.x + q * 256;
    xv[q] = i < i1 ? x[i] : make_double2(0.0, 0.0);
  }
  for (int k = k0; k < k1; ++k) {
    const double2* a = mode == 2 ? st.W + b * n : st.MV + b * st.ldb + (long long)k * n;
    double re = 0.0, im = 0.0;
    for (int q = 0; q < kDotChunk / 256; ++q) {
      long long i = i0 + threadIdx.x + q * 256;
      if (i < i1) {
        double2 u = cconjmul(a[i], xv[q]);
        re += u.x;
        im += u.y;
      }
    }
    double tre = block_sum(re, sh);
    double tim = block_sum(im, sh);
    if (threadIdx.x == 0) {
      double2* p = st.part + ((long long)b * kmax + k) * st.nchunk + chunk;
      *p = make_double2(tre, tim);
    }
  }
}

// h[b][k] = sum over chunks (fixed order)
__global__ void k_mdot_reduce(LzState st, int mode, int kmax) {
  const int b = blockIdx.x;
  if (!st.act[b]) return;
  const int j = st.jcur[b];
  int k0, k1;
  if (mode == 0) { k0 = 0; k1 = j + 1; }
  else if (mode == 1) { k0 = j; k1 = j + 1; }
  else { k0 = 0; k1 = 1; }
  for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
    const double2* p = st.part + ((long long)b * kmax + k) * st.nchunk;
    double re = 0.0, im = 0.0;
    for (int c = 0; c < st.nchunk; ++c) {
      re += p[c].x;
      im += p[c].y;
    }
    st.h[(long long)b * kmax + k] = make_double2(re, im);
    if (mode == 1) st.alpha[(long long)b * st.ncvmax + j] = re;
  }
}

// W -= alpha_j V_j + beta_{j-1} V_{j-1}
__global__ void k_three_term(LzState st, long long n) {
  const int b = blockIdx.y;
  if (!st.act[b]) return;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = st.jcur[b];
  const double a = st.alpha[(long long)b * st.ncvmax + j];
  const double2* V = st.V + b * st.ldb;
  double2 w = st.W[b * n + i];
  double2 vj = V[(long long)j * n + i];
  w.x -= a * vj.x;
  w.y -= a * vj.y;
  if (j > 0) {
    const double bt = st.beta[(long long)b * st.ncvmax + j - 1];
    double2 vm = V[(long long)(j - 1) * n + i];
    w.x -= bt * vm.x;
    w.y -= bt * vm.y;
  }
  st.W[b * n + i] = w;
}

// W -= sum_{k<=j} h_k V_k
__global__ void k_axpy_multi(LzState st, long long n, int kmax) {
  const int b = blockIdx.y;
  if (!st.act[b]) return;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = st.jcur[b];
  const double2* V = st.V + b * st.ldb;
  const double2* h = st.h + (long long)b * kmax;
  double2 acc = make_double2(0.0, 0.0);
  for (int k = 0; k <= j; ++k) acc = cadd(acc, cmul(h[k], V[(long long)k * n + i]));
  double2 w = st.W[b * n + i];
  st.W[b * n + i] = make_double2(w.x - acc.x, w.y - acc.y);
}

// y_b = M(k_b) x_b, complex node-block CSR, one thread per DOF row
__global__ void k_spmv_bsr(int nf, const int* __restrict__ rowptr, const int* __restrict__ colidx,
                           const double2* const* __restrict__ blocks, const double2* __restrict__ X,
                           long long ldx, double2* __restrict__ Y, long long ldy,
                           const int* __restrict__ act) {
  const int b = blockIdx.y;
  if (act && !act[b]) return;
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= 3 * nf) return;
  const int I = row / 3, c = row % 3;
  const double2* Mb = blocks[b];
  const double2* x = X + b * ldx;
  double2 acc = make_double2(0.0, 0.0);
  for (int s = rowptr[I]; s < rowptr[I + 1]; ++s) {
    const int J = colidx[s];
    const double2* blk = Mb + (long long)s * 9 + 3 * c;
    acc = cadd(acc, cmul(blk[0], x[3LL * J]));
    acc = cadd(acc, cmul(blk[1], x[3LL * J + 1]));
    acc = cadd(acc, cmul(blk[2], x[3LL * J + 2]));
  }
  Y[b * ldy + row] = acc;
}

// beta_j = sqrt|Re(W^H MW)|.  Breakdown (an invariant Krylov subspace) is
// declared when beta_j <= 1e-12 max_k |alpha_k|: the reference's absolute
// 1e-14 (src/hermitian_lanczos.cpp:350) depends on the units of K and M and
// fires on healthy SI-unit problems once theta drops below 1e-14.  A broken
// down problem is frozen; the host restarts it with a fresh vector.
__global__ void k_finish_beta(LzState st, int B, int kmax) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B || !st.act[b]) return;
  const int j = st.jcur[b];
  const double bb = sqrt(fabs(st.h[(long long)b * kmax].x));
  double amax = 0.0;
  for (int k = 0; k <= j; ++k) amax = fmax(amax, fabs(st.alpha[(long long)b * st.ncvmax + k]));
  st.bnew[b] = bb;
  if (!(bb > 1e-12 * amax)) {
    st.frozen[b] = 1;
    st.act[b] = 0;
  } else {
    st.beta[(long long)b * st.ncvmax + j] = bb;
  }
}

// bnew = ||W||_M only (start vector and restart continuation vector)
__global__ void k_finish_beta_start(LzState st, int B, int kmax) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B || !st.act[b]) return;
  st.bnew[b] = sqrt(fabs(st.h[(long long)b * kmax].x));
}

// V_{j+1} = W / beta, MV_{j+1} = MW / beta
__global__ void k_normalize_next(LzState st, long long n) {
  const int b = blockIdx.y;
  if (!st.act[b]) return;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = st.jcur[b];
  const double inv = 1.0 / st.bnew[b];
  double2 w = st.W[b * n + i], mw = st.MW[b * n + i];
  st.V[b * st.ldb + (long long)(j + 1) * n + i] = make_double2(w.x * inv, w.y * inv);
  st.MV[b * st.ldb + (long long)(j + 1) * n + i] = make_double2(mw.x * inv, mw.y * inv);
}

__global__ void k_advance(LzState st, int B) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B || !st.act[b]) return;
  st.jcur[b] += 1;
}

// out_b[:, c] = sum_{k < nk_b} V_b[:, k] C_b[k][c]  for c < nc_b (C row-major, ldc)
__global__ void k_combine(const double2* __restrict__ V, long long ldb, long long n,
                          const double* __restrict__ C, int ldc, long long cstride,
                          const int* __restrict__ nk, const int* __restrict__ nc,
                          double2* __restrict__ out, long long ldo) {
  const int b = blockIdx.y;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int K = nk[b], NC = nc[b];
  const double2* v = V + b * ldb;
  const double* cb = C + b * cstride;
  for (int c = 0; c < NC; ++c) {
    double re = 0.0, im = 0.0;
    for (int k = 0; k < K; ++k) {
      const double q = cb[(long long)k * ldc + c];
      const double2 x = v[(long long)k * n + i];
      re += q * x.x;
      im += q * x.y;
    }
    out[b * ldo + (long long)c * n + i] = make_double2(re, im);
  }
}

// M(k) blocks: M_c + P M_p + conj(P) M_p^T(t), pinned rows/cols zero
__global__ void k_build_phase_blocks(long long nslot, const int* __restrict__ slot_row,
                                     const int* __restrict__ colidx, const int* __restrict__ tslot,
                                     const unsigned char* __restrict__ pinned,
                                     const double* __restrict__ Xc, const double* __restrict__ Xp,
                                     const double* __restrict__ dXc, const double* __restrict__ dXp,
                                     double s, double pre, double pim, double2* __restrict__ out,
                                     int skew, double2 scale) {
  // one thread per entry of a 3x3 slot block (coalesced reads and writes)
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= 9 * nslot) return;
  const long long sl = idx / 9;
  const int e = (int)(idx - 9 * sl), ci = e / 3, cj = e % 3;
  const int I = slot_row[sl], J = colidx[sl];
  const long long tt = tslot[sl];
  const long long e1 = idx, e2 = tt * 9 + 3 * cj + ci;
  double c = Xc[e1], p = Xp[e1], q = Xp[e2];
  if (dXc) {
    c = fma(s, dXc[e1], c);
    p = fma(s, dXp[e1], p);
    q = fma(s, dXp[e2], q);
  }
  // symmetric X: Xc + P Xp + conj(P) Xp^T; skew X: Xc + P Xp - conj(P) Xp^T
  double2 v = skew ? make_double2(c + pre * (p - q), pim * (p + q)) : make_double2(c + pre * (p + q), pim * (p - q));
  v = cmul(scale, v);
  if (pinned[3 * I + ci] || pinned[3 * J + cj]) v = make_double2(0.0, 0.0);
  out[idx] = v;
}

// multi-column SpMV with one complex block array: Y[:, c] = A X[:, c], c < ncol_b
__global__ void k_spmv_multi(int nf, const int* __restrict__ rowptr, const int* __restrict__ colidx,
                             const double2* const* __restrict__ blocks, const double2* __restrict__ X,
                             long long ldb, long long n, const int* __restrict__ ncol,
                             double2* __restrict__ Y) {
  const int b = blockIdx.y;
  const int c = blockIdx.z;
  if (c >= ncol[b]) return;
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= 3 * nf) return;
  const int I = row / 3, cc = row % 3;
  const double2* Mb = blocks[b];
  const double2* x = X + b * ldb + (long long)c * n;
  double2 acc = make_double2(0.0, 0.0);
  for (int s = rowptr[I]; s < rowptr[I + 1]; ++s) {
    const int J = colidx[s];
    const double2* blk = Mb + (long long)s * 9 + 3 * cc;
    acc = cadd(acc, cmul(blk[0], x[3LL * J]));
    acc = cadd(acc, cmul(blk[1], x[3LL * J + 1]));
    acc = cadd(acc, cmul(blk[2], x[3LL * J + 2]));
  }
  Y[b * ldb + (long long)c * n + row] = acc;
}

// per column: num = ||K x - lam M x||^2, den = ||K x||^2, xmx = Re(x^H M x)
// (chunk partials, combined on the host in fixed order)
__global__ __launch_bounds__(256) void k_residual_partial(const double2* __restrict__ X,
                                                          const double2* __restrict__ KX,
                                                          const double2* __restrict__ MX,
                                                          long long ldb, long long n, long long cs,
                                                          const double* __restrict__ lam, int lam_ld,
                                                          const int* __restrict__ ncol,
                                                          double* __restrict__ part, int nchunk) {
  const int b = blockIdx.y, c = blockIdx.z;
  if (c >= ncol[b]) return;
  __shared__ double sh[8];
  const long long i0 = (long long)blockIdx.x * kDotChunk;
  const double l = lam[(long long)b * lam_ld + c];
  double num = 0.0, den = 0.0, xm = 0.0;
  for (long long i = i0 + threadIdx.x; i < i0 + kDotChunk && i < n; i += 256) {
    const long long o = b * ldb + (long long)c * cs + i;
    double2 k = KX[o], m = MX[o], x = X[o];
    double rx = k.x - l * m.x, ry = k.y - l * m.y;
    num += rx * rx + ry * ry;
    den += k.x * k.x + k.y * k.y;
    xm += x.x * m.x + x.y * m.y;
  }
  double tn = block_sum(num, sh);
  double td = block_sum(den, sh);
  double tx = block_sum(xm, sh);
  if (threadIdx.x == 0) {
    double* p = part + (((long long)b * lam_ld + c) * nchunk + blockIdx.x) * 3;
    p[0] = tn;
    p[1] = td;
    p[2] = tx;
  }
}

// full-sector mode shapes: out[dof][q] (row-major, ld = nm_ld) = T0 x_q with right
// rows times P; x_q = scale_q X[:, src_q] (scale optional), conjugated when dup_q
// (degenerate duplication)
__global__ void k_expand_modes(int nn, const int* __restrict__ free_of_full,
                               const int* __restrict__ node_class, double cr, double sr, int c1,
                               int c2, int cax, const double2* __restrict__ X, long long n,
                               const int* __restrict__ src, const int* __restrict__ conjq, int nm,
                               double pre, double pim, const double* __restrict__ scale,
                               double2* __restrict__ out) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= (long long)nn * nm) return;
  const int i = (int)(idx / nm), q = (int)(idx % nm);
  const int p = free_of_full[i];
  double2 v[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
  if (p >= 0) {
    const double2* x = X + (long long)src[q] * n + 3LL * p;
    const double f = scale ? scale[q] : 1.0;
    for (int c = 0; c < 3; ++c) {
      v[c] = make_double2(f * x[c].x, f * x[c].y);
      if (conjq[q]) v[c].y = -v[c].y;
    }
    if (node_class[i] == TM_NODE_RIGHT) {
      double2 w[3];
      w[c1] = make_double2(cr * v[c1].x - sr * v[c2].x, cr * v[c1].y - sr * v[c2].y);
      w[c2] = make_double2(sr * v[c1].x + cr * v[c2].x, sr * v[c1].y + cr * v[c2].y);
      w[cax] = v[cax];
      const double2 P = make_double2(pre, pim);
      for (int c = 0; c < 3; ++c) v[c] = cmul(P, w[c]);
    }
  }
  for (int c = 0; c < 3; ++c) out[(3LL * i + c) * nm + q] = v[c];
}

// Gram blocks: G[b][p][q] = sum_i conj(Y_p[i]) Z_q[i], p, q < ncol_b (chunk partials)
__global__ __launch_bounds__(256) void k_gram_partial(const double2* __restrict__ Y,
                                                      const double2* __restrict__ Z, long long ldb,
                                                      long long n, const int* __restrict__ ncol, int nmax,
                                                      double2* __restrict__ part, int nchunk) {
  const int b = blockIdx.y, pq = blockIdx.z;
  const int p = pq / nmax, q = pq % nmax;
  if (p >= ncol[b] || q >= ncol[b]) return;
  __shared__ double sh[8];
  const long long i0 = (long long)blockIdx.x * kDotChunk;
  const double2* y = Y + b * ldb + (long long)p * n;
  const double2* z = Z + b * ldb + (long long)q * n;
  double re = 0.0, im = 0.0;
  for (long long i = i0 + threadIdx.x; i < i0 + kDotChunk && i < n; i += 256) {
    double2 u = cconjmul(y[i], z[i]);
    re += u.x;
    im += u.y;
  }
  double tr = block_sum(re, sh);
  double ti = block_sum(im, sh);
  if (threadIdx.x == 0) part[(((long long)b * nmax + p) * nmax + q) * nchunk + blockIdx.x] = make_double2(tr, ti);
}

// out_b[:, c] = sum_{k < nk_b} V_b[:, k] Z_b[k][c] with complex coefficients (row-major, ldc)
__global__ void k_combine_c(const double2* __restrict__ V, long long ldb, long long n,
                            const double2* __restrict__ Zc, int ldc, long long cstride,
                            const int* __restrict__ nk, double2* __restrict__ out, long long ldo) {
  const int b = blockIdx.y;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int K = nk[b];
  const double2* v = V + b * ldb;
  const double2* zb = Zc + b * cstride;
  for (int c = 0; c < K; ++c) {
    double2 acc = make_double2(0.0, 0.0);
    for (int k = 0; k < K; ++k) acc = cadd(acc, cmul(zb[(long long)k * ldc + c], v[(long long)k * n + i]));
    out[b * ldo + (long long)c * n + i] = acc;
  }
}

// copy column q of every problem: dst_b[:] = src_b[:, q] (problems with q < ncol_b)
__global__ void k_copy_col(const double2* __restrict__ src, long long lds, long long n, int q,
                           const int* __restrict__ ncol, double2* __restrict__ dst, long long ldd) {
  const int b = blockIdx.y;
  if (q >= ncol[b]) return;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  dst[b * ldd + i] = src[b * lds + (long long)q * n + i];
}

__global__ void k_flag_cols(const int* __restrict__ ncol, int q, int B, int* __restrict__ act) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < B) act[b] = q < ncol[b] ? 1 : 0;
}

// Tall-skinny C = A^H B on the FP64 matrix cores, partial over row chunks.  A_b: n x ka
// (columns cs apart, problem stride lda_b), B_b: n x kb (kb <= 16); with TWO also
// C2 = A^H B2 from the same loads of A.  grid (nchunk, problems, column super-blocks):
// each block takes kTsChunk rows; its four waves are WA column groups (NAB 16-column
// blocks of A each) times 4 / WA row groups, whose sums are added in LDS in a fixed
// order.  The complex product runs as real v_mfma_f64_16x16x4 products with the
// B columns packed as [Re B | Im B] (eight complex columns per 16-wide operand):
// P = Re(A)^T [Re B | Im B] and Q = Im(A)^T [Re B | Im B], then
// Re C = P_re + Q_im and Im C = P_im - Q_re.  Lane l of a product holds A's row
// (K index) l / 16 of column (M index) l % 16 and B's packed column l % 16 of the same
// row; result register q of lane l is C[4 q + l / 16][l % 16].  Each lane reads its A
// entry as one 16-byte complex and its B entry as the one double it needs.
typedef double v4d __attribute__((ext_vector_type(4)));
template <bool TWO, int NH, int NAB, int WA>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_tsgemm_mfma(const double2* __restrict__ A, long long lda_b,
                                                     const double2* __restrict__ Bm, const double2* __restrict__ Bm2,
                                                     long long ldb_b, long long n, long long cs,
                                                     const int* __restrict__ ka_b, int kb,
                                                     const int* __restrict__ act, int kamax,
                                                     double2* __restrict__ part, double2* __restrict__ part2,
                                                     int nchunk) {
  constexpr int WR = 4 / WA, NS = (TWO ? 2 : 1) * NH, U = NAB * NS >= 8 ? 1 : 4;
  const int b = blockIdx.y;
  if (act && !act[b]) return;
  const int ka = ka_b[b];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wa = wave % WA, wr = wave / WA;
  const int l16 = lane & 15, l4 = lane >> 4;
  const int abase = (blockIdx.z * WA + wa) * NAB * 16;
  if (blockIdx.z * WA * NAB * 16 >= ka) return;  // the whole block (uniform)
  // rows of this wave: a contiguous share of the block's chunk
  const long long r0 = (long long)blockIdx.x * kTsChunk;
  const long long r1 = r0 + kTsChunk < n ? r0 + kTsChunk : n;
  constexpr int kSub = kTsChunk / WR;
  const long long rs = r0 + (long long)wr * kSub;
  const long long re = rs + kSub < r1 ? rs + kSub : r1;
  const double2* pa[NAB];
#pragma unroll
  for (int i = 0; i < NAB; ++i) {
    const int a = abase + 16 * i + l16;
    pa[i] = A + b * lda_b + (long long)(a < ka ? a : 0) * cs;
  }
  const double* pb[NS];
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    const int c = 8 * (j % NH) + (l16 & 7);
    const double2* src = (TWO && j >= NH) ? Bm2 : Bm;
    pb[j] = reinterpret_cast<const double*>(src + b * ldb_b + (long long)(c < kb ? c : 0) * cs) + (l16 >> 3);
  }
  v4d P[NAB][NS], Q[NAB][NS];
#pragma unroll
  for (int i = 0; i < NAB; ++i)
#pragma unroll
    for (int j = 0; j < NS; ++j) P[i][j] = Q[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
  // U row quads per step, all loads of a step issued before its products; enough
  // blocks per CU keep HBM busy meanwhile.  Full steps read without masks: an A column
  // at or past ka, or a B column at or past kb, reads column 0 instead and feeds only
  // results that are not stored.  The rows left over (fewer than 4 U) take one masked
  // step.
  double2 av[U][NAB];
  double bv[U][NS];
  auto products = [&]() {
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int i = 0; i < NAB; ++i)
#pragma unroll
        for (int j = 0; j < NS; ++j) {
          P[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u][i].x, bv[u][j], P[i][j], 0, 0, 0);
          Q[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u][i].y, bv[u][j], Q[i][j], 0, 0, 0);
        }
  };
  if (abase < ka && rs < re) {
    const int nfull = static_cast<int>((re - rs) / (4 * U));  // full steps (wave-uniform)
    const long long rfull = rs + (long long)nfull * 4 * U;
    for (int st = 0; st < nfull; ++st) {
      const long long r = rs + l4 + (long long)st * 4 * U;
#pragma unroll
      for (int u = 0; u < U; ++u) {
#pragma unroll
        for (int i = 0; i < NAB; ++i) av[u][i] = pa[i][r + 4 * u];
#pragma unroll
        for (int j = 0; j < NS; ++j) bv[u][j] = pb[j][2 * (r + 4 * u)];
      }
      products();
    }
    if (rfull < re) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long long ru = rfull + l4 + 4 * u;
        const bool rok = ru < re;
#pragma unroll
        for (int i = 0; i < NAB; ++i) av[u][i] = rok ? pa[i][ru] : make_double2(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < NS; ++j) bv[u][j] = rok ? pb[j][2 * ru] : 0.0;
      }
      products();
    }
  }
  // lanes l % 16 < 8 (real parts of B) assemble the complex results with lane l + 8
  // (every lane takes part in the shuffles) and store them
  auto cval = [&](int i, int j, int q) {
    const double po = __shfl_xor(P[i][j][q], 8, 64), qo = __shfl_xor(Q[i][j][q], 8, 64);
    return make_double2(P[i][j][q] + qo, po - Q[i][j][q]);
  };
  auto store = [&](int i, int j, int q, double2 v) {
    const int c = 8 * (j % NH) + l16, a = abase + 16 * i + 4 * q + l4;
    double2* dst = (TWO && j >= NH) ? part2 : part;
    if (l16 < 8 && a < ka && c < kb) dst[(((long long)b * kamax + a) * 16 + c) * nchunk + blockIdx.x] = v;
  };
  if constexpr (WR > 1) {
    // the row groups' sums added in LDS, in row-group order, by row group 0
    __shared__ double2 red[4][NAB * NS * 4][32];
    const int li = l4 * 8 + (l16 & 7);
#pragma unroll
    for (int i = 0; i < NAB; ++i)
#pragma unroll
      for (int j = 0; j < NS; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double2 v = cval(i, j, q);
          if (l16 < 8) red[wave][(i * NS + j) * 4 + q][li] = v;
        }
    __syncthreads();
    if (wr > 0) return;
#pragma unroll
    for (int i = 0; i < NAB; ++i)
#pragma unroll
      for (int j = 0; j < NS; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          double2 v = red[wa][(i * NS + j) * 4 + q][li];
          for (int g = 1; g < WR; ++g) {
            const double2 o = red[wa + g * WA][(i * NS + j) * 4 + q][li];
            v = make_double2(v.x + o.x, v.y + o.y);
          }
          store(i, j, q, v);
        }
  } else {
#pragma unroll
    for (int i = 0; i < NAB; ++i)
#pragma unroll
      for (int j = 0; j < NS; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) store(i, j, q, cval(i, j, q));
  }
}
#define TM_TSGEMM_MFMA(TWO, NH, NAB, WA)                                                                        \
  template __global__ void k_tsgemm_mfma<TWO, NH, NAB, WA>(const double2*, long long, const double2*,           \
                                                           const double2*, long long, long long, long long,     \
                                                           const int*, int, const int*, int, double2*, double2*, \
                                                           int);
#define TM_TSGEMM_MFMA_ALL(TWO, NH) \
  TM_TSGEMM_MFMA(TWO, NH, 1, 1) TM_TSGEMM_MFMA(TWO, NH, 1, 2) TM_TSGEMM_MFMA(TWO, NH, 1, 4) TM_TSGEMM_MFMA(TWO, NH, 2, 4)
TM_TSGEMM_MFMA_ALL(false, 1)
TM_TSGEMM_MFMA_ALL(false, 2)
TM_TSGEMM_MFMA_ALL(true, 1)
TM_TSGEMM_MFMA_ALL(true, 2)
#undef TM_TSGEMM_MFMA_ALL
#undef TM_TSGEMM_MFMA

// C[b][a][c] = sum over chunks (fixed order)
__global__ void k_tsgemm_reduce(const double2* __restrict__ part, const int* __restrict__ ka_b, int kb,
                                const int* __restrict__ act, int kamax, int nchunk, double2* __restrict__ C) {
  const int b = blockIdx.y;
  if (act && !act[b]) return;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = idx / 16, c = idx % 16;
  if (a >= ka_b[b] || c >= kb) return;
  const double2* p = part + (((long long)b * kamax + a) * 16 + c) * nchunk;
  double re = 0.0, im = 0.0;
  for (int q = 0; q < nchunk; ++q) {
    re += p[q].x;
    im += p[q].y;
  }
  C[((long long)b * kamax + a) * 16 + c] = make_double2(re, im);
}

// Y_b[:, c] = beta Y_b[:, c] + sum_{r0 <= a < min(ka_b, r0 + kTsMaxRows)} A_b[:, a] C_b[a][c],
// c < kc (C_b row-major kamax x 16); one thread per row.  Bases longer than kTsMaxRows
// take one launch per kTsMaxRows rows of C (r0 > 0 with beta = 1).
__global__ __launch_bounds__(256) void k_tsmm(const double2* __restrict__ A, long long lda_b, long long n,
                                              long long cs, const double2* __restrict__ C, const int* __restrict__ ka_b,
                                              int kc, const int* __restrict__ act, int kamax, double alpha,
                                              double beta, double2* __restrict__ Y, long long ldy_b,
                                              const int* __restrict__ yoff, int r0) {
  const int b = blockIdx.y;
  if (act && !act[b]) return;
  const int ka = min(ka_b[b] - r0, kTsMaxRows);
  if (r0 > 0 && ka <= 0) return;  // nothing to add to Y
  __shared__ double2 Cs[kTsMaxRows * 16];
  for (int idx = threadIdx.x; idx < ka * 16; idx += 256) Cs[idx] = C[((long long)b * kamax + r0) * 16 + idx];
  __syncthreads();
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2* Ab = A + b * lda_b + (long long)r0 * cs;
  double2 acc[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) acc[c] = make_double2(0.0, 0.0);
  for (int a = 0; a < ka; ++a) {
    const double2 x = Ab[(long long)a * cs + i];
#pragma unroll
    for (int c = 0; c < 16; ++c) cmac(acc[c], x, Cs[a * 16 + c]);
  }
  double2* Yb = Y + b * ldy_b + (yoff ? (long long)yoff[b] * cs : 0LL);
#pragma unroll
  for (int c = 0; c < 16; ++c)
    if (c < kc) {
      const double2 y = beta != 0.0 ? Yb[(long long)c * cs + i] : make_double2(0.0, 0.0);
      Yb[(long long)c * cs + i] = make_double2(beta * y.x + alpha * acc[c].x, beta * y.y + alpha * acc[c].y);
    }
}

// Work-groups are dealt round-robin over the eight XCDs (each with its own L2):
// block x takes task xcd_task(x, n) of n, so every XCD runs a contiguous run of
// the task list.
__device__ __forceinline__ int xcd_task(int x, int n) {
  const int c = x & 7, q = n >> 3, rem = n & 7;
  return c * q + min(c, rem) + (x >> 3);
}

// Multi-column SpMV with one or two complex block arrays over the same columns:
// Y1_b[:, c] (+)= A1_b X_b[:, c] and, when blocks2 is given, Y2_b[:, c] (+)= A2_b X_b[:, c],
// for c0 <= c < min(c0 + 8 CPL, cmul ncol_b); column c at offset c n.  One wave per
// node row; lane = (slot lane 0..7, column group 0..7): eight slots of the row in
// flight, each lane holding CPL columns eight apart, so a block entry loaded by a
// lane serves CPL columns (and both products share the x loads).  Tasks are
// (row quad, problem), problem-fastest, each XCD taking a contiguous run: the
// problems of a row quad run together on one XCD, and those that share a block
// array (M(k) of one harmonic) find it in that XCD's L2.
template <int CPL>
__global__ __launch_bounds__(256) void k_spmv_pair(int nf, int B, const int* __restrict__ rowptr,
                                                   const int* __restrict__ colidx,
                                                   const double2* const* __restrict__ blocks1,
                                                   const double2* const* __restrict__ blocks2,
                                                   const double2* __restrict__ X, long long ldx, long long n,
                                                   const int* __restrict__ ncol, int cmul, int c0,
                                                   double2* __restrict__ Y1, double2* __restrict__ Y2, long long ldy,
                                                   int acc1, int acc2) {
  const int t = xcd_task(blockIdx.x, gridDim.x);
  const int b = t % B, I = (t / B) * 4 + (threadIdx.x >> 6);
  const int nc = min(8 * CPL, cmul * ncol[b] - c0);
  if (nc <= 0 || I >= nf) return;  // whole waves leave together
  const int lane = threadIdx.x & 63, cg = lane & 7, sl = lane >> 3;
  const double2* A1 = blocks1[b];
  const double2* A2 = blocks2 ? blocks2[b] : nullptr;
  const double2* x = X + b * ldx + (long long)(c0 + cg) * n;
  double2 a1[CPL][3], a2[CPL][3];
#pragma unroll
  for (int j = 0; j < CPL; ++j)
#pragma unroll
    for (int rr = 0; rr < 3; ++rr) a1[j][rr] = a2[j][rr] = make_double2(0.0, 0.0);
  const int s1 = rowptr[I + 1];
  if (cg < nc) {
    int s = rowptr[I] + sl;
    // the next pass's block column is loaded one pass ahead, so only the first
    // pass waits for an index before its x gathers
    int Jn = s < s1 ? colidx[s] : 0;
    for (; s < s1; s += 8) {
      const long long J3 = 3LL * Jn;
      if (s + 8 < s1) Jn = colidx[s + 8];
      double2 xv[CPL][3];
#pragma unroll
      for (int j = 0; j < CPL; ++j) {
        const bool on = cg + 8 * j < nc;
        const double2* xj = x + (long long)(8 * j) * n + J3;
#pragma unroll
        for (int k = 0; k < 3; ++k) xv[j][k] = on ? xj[k] : make_double2(0.0, 0.0);
      }
      const double2* e1 = A1 + (long long)s * 9;
#pragma unroll
      for (int rr = 0; rr < 3; ++rr)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double2 e = e1[3 * rr + k];
#pragma unroll
          for (int j = 0; j < CPL; ++j) cmac(a1[j][rr], e, xv[j][k]);
        }
      if (A2) {
        const double2* e2 = A2 + (long long)s * 9;
#pragma unroll
        for (int rr = 0; rr < 3; ++rr)
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const double2 e = e2[3 * rr + k];
#pragma unroll
            for (int j = 0; j < CPL; ++j) cmac(a2[j][rr], e, xv[j][k]);
          }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < CPL; ++j)
#pragma unroll
    for (int rr = 0; rr < 3; ++rr)
#pragma unroll
      for (int o = 8; o < 64; o <<= 1) {
        a1[j][rr].x += __shfl_xor(a1[j][rr].x, o, 64);
        a1[j][rr].y += __shfl_xor(a1[j][rr].y, o, 64);
        if (A2) {
          a2[j][rr].x += __shfl_xor(a2[j][rr].x, o, 64);
          a2[j][rr].y += __shfl_xor(a2[j][rr].y, o, 64);
        }
      }
  if (sl != 0) return;
#pragma unroll
  for (int j = 0; j < CPL; ++j) {
    if (cg + 8 * j >= nc) break;
    const long long o = b * ldy + (long long)(c0 + cg + 8 * j) * n + 3LL * I;
#pragma unroll
    for (int rr = 0; rr < 3; ++rr) Y1[o + rr] = acc1 ? cadd(Y1[o + rr], a1[j][rr]) : a1[j][rr];
    if (A2)
#pragma unroll
      for (int rr = 0; rr < 3; ++rr) Y2[o + rr] = acc2 ? cadd(Y2[o + rr], a2[j][rr]) : a2[j][rr];
  }
}
template __global__ void k_spmv_pair<1>(int, int, const int*, const int*, const double2* const*,
                                        const double2* const*, const double2*, long long, long long, const int*, int,
                                        int, double2*, double2*, long long, int, int);
template __global__ void k_spmv_pair<2>(int, int, const int*, const int*, const double2* const*,
                                        const double2* const*, const double2*, long long, long long, const int*, int,
                                        int, double2*, double2*, long long, int, int);

// A slot is real when the phase parts of K and M vanish at it and at its
// transpose slot (Kp, dKp, Mp all zero) and its row and column nodes have no
// pinned DOF: its K(k) block is then Kc + s dKc and its M(k) block Mc for every
// harmonic, with zero imaginary parts (k_build_phase_blocks gives exactly these).
__global__ void k_real_slots(long long nslot, const int* __restrict__ slot_row, const int* __restrict__ colidx,
                             const int* __restrict__ tslot, const unsigned char* __restrict__ pinned,
                             const double* __restrict__ Kp, const double* __restrict__ dKp,
                             const double* __restrict__ Mp, unsigned char* __restrict__ real) {
  const long long sl = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (sl >= nslot) return;
  const int I = slot_row[sl], J = colidx[sl];
  const long long t = tslot[sl];
  bool r = true;
  for (int c = 0; c < 3; ++c) r = r && !pinned[3 * I + c] && !pinned[3 * J + c];
  for (int e = 0; e < 9; ++e)
    r = r && Kp[sl * 9 + e] == 0.0 && Kp[t * 9 + e] == 0.0 && dKp[sl * 9 + e] == 0.0 && dKp[t * 9 + e] == 0.0 &&
        Mp[sl * 9 + e] == 0.0 && Mp[t * 9 + e] == 0.0;
  real[sl] = r ? 1 : 0;
}

// the real slot arrays of an operator packed for k_spmv_km: per slot (Kc_e, dKc_e)
// for its 9 entries, then Mc_0..8 and a zero (224 B, one 16-byte-aligned record)
__global__ void k_pack_real_slots(long long nslot, const double* __restrict__ Kc, const double* __restrict__ dKc,
                                  const double* __restrict__ Mc, double2* __restrict__ pack) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= 14 * nslot) return;
  const long long sl = idx / 14;
  const int u = (int)(idx - 14 * sl);
  double2 v;
  if (u < 9) v = make_double2(Kc[sl * 9 + u], dKc[sl * 9 + u]);
  else if (u < 13) v = make_double2(Mc[sl * 9 + 2 * (u - 9)], Mc[sl * 9 + 2 * (u - 9) + 1]);
  else v = make_double2(Mc[sl * 9 + 8], 0.0);
  pack[idx] = v;
}

// Y1 = K(k) X and Y2 = M(k) X for the Hermitian pencil, as k_spmv_pair, with the
// real slots read from the operator's packed real arrays (Kc + s dKc, Mc; 224 B
// per slot in 14 16-byte pieces, the same for every problem of the operator and so
// found in L2 by the problems after the first) instead of the problem's complex blocks (288 B per
// slot and problem).  A real slot's products are the complex path's with the zero
// imaginary terms left out: bit-identical sums.
// One wave per node row: lane = (slot lane sl, column lane cg), eight slots of the row
// at a time.  The eight column lanes of a slot all need its whole record, so the
// wave loads the eight records once, contiguous and coalesced (112 pieces over its 64
// lanes), into its own LDS, and each lane reads its slot's record from there (a
// broadcast to the eight column lanes): the vector memory path then returns each
// record once instead of eight times.
template <int CPL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_spmv_km(int nf, int B, const int* __restrict__ rowptr,
                                                 const int* __restrict__ colidx, const SplitKM* __restrict__ sk,
                                                 const double2* const* __restrict__ kblocks,
                                                 const double2* const* __restrict__ mblocks,
                                                 const double2* __restrict__ X, long long ldx, long long n,
                                                 const int* __restrict__ ncol, int c0, double2* __restrict__ Y1,
                                                 double2* __restrict__ Y2, long long ldy) {
  constexpr int kRec = 14;  // 16-byte pieces per slot record
  __shared__ double2 rec[4][8 * kRec];
  const int t = xcd_task(blockIdx.x, gridDim.x);
  const int b = t % B, I = (t / B) * 4 + (threadIdx.x >> 6);
  const int nc = min(8 * CPL, ncol[b] - c0);
  if (nc <= 0 || I >= nf) return;  // whole waves leave together
  const int lane = threadIdx.x & 63, cg = lane & 7, sl = lane >> 3;
  double2* wrec = rec[threadIdx.x >> 6];
  const SplitKM q = sk[b];
  const double2* A1 = kblocks[b];
  const double2* A2 = mblocks[b];
  const double2* x = X + b * ldx + (long long)(c0 + cg) * n;
  double2 a1[CPL][3], a2[CPL][3];
#pragma unroll
  for (int j = 0; j < CPL; ++j)
#pragma unroll
    for (int rr = 0; rr < 3; ++rr) a1[j][rr] = a2[j][rr] = make_double2(0.0, 0.0);
  const int r0 = rowptr[I], s1 = rowptr[I + 1];
  const bool col_on = cg < nc;
  int Jn = r0 + sl < s1 ? colidx[r0 + sl] : 0, Rn = r0 + sl < s1 ? q.real[r0 + sl] : 0;
  for (int base = r0; base < s1; base += 8) {  // wave-uniform trip count
    const int s = base + sl;
    const long long J3 = 3LL * Jn;
    const int re = Rn;
    if (s + 8 < s1) {
      Jn = colidx[s + 8];
      Rn = q.real[s + 8];
    }
    // the eight records of slots base .. base + 7 (those of the row) into this wave's LDS
    const double2* src = q.pack + (long long)base * kRec;
    const int npiece = min(8, s1 - base) * kRec;
    for (int e = lane; e < npiece; e += 64) wrec[e] = src[e];
    __builtin_amdgcn_s_waitcnt(0xc07f);  // this wave's LDS writes done before any lane reads
    __builtin_amdgcn_wave_barrier();
    if (s < s1 && col_on) {
      double2 xv[CPL][3];
#pragma unroll
      for (int j = 0; j < CPL; ++j) {
        const bool on = cg + 8 * j < nc;
        const double2* xj = x + (long long)(8 * j) * n + J3;
#pragma unroll
        for (int k = 0; k < 3; ++k) xv[j][k] = on ? xj[k] : make_double2(0.0, 0.0);
      }
      if (re) {
        const double2* pk = wrec + sl * kRec;
        double mcv[10];
#pragma unroll
        for (int u = 0; u < 5; ++u) {
          const double2 m2 = pk[9 + u];
          mcv[2 * u] = m2.x;
          mcv[2 * u + 1] = m2.y;
        }
#pragma unroll
        for (int rr = 0; rr < 3; ++rr)
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const double2 kk = pk[3 * rr + k];
            const double kv = fma(q.s, kk.y, kk.x);
            const double mv = mcv[3 * rr + k];
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
              a1[j][rr].x = fma(kv, xv[j][k].x, a1[j][rr].x);
              a1[j][rr].y = fma(kv, xv[j][k].y, a1[j][rr].y);
              a2[j][rr].x = fma(mv, xv[j][k].x, a2[j][rr].x);
              a2[j][rr].y = fma(mv, xv[j][k].y, a2[j][rr].y);
            }
          }
      } else {
        const double2* e1 = A1 + (long long)s * 9;
        const double2* e2 = A2 + (long long)s * 9;
#pragma unroll
        for (int rr = 0; rr < 3; ++rr)
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const double2 u = e1[3 * rr + k], v = e2[3 * rr + k];
#pragma unroll
            for (int j = 0; j < CPL; ++j) {
              cmac(a1[j][rr], u, xv[j][k]);
              cmac(a2[j][rr], v, xv[j][k]);
            }
          }
      }
    }
    __builtin_amdgcn_s_waitcnt(0xc07f);  // every lane's reads of the records done before the next writes
    __builtin_amdgcn_wave_barrier();
  }
#pragma unroll
  for (int j = 0; j < CPL; ++j)
#pragma unroll
    for (int rr = 0; rr < 3; ++rr)
#pragma unroll
      for (int o = 8; o < 64; o <<= 1) {
        a1[j][rr].x += __shfl_xor(a1[j][rr].x, o, 64);
        a1[j][rr].y += __shfl_xor(a1[j][rr].y, o, 64);
        a2[j][rr].x += __shfl_xor(a2[j][rr].x, o, 64);
        a2[j][rr].y += __shfl_xor(a2[j][rr].y, o, 64);
      }
  if (sl != 0) return;
#pragma unroll
  for (int j = 0; j < CPL; ++j) {
    if (cg + 8 * j >= nc) break;
    const long long o = b * ldy + (long long)(c0 + cg + 8 * j) * n + 3LL * I;
#pragma unroll
    for (int rr = 0; rr < 3; ++rr) {
      Y1[o + rr] = a1[j][rr];
      Y2[o + rr] = a2[j][rr];
    }
  }
}
template __global__ void k_spmv_km<1>(int, int, const int*, const int*, const SplitKM*, const double2* const*,
                                      const double2* const*, const double2*, long long, long long, const int*, int,
                                      double2*, double2*, long long);
template __global__ void k_spmv_km<2>(int, int, const int*, const int*, const SplitKM*, const double2* const*,
                                      const double2* const*, const double2*, long long, long long, const int*, int,
                                      double2*, double2*, long long);

// Coriolis pencil: from BX = [M a; M b] (per column), AX = [M b; M a] (the
// -D b term is added after), c < ncol_b, n rows per half, columns cs apart
__global__ void k_qep_swap_halves(const double2* __restrict__ BX, double2* __restrict__ AX, long long ld, long long n,
                                  long long cs, const int* __restrict__ ncol) {
  const int b = blockIdx.y, c = blockIdx.z;
  if (c >= ncol[b]) return;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long o = b * ld + (long long)c * cs;
  AX[o + i] = BX[o + n + i];
  AX[o + n + i] = BX[o + i];
}

// dst_b[doff + i, d0_b + c] = alpha src_b[soff + i, s0_b + c] + beta dst (i < len, c < ncol),
// columns `stride` apart: halves of the 2n-long vectors of the Coriolis pencil
__global__ void k_axpy_half(const double2* __restrict__ src, long long lds, const int* __restrict__ s0v,
                            long long soff, double2* __restrict__ dst, long long ldd, const int* __restrict__ d0v,
                            long long doff, int ncol, long long len, long long stride, double alpha, double beta,
                            const int* __restrict__ act) {
  const int b = blockIdx.y;
  if (act && !act[b]) return;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= len) return;
  const int s0 = s0v ? s0v[b] : 0, d0 = d0v ? d0v[b] : 0;
  for (int c = 0; c < ncol; ++c) {
    const double2 x = src[b * lds + (long long)(s0 + c) * stride + soff + i];
    double2* y = dst + b * ldd + (long long)(d0 + c) * stride + doff + i;
    const double2 yo = beta != 0.0 ? *y : make_double2(0.0, 0.0);
    *y = make_double2(alpha * x.x + beta * yo.x, alpha * x.y + beta * yo.y);
  }
}

// per column: Re(x_head^H y_head) over the first len entries (chunk partials), columns stride apart
__global__ __launch_bounds__(256) void k_head_dot(const double2* __restrict__ X, const double2* __restrict__ Y,
                                                  long long ldb, long long stride, long long len,
                                                  const int* __restrict__ ncol, int ldc, double* __restrict__ part,
                                                  int nchunk) {
  const int b = blockIdx.y, c = blockIdx.z;
  if (c >= ncol[b]) return;
  __shared__ double sh[8];
  const long long i0 = (long long)blockIdx.x * kDotChunk;
  double s = 0.0;
  for (long long i = i0 + threadIdx.x; i < i0 + kDotChunk && i < len; i += 256) {
    const long long o = b * ldb + (long long)c * stride + i;
    s += X[o].x * Y[o].x + X[o].y * Y[o].y;
  }
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) part[((long long)b * ldc + c) * nchunk + blockIdx.x] = t;
}

// dst_b[:, d0_b + c] = src_b[:, s0_b + c] for c < ncol (problems with act_b); the
// per-problem column offsets s0v / d0v are optional (else 0)
__global__ void k_copy_cols(const double2* __restrict__ src, long long lds, const int* __restrict__ s0v,
                            double2* __restrict__ dst, long long ldd, const int* __restrict__ d0v, int ncol,
                            long long n, const int* __restrict__ act) {
  const int b = blockIdx.y;
  if (act && !act[b]) return;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s0 = s0v ? s0v[b] : 0, d0 = d0v ? d0v[b] : 0;
  for (int c = 0; c < ncol; ++c) dst[b * ldd + (long long)(d0 + c) * n + i] = src[b * lds + (long long)(s0 + c) * n + i];
}

// Refinement of a shift-invert solve with the complex-FP32 factor: for c < ncol,
//   R_b[:, c] = Y_b[:, s0_b + c] - (K W_b[:, c] - sigma_b M W_b[:, c])
// with Y = M V (the right-hand sides), KW and MW the FP64 products of the current
// solution W; R overwrites KW.
__global__ void k_mixed_residual(const double2* __restrict__ Y, long long ldy, const int* __restrict__ s0v,
                                 double2* __restrict__ KW, const double2* __restrict__ MW, long long ldw,
                                 const ProbParam* __restrict__ params, int ncol, long long n,
                                 const int* __restrict__ act) {
  const int b = blockIdx.y;
  if (act && !act[b]) return;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double sg = params[b].sigma;
  const int s0 = s0v ? s0v[b] : 0;
  for (int c = 0; c < ncol; ++c) {
    const double2 y = Y[b * ldy + (long long)(s0 + c) * n + i];
    const long long e = b * ldw + (long long)c * n + i;
    const double2 k = KW[e], m = MW[e];
    KW[e] = make_double2(y.x - k.x + sg * m.x, y.y - k.y + sg * m.y);
  }
}

// Partial sums of |R_b[:, c]|^2 and |Y_b[:, s0_b + c]|^2 over kDotChunk rows per block
// (grid x), reduced on the host in chunk order: part[((b ncol + c) nchunk + x) 2 + {0, 1}]
__global__ __launch_bounds__(256) void k_colnorm_partial(const double2* __restrict__ R, long long ldr,
                                                         const double2* __restrict__ Y, long long ldy,
                                                         const int* __restrict__ s0v, int ncol, long long n,
                                                         const int* __restrict__ act, double* __restrict__ part,
                                                         int nchunk) {
  const int b = blockIdx.y;
  if (act && !act[b]) return;
  const int s0 = s0v ? s0v[b] : 0;
  const long long r0 = (long long)blockIdx.x * kDotChunk;
  __shared__ double red[2][256];
  for (int c = 0; c < ncol; ++c) {
    double sr = 0.0, sy = 0.0;
    for (long long i = r0 + threadIdx.x; i < r0 + kDotChunk && i < n; i += 256) {
      const double2 r = R[b * ldr + (long long)c * n + i], y = Y[b * ldy + (long long)(s0 + c) * n + i];
      sr = fma(r.x, r.x, fma(r.y, r.y, sr));
      sy = fma(y.x, y.x, fma(y.y, y.y, sy));
    }
    red[0][threadIdx.x] = sr;
    red[1][threadIdx.x] = sy;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) {
        red[0][threadIdx.x] += red[0][threadIdx.x + o];
        red[1][threadIdx.x] += red[1][threadIdx.x + o];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      part[(((long long)b * ncol + c) * nchunk + blockIdx.x) * 2] = red[0][0];
      part[(((long long)b * ncol + c) * nchunk + blockIdx.x) * 2 + 1] = red[1][0];
    }
    __syncthreads();
  }
}

// W_b[:, c] += D_b[:, c], c < ncol
__global__ void k_add_cols(double2* __restrict__ W, const double2* __restrict__ D, long long ld, int ncol,
                           long long n, const int* __restrict__ act) {
  const int b = blockIdx.y;
  if (act && !act[b]) return;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < ncol; ++c) {
    const long long e = b * ld + (long long)c * n + i;
    W[e] = make_double2(W[e].x + D[e].x, W[e].y + D[e].y);
  }
}

// Y_b[:, c] = |A_b| |X_b[:, c]| (entrywise moduli; real result in .x) for c < ncol_b
// |A_q| |x| for up to three block arrays over the same columns (blocks3 optional):
// Y_q[:, c] = |A_q| |X[:, c]| (real, in the .x part), c < ncol_b.  One wave per node
// row, lanes = 8 block slots x 8 columns (columns 8 blockIdx.y ..), as k_spmv_pair:
// each block row is read once for all the columns and |x| is formed once for all
// the arrays; (row quad, problem) tasks problem-fastest, each XCD a contiguous run.
// With sk (the Hermitian pencil: no third array), real slots read the packed real
// arrays as k_spmv_km does.  Moduli as the hardware square root of re^2 + im^2 (no IEEE refinement): operator
// entries and vectors are far from the overflow range, and the products only size a
// rounding floor.
__global__ __launch_bounds__(256) void k_spmv_abs(int nf, const int* __restrict__ rowptr,
                                                  const int* __restrict__ colidx,
                                                  const double2* const* __restrict__ blocks1,
                                                  const double2* const* __restrict__ blocks2,
                                                  const double2* const* __restrict__ blocks3,
                                                  const double2* __restrict__ X, long long ldx, long long cs,
                                                  const int* __restrict__ ncol, double2* __restrict__ Y1,
                                                  double2* __restrict__ Y2, double2* __restrict__ Y3, long long ldy,
                                                  const SplitKM* __restrict__ sk) {
  const int B = gridDim.x / ((nf + 3) / 4);
  const int t = xcd_task(blockIdx.x, gridDim.x);
  const int b = t % B, I = (t / B) * 4 + (threadIdx.x >> 6);
  if (I >= nf) return;
  const int lane = threadIdx.x & 63, c = 8 * blockIdx.y + (lane & 7), sl = lane >> 3;
  const bool on = c < ncol[b];
  const double2* A[3] = {blocks1[b], blocks2[b], blocks3 ? blocks3[b] : nullptr};
  const double2* x = X + b * ldx + (long long)c * cs;
  double acc[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  const int s1 = rowptr[I + 1];
  if (on)
    for (int s = rowptr[I] + sl; s < s1; s += 8) {
      const long long J3 = 3LL * colidx[s];
      double ax[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const double2 v = x[J3 + q];
        ax[q] = __builtin_amdgcn_sqrt(fma(v.x, v.x, v.y * v.y));
      }
      if (sk && sk[b].real[s]) {  // real K and M entries (pair k_spmv_km): their moduli are |entry|
        const double2* pk = sk[b].pack + (long long)s * 14;
        const double sb = sk[b].s;
        double mcv[10];
#pragma unroll
        for (int u = 0; u < 5; ++u) {
          const double2 m2 = pk[9 + u];
          mcv[2 * u] = m2.x;
          mcv[2 * u + 1] = m2.y;
        }
#pragma unroll
        for (int rr = 0; rr < 3; ++rr)
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            const double2 kk = pk[3 * rr + q];
            acc[0][rr] = fma(fabs(fma(sb, kk.y, kk.x)), ax[q], acc[0][rr]);
            acc[1][rr] = fma(fabs(mcv[3 * rr + q]), ax[q], acc[1][rr]);
          }
        continue;
      }
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        if (!A[m]) continue;
        const double2* blk = A[m] + (long long)s * 9;
#pragma unroll
        for (int rr = 0; rr < 3; ++rr)
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            const double2 e = blk[3 * rr + q];
            acc[m][rr] = fma(__builtin_amdgcn_sqrt(fma(e.x, e.x, e.y * e.y)), ax[q], acc[m][rr]);
          }
      }
    }
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int rr = 0; rr < 3; ++rr)
#pragma unroll
      for (int o = 8; o < 64; o <<= 1) acc[m][rr] += __shfl_xor(acc[m][rr], o, 64);
  if (sl == 0 && on) {
    double2* Y[3] = {Y1, Y2, Y3};
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      if (!A[m]) continue;
      double2* y = Y[m] + b * ldy + (long long)c * cs + 3LL * I;
#pragma unroll
      for (int rr = 0; rr < 3; ++rr) y[rr] = make_double2(acc[m][rr], 0.0);
    }
  }
}

// per column: sum_i (AK_i + |lam| AM_i)^2, or with AD: sum_i (AK_i + |lam| AD_i + lam^2 AM_i)^2
// (real parts of n entries, columns cs apart), chunk partials
__global__ __launch_bounds__(256) void k_floor_partial(const double2* __restrict__ AK, const double2* __restrict__ AD,
                                                       const double2* __restrict__ AM, long long ldb, long long n,
                                                       long long cs, const double* __restrict__ lam,
                                                       int lam_ld, const int* __restrict__ ncol,
                                                       double* __restrict__ part, int nchunk) {
  const int b = blockIdx.y, c = blockIdx.z;
  if (c >= ncol[b]) return;
  __shared__ double sh[8];
  const long long i0 = (long long)blockIdx.x * kDotChunk;
  const double l = fabs(lam[(long long)b * lam_ld + c]);
  double s = 0.0;
  for (long long i = i0 + threadIdx.x; i < i0 + kDotChunk && i < n; i += 256) {
    const long long o = b * ldb + (long long)c * cs + i;
    const double t = AD ? AK[o].x + l * AD[o].x + l * l * AM[o].x : AK[o].x + l * AM[o].x;
    s += t * t;
  }
  const double ts = block_sum(s, sh);
  if (threadIdx.x == 0) part[((long long)b * lam_ld + c) * nchunk + blockIdx.x] = ts;
}

__global__ void k_real_to_complex(const double* __restrict__ x, double2* __restrict__ y, long long n) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) y[i] = make_double2(x[i], 0.0);
}

}  // namespace tme
