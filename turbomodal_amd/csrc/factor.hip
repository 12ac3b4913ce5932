// Supernodal multifrontal LDL^H of A(k) = K(k) - sigma M(k) and its batched
// triangular solves (gfx950).
//
// The reference factors each harmonic separately with Eigen's scalar
// SimplicialLDLT (src/hermitian_lanczos.cpp:12-56; real harmonics through
// include/turbomodal/sym_shift_invert_ldlt.hpp:25-112) - LDL^T/LDL^H without
// pivoting.  Here every front of the nested-dissection tree is a dense block
// held as an m x w panel (kept: it becomes L and D) plus an r x r update
// matrix (consumed by the parent).  A tree level is one set of launches over
// all fronts of that level and all problems of the batch (grid.y = problem):
//
//   k_front_assemble  zero the front, add A(k) blocks (formed on the fly from
//                     the phase-split operator arrays: K_c + s dK_c - sigma M_c
//                     + P(.)_p + conj(P)(.)_p^T), extend-add the children's
//                     update matrices (children in fixed order -> deterministic)
//   k_diag_ldl        unblocked LDL^H of a 32-wide diagonal block in LDS
//   k_panel_trsm      L21 = F21 L11^-H D^-1, one row per thread
//   k_schur_tile      trailing update F22 -= L21 D L21^H on 64x64 tiles
//
// Triangular solves reuse the same tree: forward (leaves -> root) with update
// vectors passed to the parent, diagonal scaling fused into the backward
// sweep, backward (root -> leaves) gathering already-solved ancestor values.
#include "kernels.hpp"

namespace tme {

namespace {

constexpr int NBMAX = 32;
constexpr int TILE = 64;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// conj(a) * b
__device__ __forceinline__ double2 cconjmul(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// front element (i, j), i >= j: panel for pivot columns, update matrix otherwise
__device__ __forceinline__ double2* front_at(double2* Pn, double2* Un, int m, int w, int r, int i,
                                             int j) {
  return j < w ? Pn + i + (long long)j * m : Un + (i - w) + (long long)(j - w) * r;
}

__device__ __forceinline__ double shfl_d(double v, int src) { return __shfl(v, src, 64); }

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

struct FactorArgs {
  const DevSn* sn;
  const int* struct_nodes;
  const int* relmap;
  const int* asm_slot;
  const int* asm_pos;
  const int* tslot;
  const unsigned char* pinned;
  OpArrays op;
  double2* L;
  double2* U;
  long long PE, UE;
  int* status;
  const ProbParam* params;
};

__global__ __launch_bounds__(256) void k_front_assemble(FactorArgs fa, const AsmTask* __restrict__ tasks,
                                                        int t0) {
  const AsmTask t = tasks[t0 + blockIdx.x];
  const int b = blockIdx.y;
  const DevSn s = fa.sn[t.sn];
  const int w = 3 * s.ncol, r = 3 * s.nstr, m = w + r;
  double2* Pn = fa.L + b * fa.PE + s.poff;
  double2* Un = fa.U + b * fa.UE + s.uoff;
  const int tid = threadIdx.x;
  const int j0 = 3 * t.n0, j1 = 3 * t.n1;
  for (int j = j0; j < j1; ++j)
    for (int i = j + tid; i < m; i += 256) *front_at(Pn, Un, m, w, r, i, j) = make_double2(0.0, 0.0);
  __syncthreads();
  const ProbParam pp = fa.params[b];
  const int nent = (t.a1 - t.a0) * 9;
  for (int idx = tid; idx < nent; idx += 256) {
    const int q = t.a0 + idx / 9, cc = idx % 9, ci = cc / 3, cj = cc % 3;
    const int pos = fa.asm_pos[q];
    const int rowpos = pos >> 16, colpos = pos & 0xffff;
    const int gi = 3 * rowpos + ci, gj = 3 * colpos + cj;
    if (gi < gj) continue;
    const int rnode = rowpos < s.ncol ? s.col0 + rowpos : fa.struct_nodes[s.str0 + rowpos - s.ncol];
    const int cnode = s.col0 + colpos;
    double2 v;
    if (fa.pinned[3 * rnode + ci] || fa.pinned[3 * cnode + cj]) {
      v = make_double2(gi == gj ? 1.0 : 0.0, 0.0);
    } else {
      const long long sl = fa.asm_slot[q];
      const long long tt = fa.tslot[sl];
      const long long e1 = sl * 9 + 3 * ci + cj;
      const long long e2 = tt * 9 + 3 * cj + ci;
      double c = fa.op.Kc[e1] + pp.s * fa.op.dKc[e1] - pp.sigma * fa.op.Mc[e1];
      double p = fa.op.Kp[e1] + pp.s * fa.op.dKp[e1] - pp.sigma * fa.op.Mp[e1];
      double qv = fa.op.Kp[e2] + pp.s * fa.op.dKp[e2] - pp.sigma * fa.op.Mp[e2];
      v = make_double2(c + pp.pre * (p + qv), pp.pim * (p - qv));
    }
    *front_at(Pn, Un, m, w, r, gi, gj) = v;
  }
  __syncthreads();
  for (int k = 0; k < 2; ++k) {
    const int c = k == 0 ? s.child0 : s.child1;
    if (c < 0) continue;
    const DevSn cs = fa.sn[c];
    const int rc = 3 * cs.nstr;
    const double2* Uc = fa.U + b * fa.UE + cs.uoff;
    const int* rm = fa.relmap + cs.rel0;
    for (int jc = 3 * t.clo[k]; jc < 3 * t.chi[k]; ++jc) {
      const int gj = 3 * rm[jc / 3] + jc % 3;
      for (int ic = jc + tid; ic < rc; ic += 256) {
        const int gi = 3 * rm[ic / 3] + ic % 3;
        double2* dst = front_at(Pn, Un, m, w, r, gi, gj);
        *dst = cadd(*dst, Uc[ic + (long long)jc * rc]);
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_diag_ldl(FactorArgs fa, const int* __restrict__ tasks, int t0,
                                                  int kb) {
  const int id = tasks[t0 + blockIdx.x];
  const int b = blockIdx.y;
  const DevSn s = fa.sn[id];
  const int w = 3 * s.ncol, m = w + 3 * s.nstr;
  const int jb = min(NBMAX, w - kb);
  double2* Pn = fa.L + b * fa.PE + s.poff;
  __shared__ double2 A[NBMAX][NBMAX + 1];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < jb * jb; idx += 256) {
    int i = idx % jb, j = idx / jb;
    if (i >= j) A[i][j] = Pn[(kb + i) + (long long)(kb + j) * m];
  }
  __syncthreads();
  for (int j = 0; j < jb; ++j) {
    const double d = A[j][j].x;
    if (tid == 0 && (d == 0.0 || !isfinite(d))) atomicOr(&fa.status[b], 1);
    const double inv = 1.0 / d;
    for (int i = j + 1 + tid; i < jb; i += 256) A[i][j] = make_double2(A[i][j].x * inv, A[i][j].y * inv);
    __syncthreads();
    const int nt = jb - j - 1;
    for (int idx = tid; idx < nt * nt; idx += 256) {
      int ii = j + 1 + idx % nt, kk = j + 1 + idx / nt;
      if (ii >= kk) {
        double2 l = A[ii][j], lk = A[kk][j];
        double2 u = cmulc(l, lk);
        A[ii][kk] = make_double2(A[ii][kk].x - u.x * d, A[ii][kk].y - u.y * d);
      }
    }
    __syncthreads();
  }
  for (int idx = tid; idx < jb * jb; idx += 256) {
    int i = idx % jb, j = idx / jb;
    if (i > j) Pn[(kb + i) + (long long)(kb + j) * m] = A[i][j];
    else if (i == j) Pn[(kb + i) + (long long)(kb + j) * m] = make_double2(A[i][i].x, 0.0);
  }
}

__global__ __launch_bounds__(256) void k_panel_trsm(FactorArgs fa, const RowTask* __restrict__ tasks,
                                                    int t0, int kb) {
  const RowTask t = tasks[t0 + blockIdx.x];
  const int b = blockIdx.y;
  const DevSn s = fa.sn[t.sn];
  const int w = 3 * s.ncol, m = w + 3 * s.nstr;
  const int jb = min(NBMAX, w - kb);
  double2* Pn = fa.L + b * fa.PE + s.poff;
  __shared__ double2 Lk[NBMAX][NBMAX + 1];
  __shared__ double dk[NBMAX];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < jb * jb; idx += 256) {
    int i = idx % jb, j = idx / jb;
    if (i > j) Lk[i][j] = Pn[(kb + i) + (long long)(kb + j) * m];
    if (i == j) dk[i] = Pn[(kb + i) + (long long)(kb + i) * m].x;
  }
  __syncthreads();
  const int i = t.r0 + tid;
  if (i >= t.r1) return;
  double2 x[NBMAX];
  for (int j = 0; j < jb; ++j) x[j] = Pn[i + (long long)(kb + j) * m];
  for (int j = 0; j < jb; ++j) {
    double2 v = x[j];
    for (int k = 0; k < j; ++k) v = csub(v, cmulc(x[k], Lk[j][k]));
    x[j] = v;
  }
  for (int j = 0; j < jb; ++j) {
    const double inv = 1.0 / dk[j];
    Pn[i + (long long)(kb + j) * m] = make_double2(x[j].x * inv, x[j].y * inv);
  }
}

__global__ __launch_bounds__(256) void k_schur_tile(FactorArgs fa, const TileTask* __restrict__ tasks,
                                                    int t0, int kb) {
  const TileTask t = tasks[t0 + blockIdx.x];
  const int b = blockIdx.y;
  const DevSn s = fa.sn[t.sn];
  const int w = 3 * s.ncol, r = 3 * s.nstr, m = w + r;
  const int jb = min(NBMAX, w - kb);
  double2* Pn = fa.L + b * fa.PE + s.poff;
  double2* Un = fa.U + b * fa.UE + s.uoff;
  __shared__ double2 Li[TILE][NBMAX + 1];
  __shared__ double2 Lj[TILE][NBMAX + 1];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < TILE * jb; idx += 256) {
    const int ii = idx % TILE, k = idx / TILE;
    const double d = Pn[(kb + k) + (long long)(kb + k) * m].x;
    const int gi = t.i0 + ii, gj = t.j0 + ii;
    double2 a = gi < m ? Pn[gi + (long long)(kb + k) * m] : make_double2(0.0, 0.0);
    Li[ii][k] = make_double2(a.x * d, a.y * d);
    Lj[ii][k] = gj < m ? Pn[gj + (long long)(kb + k) * m] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  const int tx = tid % 16, ty = tid / 16;
  double2 acc[4][4];
  for (int p = 0; p < 4; ++p)
    for (int q = 0; q < 4; ++q) acc[p][q] = make_double2(0.0, 0.0);
  for (int k = 0; k < jb; ++k) {
    double2 a[4], bq[4];
    for (int p = 0; p < 4; ++p) a[p] = Li[tx + 16 * p][k];
    for (int q = 0; q < 4; ++q) bq[q] = Lj[ty + 16 * q][k];
    for (int p = 0; p < 4; ++p)
      for (int q = 0; q < 4; ++q) acc[p][q] = cadd(acc[p][q], cmulc(a[p], bq[q]));
  }
  for (int q = 0; q < 4; ++q) {
    const int gj = t.j0 + ty + 16 * q;
    if (gj >= m) continue;
    for (int p = 0; p < 4; ++p) {
      const int gi = t.i0 + tx + 16 * p;
      if (gi >= m || gi < gj) continue;
      double2* dst = front_at(Pn, Un, m, w, r, gi, gj);
      *dst = csub(*dst, acc[p][q]);
    }
  }
}

// ---------------------------------------------------------------------------
// triangular solves
// ---------------------------------------------------------------------------
struct SolveArgs {
  const DevSn* sn;
  const int* struct_nodes;
  const int* relmap;
  const double2* L;
  double2* T;
  long long PE, TE;
  double2* X;
  long long ldx;
  const int* active;
};

__global__ __launch_bounds__(256) void k_fwd_asm(SolveArgs sa, const int* __restrict__ level_sn, int s0) {
  const int b = blockIdx.y;
  if (sa.active && !sa.active[b]) return;
  const DevSn s = sa.sn[level_sn[s0 + blockIdx.x]];
  const int w = 3 * s.ncol, r = 3 * s.nstr;
  double2* x = sa.X + b * sa.ldx + 3LL * s.col0;
  double2* ts = sa.T + b * sa.TE + s.toff;
  const int tid = threadIdx.x;
  for (int q = tid; q < r; q += 256) ts[q] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int k = 0; k < 2; ++k) {
    const int c = k == 0 ? s.child0 : s.child1;
    if (c < 0) continue;
    const DevSn cs = sa.sn[c];
    const int rc = 3 * cs.nstr;
    const double2* tc = sa.T + b * sa.TE + cs.toff;
    const int* rm = sa.relmap + cs.rel0;
    for (int q = tid; q < rc; q += 256) {
      const int pos = 3 * rm[q / 3] + q % 3;
      if (pos < w) x[pos] = cadd(x[pos], tc[q]);
      else ts[pos - w] = cadd(ts[pos - w], tc[q]);
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_fwd_trsv(SolveArgs sa, const int* __restrict__ level_sn, int s0) {
  const int b = blockIdx.y;
  if (sa.active && !sa.active[b]) return;
  const DevSn s = sa.sn[level_sn[s0 + blockIdx.x]];
  const int w = 3 * s.ncol, m = w + 3 * s.nstr;
  const double2* Pn = sa.L + b * sa.PE + s.poff;
  double2* y = sa.X + b * sa.ldx + 3LL * s.col0;
  __shared__ double2 yb[64];
  const int tid = threadIdx.x;
  for (int jb0 = 0; jb0 < w; jb0 += 64) {
    const int nb = min(64, w - jb0);
    if (tid < 64) {
      const int lane = tid;
      double2 yi = lane < nb ? y[jb0 + lane] : make_double2(0.0, 0.0);
      for (int j = 0; j < nb; ++j) {
        double2 yj = make_double2(shfl_d(yi.x, j), shfl_d(yi.y, j));
        if (lane > j && lane < nb) yi = csub(yi, cmul(Pn[(jb0 + lane) + (long long)(jb0 + j) * m], yj));
      }
      if (lane < nb) {
        y[jb0 + lane] = yi;
        yb[lane] = yi;
      }
    }
    __syncthreads();
    for (int i = jb0 + nb + tid; i < w; i += 256) {
      double2 acc = make_double2(0.0, 0.0);
      for (int j = 0; j < nb; ++j) acc = cadd(acc, cmul(Pn[i + (long long)(jb0 + j) * m], yb[j]));
      y[i] = csub(y[i], acc);
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_fwd_update(SolveArgs sa, const RowTask* __restrict__ tasks, int t0) {
  const int b = blockIdx.y;
  if (sa.active && !sa.active[b]) return;
  const RowTask t = tasks[t0 + blockIdx.x];
  const DevSn s = sa.sn[t.sn];
  const int w = 3 * s.ncol, m = w + 3 * s.nstr;
  const double2* Pn = sa.L + b * sa.PE + s.poff;
  const double2* y = sa.X + b * sa.ldx + 3LL * s.col0;
  double2* ts = sa.T + b * sa.TE + s.toff;
  __shared__ double2 ys[512];
  const int tid = threadIdx.x;
  const int q = t.r0 + tid;
  double2 acc = make_double2(0.0, 0.0);
  for (int c0 = 0; c0 < w; c0 += 512) {
    const int nc = min(512, w - c0);
    __syncthreads();
    for (int j = tid; j < nc; j += 256) ys[j] = y[c0 + j];
    __syncthreads();
    if (q < t.r1)
      for (int j = 0; j < nc; ++j) acc = cadd(acc, cmul(Pn[(w + q) + (long long)(c0 + j) * m], ys[j]));
  }
  if (q < t.r1) ts[q] = csub(ts[q], acc);
}

// x_j = y_j / d_j - sum_q conj(L[w+q, j]) x[struct(q)]  (one wave per pivot column)
__global__ __launch_bounds__(256) void k_bwd_update(SolveArgs sa, const RowTask* __restrict__ tasks, int t0) {
  const int b = blockIdx.y;
  if (sa.active && !sa.active[b]) return;
  const RowTask t = tasks[t0 + blockIdx.x];
  const DevSn s = sa.sn[t.sn];
  const int w = 3 * s.ncol, r = 3 * s.nstr, m = w + r;
  const double2* Pn = sa.L + b * sa.PE + s.poff;
  double2* xb = sa.X + b * sa.ldx;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int j = t.r0 + wave;
  if (j >= t.r1) return;
  double re = 0.0, im = 0.0;
  const double2* col = Pn + (long long)j * m + w;
  for (int q = lane; q < r; q += 64) {
    const int node = sa.struct_nodes[s.str0 + q / 3];
    const double2 xv = xb[3LL * node + q % 3];
    const double2 l = col[q];
    double2 u = cconjmul(l, xv);
    re += u.x;
    im += u.y;
  }
  re = wave_sum(re);
  im = wave_sum(im);
  if (lane == 0) {
    double2* xj = xb + 3LL * s.col0 + j;
    const double d = Pn[j + (long long)j * m].x;
    *xj = make_double2(xj->x / d - re, xj->y / d - im);
  }
}

__global__ __launch_bounds__(256) void k_bwd_trsv(SolveArgs sa, const int* __restrict__ level_sn, int s0) {
  const int b = blockIdx.y;
  if (sa.active && !sa.active[b]) return;
  const DevSn s = sa.sn[level_sn[s0 + blockIdx.x]];
  const int w = 3 * s.ncol, m = w + 3 * s.nstr;
  const double2* Pn = sa.L + b * sa.PE + s.poff;
  double2* x = sa.X + b * sa.ldx + 3LL * s.col0;
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int nblk = (w + 63) / 64;
  for (int bi = nblk - 1; bi >= 0; --bi) {
    const int jb0 = bi * 64, nb = min(64, w - jb0), ib0 = jb0 + nb;
    // subtract contributions of the already solved rows below the block
    for (int jj = wave; jj < nb; jj += 4) {
      const int j = jb0 + jj;
      double re = 0.0, im = 0.0;
      for (int i = ib0 + lane; i < w; i += 64) {
        double2 u = cconjmul(Pn[i + (long long)j * m], x[i]);
        re += u.x;
        im += u.y;
      }
      re = wave_sum(re);
      im = wave_sum(im);
      if (lane == 0) x[j] = make_double2(x[j].x - re, x[j].y - im);
    }
    __syncthreads();
    if (wave == 0) {
      double2 v = lane < nb ? x[jb0 + lane] : make_double2(0.0, 0.0);
      for (int j = nb - 1; j >= 0; --j) {
        double2 xj = make_double2(shfl_d(v.x, j), shfl_d(v.y, j));
        if (lane < j) v = csub(v, cconjmul(Pn[(jb0 + j) + (long long)(jb0 + lane) * m], xj));
      }
      if (lane < nb) x[jb0 + lane] = v;
    }
    __syncthreads();
  }
}

}  // namespace

static FactorArgs factor_args(const Mesh& M, const Plan& P, const OpArrays& ops, const BatchBufs& bb) {
  FactorArgs fa;
  fa.sn = P.d_sn.as<DevSn>();
  fa.struct_nodes = M.d_struct.as<int>();
  fa.relmap = M.d_relmap.as<int>();
  fa.asm_slot = M.d_asm_slot.as<int>();
  fa.asm_pos = M.d_asm_pos.as<int>();
  fa.tslot = M.d_tslot.as<int>();
  fa.pinned = M.d_pinned.as<unsigned char>();
  fa.op = ops;
  fa.L = bb.L;
  fa.U = bb.U;
  fa.PE = P.panel_entries;
  fa.UE = P.upool_entries;
  fa.status = bb.status;
  fa.params = bb.params;
  return fa;
}

void factor_batch(Context* ctx, const Mesh& M, const Plan& P, const OpArrays& ops,
                  const BatchBufs& bb) {
  FactorArgs fa = factor_args(M, P, ops, bb);
  const int B = bb.B;
  KTimer t(ctx, "factor", 0, M.S.flops * B);
  for (const LevelPlan& L : P.levels) {
    if (L.nasm)
      hipLaunchKernelGGL(k_front_assemble, dim3(L.nasm, B), dim3(256), 0, ctx->stream, fa,
                         P.d_asm.as<AsmTask>(), L.asm0);
    for (const FactorStep& st : L.steps) {
      if (st.ndiag)
        hipLaunchKernelGGL(k_diag_ldl, dim3(st.ndiag, B), dim3(256), 0, ctx->stream, fa,
                           P.d_diag.as<int>(), st.diag0, st.kb);
      if (st.ntrsm)
        hipLaunchKernelGGL(k_panel_trsm, dim3(st.ntrsm, B), dim3(256), 0, ctx->stream, fa,
                           P.d_row.as<RowTask>(), st.trsm0, st.kb);
      if (st.ntile)
        hipLaunchKernelGGL(k_schur_tile, dim3(st.ntile, B), dim3(256), 0, ctx->stream, fa,
                           P.d_tile.as<TileTask>(), st.tile0, st.kb);
    }
  }
  TM_HIP(hipGetLastError());
}

void trisolve_batch(Context* ctx, const Mesh& M, const Plan& P, const BatchBufs& bb, double2* X,
                    long long ldx, const int* active) {
  SolveArgs sa;
  sa.sn = P.d_sn.as<DevSn>();
  sa.struct_nodes = M.d_struct.as<int>();
  sa.relmap = M.d_relmap.as<int>();
  sa.L = bb.L;
  sa.T = bb.T;
  sa.PE = P.panel_entries;
  sa.TE = P.tvec_entries;
  sa.X = X;
  sa.ldx = ldx;
  sa.active = active;
  const int B = bb.B;
  KTimer t(ctx, "trisolve", P.trisolve_bytes_c16 * B, 0);
  const int* lsn = P.d_level_sn.as<int>();
  for (const LevelPlan& L : P.levels) {
    hipLaunchKernelGGL(k_fwd_asm, dim3(L.nsn, B), dim3(256), 0, ctx->stream, sa, lsn, L.sn0);
    hipLaunchKernelGGL(k_fwd_trsv, dim3(L.nsn, B), dim3(256), 0, ctx->stream, sa, lsn, L.sn0);
    if (L.nfup)
      hipLaunchKernelGGL(k_fwd_update, dim3(L.nfup, B), dim3(256), 0, ctx->stream, sa,
                         P.d_fup.as<RowTask>(), L.fup0);
  }
  for (int h = static_cast<int>(P.levels.size()) - 1; h >= 0; --h) {
    const LevelPlan& L = P.levels[h];
    if (L.nbup)
      hipLaunchKernelGGL(k_bwd_update, dim3(L.nbup, B), dim3(256), 0, ctx->stream, sa,
                         P.d_bup.as<RowTask>(), L.bup0);
    hipLaunchKernelGGL(k_bwd_trsv, dim3(L.nsn, B), dim3(256), 0, ctx->stream, sa, lsn, L.sn0);
  }
  TM_HIP(hipGetLastError());
}

}  // namespace tme
