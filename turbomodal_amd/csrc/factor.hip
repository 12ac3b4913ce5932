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
//   k_front_assemble  each entry of the pivot columns written once with the
//                     children's update entries that land on it (pulled through
//                     inverse relmaps, children in fixed order -> deterministic),
//                     then the A(k) blocks added (formed on the fly from the
//                     phase-split operator arrays: K_c + s dK_c - sigma M_c +
//                     P(.)_p + conj(P)(.)_p^T)
//   k_diag_ldl        unblocked LDL^H of a 32-wide diagonal block in LDS
//   k_panel_trsm      L21 = F21 L11^-H D^-1, one row per thread
//   k_schur_tile      trailing update F22 -= L21 D L21^H on 64x64 tiles; the
//                     first trailing pass of a front writes its update matrix
//                     (which A(k) does not touch) from the children's update
//                     entries, pulled as the assembly pulls them, minus the
//                     product, so the update matrix is written once and never
//                     re-read before its second pass
//
// Triangular solves reuse the same tree: forward (leaves -> root) with update
// vectors passed to the parent, diagonal scaling fused into the backward
// sweep, backward (root -> leaves) gathering already-solved ancestor values.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <shared_mutex>
#include <tuple>
#include <type_traits>

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
// the same for the complex-FP32 factor
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// Entry of a real factor (a real harmonic, k = 0 or N/2 with N even, whose A(k) is
// real symmetric, as the reference's real path: src/cyclic_solver.cpp:673-723): 8 bytes,
// read by the same kernels as a complex entry.  Its imaginary part `y` takes no
// storage and reads as zero; what is written to it is dropped, which is exact here:
// every value the kernels write there is the zero imaginary part of a product or sum
// of real entries (the phase of a real harmonic is exactly +-1).
struct ZeroIm {
  __host__ __device__ operator double() const { return 0.0; }
  __host__ __device__ ZeroIm& operator=(double) { return *this; }
};
struct rdouble {
  double x;
  [[no_unique_address]] ZeroIm y;
};
static_assert(sizeof(rdouble) == sizeof(double), "a real factor entry is one double");
__device__ __forceinline__ rdouble cmul(rdouble a, rdouble b) { return {a.x * b.x, {}}; }
__device__ __forceinline__ rdouble cmulc(rdouble a, rdouble b) { return {a.x * b.x, {}}; }
__device__ __forceinline__ rdouble cadd(rdouble a, rdouble b) { return {a.x + b.x, {}}; }
__device__ __forceinline__ rdouble csub(rdouble a, rdouble b) { return {a.x - b.x, {}}; }

// element type of a factor: complex FP64 (double2), complex FP32 (float2) or real FP64
template <class CT> struct real_of;
template <> struct real_of<double2> { using type = double; };
template <> struct real_of<float2> { using type = float; };
template <> struct real_of<rdouble> { using type = double; };
template <class CT> using real_t = typename real_of<CT>::type;
template <class CT> constexpr bool is_real_elem = false;
template <> constexpr bool is_real_elem<rdouble> = true;
template <class CT> __device__ __forceinline__ CT mk(double re, double im);
template <> __device__ __forceinline__ double2 mk<double2>(double re, double im) { return make_double2(re, im); }
template <> __device__ __forceinline__ float2 mk<float2>(double re, double im) {
  return make_float2((float)re, (float)im);
}
template <> __device__ __forceinline__ rdouble mk<rdouble>(double re, double) { return {re, {}}; }
// a factor entry as complex FP64 (the triangular sweeps accumulate in FP64)
__device__ __forceinline__ double2 ldc(double2 v) { return v; }
__device__ __forceinline__ double2 ldc(float2 v) { return make_double2(v.x, v.y); }
__device__ __forceinline__ double2 ldc(rdouble v) { return make_double2(v.x, 0.0); }
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

// A factor entry as the triangular sweeps use it: complex (double2) for complex
// factors, real (double) for a real factor, whose products with the complex vectors
// then take two FMAs instead of four and whose staged tiles half the LDS.  Leaving out
// the zero imaginary terms changes no result but the sign of an exact zero.
template <class LT> struct sweep_entry { using type = double2; };
template <> struct sweep_entry<rdouble> { using type = double; };
template <class LT> using se_t = typename sweep_entry<LT>::type;
__device__ __forceinline__ double2 lde(double2 v) { return v; }
__device__ __forceinline__ double2 lde(float2 v) { return make_double2(v.x, v.y); }
__device__ __forceinline__ double lde(rdouble v) { return v.x; }
template <class E> __device__ __forceinline__ E eunit(bool one);
template <> __device__ __forceinline__ double2 eunit<double2>(bool one) { return make_double2(one ? 1.0 : 0.0, 0.0); }
template <> __device__ __forceinline__ double eunit<double>(bool one) { return one ? 1.0 : 0.0; }
// acc += a b and acc += conj(a) b
__device__ __forceinline__ void emac(double2& acc, double2 a, double2 b) { cmac(acc, a, b); }
__device__ __forceinline__ void emac(double2& acc, double a, double2 b) {
  acc.x = fma(a, b.x, acc.x);
  acc.y = fma(a, b.y, acc.y);
}
__device__ __forceinline__ void emacc(double2& acc, double2 a, double2 b) { cmacc(acc, a, b); }
__device__ __forceinline__ void emacc(double2& acc, double a, double2 b) { emac(acc, a, b); }
__device__ __forceinline__ void cmac(float2& acc, float2 a, float2 b) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.x = fmaf(-a.y, b.y, acc.x);
  acc.y = fmaf(a.x, b.y, acc.y);
  acc.y = fmaf(a.y, b.x, acc.y);
}
__device__ __forceinline__ void cmac(rdouble& acc, rdouble a, rdouble b) { acc.x = fma(a.x, b.x, acc.x); }

// front element (i, j), i >= j: panel for pivot columns, update matrix otherwise
template <class CT>
__device__ __forceinline__ CT* front_at(CT* Pn, CT* Un, int m, int w, int r, int i, int j) {
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
  const int* invmap;
  const int* asm_slot;
  const int* asm_pos;
  const int* tslot;
  const unsigned char* pinned;
  OpArrays op;
  void* L;  // factor panels and update matrices: double2 or float2 (the kernels' CT)
  void* U;
  long long PE, UE;
  int* status;
  int* pivinfo;           // per problem: negative pivots, smallest |d_j| / |a_jj| (float bits); may be null
  const int* diag_slot;   // per free node: slot of its diagonal block
  const ProbParam* params;
  const int* active;  // per problem: factor it (null: all)
  const OpArrays* ops_b;  // per problem operator arrays (null: op for all)
  const int* lead;    // per problem: whose storage holds the shared fronts (null: its own)
};

// Diagonal entry (i, i) of A(k) = K + s dK - sigma M (+ the phase parts) at DOF comp of
// free node I, before elimination (1 at a pinned DOF), and |sigma| times that of M(k)
// (0 at a pinned DOF).  On the diagonal block the transposed slot is the slot itself,
// so the phase parts add 2 Re(P) times the real diagonal of the coupling block.
__device__ __forceinline__ double2 diag_entry(const FactorArgs& fa, const OpArrays& op, const ProbParam& pp,
                                              int node, int comp) {
  if (fa.pinned[3 * node + comp]) return make_double2(1.0, 0.0);
  const long long e = (long long)fa.diag_slot[node] * 9 + 4 * comp;
  const double m = op.Mc[e] + 2.0 * pp.pre * op.Mp[e];
  const double c = op.Kc[e] + pp.s * op.dKc[e] + 2.0 * pp.pre * (op.Kp[e] + pp.s * op.dKp[e]);
  return make_double2(c - pp.sigma * m, fabs(pp.sigma * m));
}

// The children's update matrices of front s in problem b, as a front entry pulls
// them: for child k (child0 first) the inverse relmap (front node -> the child's
// update node, or -1), the matrix and its order.
template <class CT>
struct ChildUpdates {
  const int* inv[2] = {nullptr, nullptr};
  const CT* U[2] = {nullptr, nullptr};
  int rc[2] = {0, 0};
  __device__ ChildUpdates() = default;
  __device__ ChildUpdates(const FactorArgs& fa, const DevSn& s, int b) {
    int o = s.inv0;
    for (int k = 0; k < 2; ++k) {
      const int c = k == 0 ? s.child0 : s.child1;
      if (c < 0) continue;
      const DevSn cs = fa.sn[c];
      inv[k] = fa.invmap + o;
      o += s.ncol + s.nstr;
      // a shared child's update matrix is in its class leader's storage
      const int cb = cs.shared && fa.lead ? fa.lead[b] : b;
      U[k] = static_cast<const CT*>(fa.U) + cb * fa.UE + cs.uoff;
      rc[k] = 3 * cs.nstr;
    }
  }
  // child k's node for front node `node`, or -1
  __device__ __forceinline__ int node_of(int k, int node) const { return inv[k] ? inv[k][node] : -1; }
  // child k's update entry at front (row, column) whose nodes map to qi, qj (both >= 0)
  __device__ __forceinline__ CT at(int k, int qi, int qj, int i, int j) const {
    return U[k][(3 * qi + i % 3) + (long long)(3 * qj + j % 3) * rc[k]];
  }
};

// Pivot columns only (the assembly tasks cover the front's pivot nodes); the update
// matrix is written by the first trailing Schur pass.
template <class CT>
__global__ __launch_bounds__(256) void k_front_assemble(FactorArgs fa, const AsmTask* __restrict__ tasks,
                                                        int t0) {
  const AsmTask t = tasks[t0 + blockIdx.x];
  const int b = blockIdx.y;
  if (fa.active && !fa.active[b]) return;
  const DevSn s = fa.sn[t.sn];
  const int w = 3 * s.ncol, r = 3 * s.nstr, m = w + r;
  CT* Pn = static_cast<CT*>(fa.L) + b * fa.PE + s.poff;
  CT* Un = static_cast<CT*>(fa.U) + b * fa.UE + s.uoff;
  const int tid = threadIdx.x;
  const int j0 = 3 * t.n0, j1 = 3 * t.n1;  // pivot columns: j1 <= w
  // children's update entries pulled through the inverse relmaps: every lower
  // entry of columns [j0, j1) is written once, child0's term first
  const ChildUpdates<CT> cu(fa, s, b);
  // wave v takes columns j0 + v, j0 + v + 4, ...: four columns in flight per block
  const int lane = tid & 63;
  for (int j = j0 + (tid >> 6); j < j1; j += 4) {
    int qj[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) qj[k] = cu.node_of(k, j / 3);
    CT* col = Pn + (long long)j * m;
    for (int i = j + lane; i < m; i += 64) {
      CT v = mk<CT>(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        if (qj[k] < 0) continue;
        const int qi = cu.node_of(k, i / 3);
        if (qi >= 0) v = cadd(v, cu.at(k, qi, qj[k], i, j));
      }
      col[i] = v;
    }
  }
  __syncthreads();
  // A(k) entries added (pinned DOFs: identity rows and columns)
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
    CT* dst = front_at(Pn, Un, m, w, r, gi, gj);
    if (fa.pinned[3 * rnode + ci] || fa.pinned[3 * cnode + cj]) {
      *dst = mk<CT>(gi == gj ? 1.0 : 0.0, 0.0);
    } else {
      const long long sl = fa.asm_slot[q];
      const long long tt = fa.tslot[sl];
      const long long e1 = sl * 9 + 3 * ci + cj;
      const long long e2 = tt * 9 + 3 * cj + ci;
      const OpArrays& op = fa.ops_b ? fa.ops_b[b] : fa.op;
      double c = op.Kc[e1] + pp.s * op.dKc[e1] - pp.sigma * op.Mc[e1];
      double p = op.Kp[e1] + pp.s * op.dKp[e1] - pp.sigma * op.Mp[e1];
      double qv = op.Kp[e2] + pp.s * op.dKp[e2] - pp.sigma * op.Mp[e2];
      *dst = cadd(*dst, mk<CT>(c + pp.pre * (p + qv), pp.pim * (p - qv)));
    }
  }
}

// pivot statistics of the problems about to be factored: no negative pivot yet,
// smallest relative pivot +inf
__global__ void k_pivinfo_init(int* pivinfo, const int* active, int B) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B || (active && !active[b])) return;
  pivinfo[2 * b] = 0;
  pivinfo[2 * b + 1] = (int)__float_as_uint(INFINITY);
}

// Shared fronts: ws = [leader mask B][status B][pivot statistics 2 B].  The problems
// about to be factored that lead their class are marked, and the statistics of their
// shared fronts cleared.
__global__ void k_share_prep(const int* __restrict__ lead, const int* __restrict__ active, int B, int* ws) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  ws[b] = (!active || active[b]) && lead[b] == b ? 1 : 0;
  ws[B + b] = 0;
  ws[2 * B + 2 * b] = 0;
  ws[2 * B + 2 * b + 1] = (int)__float_as_uint(INFINITY);
}

// Every factored problem receives its class's shared-front statistics: status bits,
// negative pivots, smallest relative pivot (non-negative floats order as their bit
// patterns).  The shared fronts' panels stay in the leader's storage, where the
// triangular sweeps of every member read them.
__global__ void k_share_finish(const int* __restrict__ lead, const int* __restrict__ active, int B,
                               int* __restrict__ status, int* __restrict__ pivinfo, const int* __restrict__ ws) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B || (active && !active[b])) return;
  const int lb = lead[b];
  status[b] |= ws[B + lb];
  if (pivinfo) {
    pivinfo[2 * b] += ws[2 * B + 2 * lb];
    pivinfo[2 * b + 1] = (int)min((unsigned)pivinfo[2 * b + 1], (unsigned)ws[2 * B + 2 * lb + 1]);
  }
}

template <class CT>
__global__ __launch_bounds__(256) void k_diag_ldl(FactorArgs fa, const int* __restrict__ tasks, int t0,
                                                  int kb) {
  const int id = tasks[t0 + blockIdx.x];
  const int b = blockIdx.y;
  if (fa.active && !fa.active[b]) return;
  const DevSn s = fa.sn[id];
  const int w = 3 * s.ncol, m = w + 3 * s.nstr;
  const int jb = min(NBMAX, w - kb);
  CT* Pn = static_cast<CT*>(fa.L) + b * fa.PE + s.poff;
  using RT = real_t<CT>;
  __shared__ CT A[NBMAX][NBMAX + 1];
  const int tid = threadIdx.x;
  // thread (tx, ty) of a 32 x 8 grid takes row tx of columns ty, ty + 8, ...
  constexpr int NY = 256 / NBMAX;
  const int tx = tid % NBMAX, ty = tid / NBMAX;
  // the block's loads issued together, ahead of the diagonal-entry lookups below, and
  // stored to LDS after them, so that their latencies overlap
  CT ra[NBMAX / NY];
#pragma unroll
  for (int u = 0; u < NBMAX / NY; ++u) {
    const int j = ty + NY * u;
    ra[u] = j < jb && tx >= j && tx < jb ? Pn[(kb + tx) + (long long)(kb + j) * m] : mk<CT>(0.0, 0.0);
  }
  // the diagonals of A(k) and of sigma M(k) at these columns before elimination, for the
  // pivot-size test
  __shared__ double2 a0[NBMAX];
  if (tid < jb) {
    const int g = kb + tid;
    a0[tid] = diag_entry(fa, fa.ops_b ? fa.ops_b[b] : fa.op, fa.params[b], s.col0 + g / 3, g % 3);
  }
#pragma unroll
  for (int u = 0; u < NBMAX / NY; ++u) {
    const int j = ty + NY * u;
    if (j < jb && tx >= j && tx < jb) A[tx][j] = ra[u];
  }
  __syncthreads();
  // Step j: L[i][j] = A[i][j] / d_j below the pivot, and the trailing lower triangle
  // less L[i][j] d_j conj(L[k][j]).  Each thread scales the two column entries it uses
  // itself (the same product as the stored one), so a step needs one barrier: column j
  // is read and never written during step j, and its scaled values are stored during
  // step j + 1, when no thread reads it any more.
  RT inv_prev = RT(0);
  for (int j = 0; j < jb; ++j) {
    const RT d = A[j][j].x;
    const RT inv = RT(1) / d;
    if (j > 0 && tid >= j && tid < jb)
      A[tid][j - 1] = mk<CT>(A[tid][j - 1].x * inv_prev, A[tid][j - 1].y * inv_prev);
    const int ii = j + 1 + tx;  // entries (ii, kk), j < kk <= ii < jb, kk - j - 1 = ty mod NY
    if (ii < jb) {
      const CT l = mk<CT>(A[ii][j].x * inv, A[ii][j].y * inv);
      for (int kk = j + 1 + ty; kk <= ii; kk += NY) {
        const CT lk = mk<CT>(A[kk][j].x * inv, A[kk][j].y * inv);
        CT u = cmulc(l, lk);
        A[ii][kk] = mk<CT>(A[ii][kk].x - u.x * d, A[ii][kk].y - u.y * d);
      }
    }
    inv_prev = inv;
    __syncthreads();
  }
  // pivot statistics, one lane of the last wave per pivot (its columns of the inverse
  // below are the shortest; d_j = A[j][j], which no step after j changes): zero or
  // non-finite pivots, the negative-pivot count, the smallest pivot relative to its
  // entry before elimination, and near-singularity: the pivot has all but vanished
  // (relative to the entry it came from) and what cancelled it is the shift (the pivot
  // is small against sigma M_jj too): the shift sits on an eigenvalue.  A pivot that is
  // small only against K_jj, as sigma = 1 leaves above the rigid-body modes of a free
  // structure, is not.  (Lanes past jb: none of these; the first shuffle step pairs
  // every lane below 32 with one that holds +inf, so the minimum skips NaN ratios as a
  // sequential fminf would.)
  constexpr int kStatLane0 = 256 - 64;
  const int pl = tid - kStatLane0;
  int nneg = 0, tiny = 0, bad = 0;
  float rmin = INFINITY;
  if (pl >= 0) {
    const bool on = pl < jb;
    const RT d = on ? A[pl][pl].x : RT(1);
    const double d0 = on ? a0[pl].x : 1.0, s0 = on ? a0[pl].y : 1.0;
    bad = __ballot(on && (d == 0.0 || !isfinite(d))) != 0;
    nneg = __popcll(__ballot(on && d < 0.0));
    tiny = __ballot(on && fabs((double)d) <= kTinyPivot * fabs(d0) && fabs((double)d) <= kTinyShiftPivot * s0) != 0;
    rmin = on ? (float)(fabs((double)d) / fabs(d0)) : INFINITY;
    for (int o = 32; o > 0; o >>= 1) rmin = fminf(rmin, __shfl_xor(rmin, o, 64));
  }
  // the inverse of the unit lower block, column c by forward substitution, kept
  // transposed in the block's strictly upper triangle (no other kernel reads it)
  // for the panel solves of this step
  // Eight lanes per column (a group of eight consecutive lanes of one wave): each
  // row's dot product split over the group and summed by shuffles within it.  A
  // group's LDS writes precede its next reads (one wave, in order; the waits keep
  // the compiler from moving them).
  __shared__ CT Li[NBMAX][NBMAX + 1];
  {
    const int c = tid >> 3, l8 = tid & 7;
    if (c < jb) {
      if (l8 == 0) Li[c][c] = mk<CT>(1.0, 0.0);
      __builtin_amdgcn_s_waitcnt(0xc07f);
      __builtin_amdgcn_wave_barrier();
      for (int i = c + 1; i < jb; ++i) {
        CT v = mk<CT>(0.0, 0.0);
        for (int k = c + l8; k < i; k += 8) v = csub(v, cmul(A[i][k], Li[k][c]));
#pragma unroll
        for (int o = 1; o < 8; o <<= 1) {
          v.x += __shfl_xor(v.x, o, 64);
          if constexpr (!is_real_elem<CT>) v.y += __shfl_xor(v.y, o, 64);
        }
        if (l8 == 0) Li[i][c] = v;
        __builtin_amdgcn_s_waitcnt(0xc07f);
        __builtin_amdgcn_wave_barrier();
      }
    }
  }
  __syncthreads();
  if (pl == 0) {
    if (bad) atomicOr(&fa.status[b], 1);
    if (tiny) atomicOr(&fa.status[b], 2);
    if (fa.pivinfo) {
      if (nneg) atomicAdd(&fa.pivinfo[2 * b], nneg);
      // non-negative floats order as their bit patterns
      atomicMin(reinterpret_cast<unsigned int*>(&fa.pivinfo[2 * b + 1]), __float_as_uint(rmin));
    }
  }
  for (int j = ty; j < jb; j += NY) {
    const int i = tx;
    if (i > j && i < jb) {
      Pn[(kb + i) + (long long)(kb + j) * m] = A[i][j];
      Pn[(kb + j) + (long long)(kb + i) * m] = Li[i][j];
    } else if (i == j) {
      Pn[(kb + i) + (long long)(kb + j) * m] = mk<CT>(A[i][i].x, 0.0);
    }
  }
}

template <class CT>
__global__ __launch_bounds__(256) void k_panel_trsm(FactorArgs fa, const RowTask* __restrict__ tasks,
                                                    int t0, int kb) {
  const RowTask t = tasks[t0 + blockIdx.x];
  const int b = blockIdx.y;
  if (fa.active && !fa.active[b]) return;
  const DevSn s = fa.sn[t.sn];
  const int w = 3 * s.ncol, m = w + 3 * s.nstr;
  const int jb = min(NBMAX, w - kb);
  CT* Pn = static_cast<CT*>(fa.L) + b * fa.PE + s.poff;
  using RT = real_t<CT>;
  // Li = inverse of the unit lower diagonal block (zero above the diagonal and
  // outside jb x jb), so that x = b L11^-H is a product with static indices;
  // k_diag_ldl left it transposed in the block's strictly upper triangle
  __shared__ CT Li[NBMAX][NBMAX + 1];
  __shared__ RT dk[NBMAX];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < NBMAX * NBMAX; idx += 256) {
    const int i = idx / NBMAX, j = idx % NBMAX;  // row i of Li is a contiguous piece of column kb + i
    Li[i][j] = (i > j && i < jb) ? Pn[(kb + j) + (long long)(kb + i) * m]
                                 : mk<CT>(i == j && i < jb ? 1.0 : 0.0, 0.0);
    if (i == j && i < jb) dk[i] = Pn[(kb + i) + (long long)(kb + i) * m].x;
  }
  __syncthreads();
  const int i = t.r0 + tid;
  if (i >= t.r1) return;
  CT x[NBMAX];
#pragma unroll
  for (int k = 0; k < NBMAX; ++k) x[k] = k < jb ? Pn[i + (long long)(kb + k) * m] : mk<CT>(0.0, 0.0);
  // v = sum_{k <= j} x_k conj(Li_jk): the terms above the diagonal of Li are exact
  // zeros, and adding their products changed nothing (the sum starts at +0 and never
  // becomes -0, and v + (+-0) = v otherwise), so the sums stop at the diagonal
#pragma unroll
  for (int j = 0; j < NBMAX; ++j) {
    if (j < jb) {  // uniform; a constant trip count keeps x[] in registers
      CT v = mk<CT>(0.0, 0.0);
#pragma unroll
      for (int k = 0; k <= j; ++k) {  // v += x_k conj(L_jk)
        v.x = fma(x[k].x, Li[j][k].x, v.x);
        if constexpr (!is_real_elem<CT>) {
          v.x = fma(x[k].y, Li[j][k].y, v.x);
          v.y = fma(x[k].y, Li[j][k].x, v.y);
          v.y = fma(-x[k].x, Li[j][k].y, v.y);
        }
      }
      const RT inv = RT(1) / dk[j];
      Pn[i + (long long)(kb + j) * m] = mk<CT>(v.x * inv, v.y * inv);
    }
  }
}

// F[i][j] -= sum_{k in [tk, tk+kd)} L[i][k] d_k conj(L[j][k]) on a 64x64 tile,
// source columns staged through LDS (one read-modify-write of the
// tile per super-panel instead of one per staged chunk).  The products run on the
// FP64 matrix cores: each wave owns a 32x32 quarter of the tile as 2x2 blocks of
// 16x16, and every complex block product is four real v_mfma_f64_16x16x4
// (re re + im im into the real part, im re - re im into the imaginary part).
// Operand layout of v_mfma_f64_16x16x4: lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16]; result register j of lane l is C[4 j + l / 16][l % 16].
// Columns are staged KC at a time (KC = 16: 34 KB of LDS per block for complex FP64,
// so three to four blocks share a CU and one's staging overlaps another's products);
// every KC accumulates the same products in the same order (k ascending, four at a
// time), so the variants give bit-identical results.
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
// one 16x16x4 real product on the matrix cores, FP64 or FP32 (FP32 accumulation)
template <class RT> struct v4of;
template <> struct v4of<double> { using type = v4d; };
template <> struct v4of<float> { using type = v4f; };
__device__ __forceinline__ v4d mfma16(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma16(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
// row of result register j of lane l in a 16x16 block (column l % 16): measured on
// gfx950 with an in-bounds probe, v_mfma_f64_16x16x4 puts C[4 j + l / 16] there and
// v_mfma_f32_16x16x4 C[4 (l / 16) + j]
__device__ __forceinline__ int mfma_row(double, int j, int l4) { return 4 * j + l4; }
__device__ __forceinline__ int mfma_row(float, int j, int l4) { return 4 * l4 + j; }

// Work-groups are dealt round-robin over the eight XCDs (each with its own L2):
// in a row of n blocks, the blocks with the same x mod 8 share an XCD.  Block x
// takes task xcd_task(x, n), so every XCD runs a contiguous run of the task list,
// whose neighbouring tasks share their source panels.
__device__ __forceinline__ int xcd_task(int x, int n) {
  const int c = x & 7, q = n >> 3, rem = n & 7;
  return c * q + min(c, rem) + (x >> 3);
}

template <class CT, int KC, bool PULL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_schur_tile(FactorArgs fa, const TileTask* __restrict__ tasks,
                                                    int t0, int kb) {
  (void)kb;
  const TileTask t = tasks[t0 + xcd_task(blockIdx.x, gridDim.x)];
  const int b = blockIdx.y;
  if (fa.active && !fa.active[b]) return;
  const DevSn s = fa.sn[t.sn];
  const int w = 3 * s.ncol, r = 3 * s.nstr, m = w + r;
  const int jlim = t.jend < 0 ? m : min(t.jend, w);
  using RT = real_t<CT>;
  using V4 = typename v4of<RT>::type;
  CT* Pn = static_cast<CT*>(fa.L) + b * fa.PE + s.poff;
  CT* Un = static_cast<CT*>(fa.U) + b * fa.UE + s.uoff;
  __shared__ CT Li[TILE][KC + 1];
  __shared__ CT Lj[TILE][KC + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int r0 = 32 * (wv & 1), q0 = 32 * (wv >> 1);  // this wave's rows / columns in the tile
  const int l16 = lane & 15, l4 = lane >> 4;
  V4 cr[2][2], ci[2][2];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      cr[p][q] = V4{0, 0, 0, 0};
      ci[p][q] = V4{0, 0, 0, 0};
    }
  // 16 x 16 blocks of this wave whose result is written: not wholly above the
  // diagonal, below row m, or right of column jlim (uniform over the wave)
  bool need[2][2];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int gi1 = t.i0 + r0 + 16 * p + 15, gi0 = t.i0 + r0 + 16 * p, gj0 = t.j0 + q0 + 16 * q;
      need[p][q] = gi1 >= gj0 && gi0 < m && gj0 < jlim;
    }
  // staging of one chunk: this thread's KC * TILE / 256 elements of each
  // panel, loaded one chunk ahead so the loads overlap the previous products
  constexpr int PER = TILE * KC / 256;
  CT ri[PER], rj[PER];
  auto fetch = [&](int c0) {
    const int jb = min(KC, t.tk + t.kd - c0);
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int idx = tid + 256 * e, ii = idx % TILE, k = idx / TILE;
      const int gi = t.i0 + ii, gj = t.j0 + ii;
      const bool in = k < jb;
      const long long col = (long long)(c0 + k) * m;
      ri[e] = in && gi < m ? Pn[gi + col] : mk<CT>(0.0, 0.0);
      rj[e] = in && gj < m ? Pn[gj + col] : mk<CT>(0.0, 0.0);
    }
  };
  __shared__ RT dk[KC];
  fetch(t.tk);
  // The first trailing pass of the front (source columns from 0, no column limit)
  // writes the update-matrix entries (columns >= w) for the first time: they start
  // from the children's update entries, child0's term first, as the assembly pulls
  // them for the pivot columns.  The children's nodes of this lane's rows and
  // columns are looked up here, after the first chunk's loads are issued and ahead of
  // the products (-1: no entry); other tiles do not touch the children.
  // (PULL: the launch holds such tiles; the other launches leave this code out)
  const bool pull = PULL && t.tk == 0 && t.jend < 0 && t.j0 + TILE > w;
  ChildUpdates<CT> cu;
  int cqi[2][2][4], cqj[2][2];
  if (pull) {
    cu = ChildUpdates<CT>(fa, s, b);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int gj = t.j0 + q0 + 16 * q + l16;
        const int v = cu.node_of(k, min(gj, m - 1) / 3);
        cqj[k][q] = gj < m && gj >= w ? v : -1;
      }
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int gi = t.i0 + r0 + 16 * p + mfma_row(RT(0), j, l4);
          const int v = cu.node_of(k, min(gi, m - 1) / 3);
          cqi[k][p][j] = gi < m ? v : -1;
        }
    }
  }
  // Entry (gi, gj) of block (p, q), register j, before this update, in two steps so
  // that the loads of several entries are in flight together: pre_load issues them
  // (no branch: an entry outside the lower triangle or the front, valid false, and
  // an absent child term read the front's first entry and are dropped), pre_value
  // combines.  On the first pass an update-matrix entry is the children's sum
  // (child0's term first; adding +0 to the sum, which starts at +0, changes no bit)
  // and a pivot-column entry the entry itself.
  struct Pre {
    CT a, c;
    bool ok0, ok1;
  };
  auto pre_load = [&](int p, int q, int j, int gi, int gj, bool valid) -> Pre {
    CT* own = valid ? front_at(Pn, Un, m, w, r, gi, gj) : Pn;
    Pre e;
    if (!pull) {
      e.a = *own;
      e.c = e.a;
      e.ok0 = e.ok1 = false;
      return e;
    }
    const CT* src[2];
    bool ok[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      ok[k] = valid && gj >= w && cqj[k][q] >= 0 && cqi[k][p][j] >= 0;
      // an absent term reads the front's first entry (one line, in cache), not the
      // entry's own location: that would be a wasted HBM read per absent term
      src[k] = ok[k] ? &cu.U[k][(3 * cqi[k][p][j] + gi % 3) + (long long)(3 * cqj[k][q] + gj % 3) * cu.rc[k]]
                     : (k == 0 && gj < w ? own : Pn);
    }
    e.a = *src[0];
    e.c = *src[1];
    e.ok0 = ok[0];
    e.ok1 = ok[1];
    return e;
  };
  // keeps the compiler from moving a load of pre_load into a branch of pre_value
  auto pre_hold = [](const Pre& e) {
    asm volatile("" ::"v"(e.a.x), "v"(e.c.x));
    if constexpr (!is_real_elem<CT>) asm volatile("" ::"v"(e.a.y), "v"(e.c.y));
  };
  auto pre_value = [&](const Pre& e, int gj) -> CT {
    if (!pull || gj < w) return e.a;  // the entry itself
    const CT z = mk<CT>(0.0, 0.0);
    return cadd(cadd(z, e.ok0 ? e.a : z), e.ok1 ? e.c : z);
  };
  // A first-pass tile of an FP64 factor starts its accumulators at the entries' values
  // before the update (gathered while the first chunk's loads are in flight) and
  // accumulates the negated products (dk holds -d), so that its epilogue only stores.
  // The complex FP32 factor keeps old - sum(products) with the sum from zero: each
  // accumulation step rounds at the accumulator's magnitude, and starting it at the
  // entry cost the FP32 factor enough accuracy to double the FP64 refinement steps of
  // its solves (1049 -> 2106 sweep pairs per C4 sweep).
  constexpr bool kAccInit = !std::is_same<CT, float2>::value;
  if (kAccInit && pull) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int gj = t.j0 + q0 + 16 * q + l16;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        if (!need[p][q]) continue;
        Pre pre[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int gi = t.i0 + r0 + 16 * p + mfma_row(RT(0), j, l4);
          pre[j] = pre_load(p, q, j, gi, gj, gj < jlim && gi < m && gi >= gj);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) pre_hold(pre[j]);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const CT o = pre_value(pre[j], gj);
          cr[p][q][j] = o.x;
          ci[p][q][j] = o.y;
        }
      }
    }
  }
  for (int c0 = t.tk; c0 < t.tk + t.kd; c0 += KC) {
    const int jb = min(KC, t.tk + t.kd - c0);
    __syncthreads();
    if (tid < KC) {
      const RT dv = tid < jb ? Pn[(c0 + tid) + (long long)(c0 + tid) * m].x : RT(0);
      dk[tid] = kAccInit && pull ? -dv : dv;
    }
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int idx = tid + 256 * e, ii = idx % TILE, k = idx / TILE;
      Li[ii][k] = ri[e];
      Lj[ii][k] = rj[e];
    }
    __syncthreads();
    if (c0 + KC < t.tk + t.kd) fetch(c0 + KC);
    const int kend = (jb + 3) & ~3;  // entries past jb are zero
    // operands of step kk + 4 read from LDS while the products of step kk run
    CT va[2], vb[2];
    RT d;
    auto operands = [&](int kk, CT* oa, CT* ob, RT& od) {
      const int kx = kk + l4;
      od = dk[kx];
#pragma unroll
      for (int p = 0; p < 2; ++p) oa[p] = Li[r0 + 16 * p + l16][kx];
#pragma unroll
      for (int q = 0; q < 2; ++q) ob[q] = Lj[q0 + 16 * q + l16][kx];
    };
    operands(0, va, vb, d);
    for (int kk = 0; kk < kend; kk += 4) {
      // unconditional (the last step re-reads columns of this chunk and drops them):
      // the products then wait only for the loads of the step before
      CT na[2], nb[2];
      RT nd;
      operands(min(kk + 4, KC - 4), na, nb, nd);
      CT a[2];
#pragma unroll
      for (int p = 0; p < 2; ++p) a[p] = mk<CT>(va[p].x * d, va[p].y * d);
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          if (!need[p][q]) continue;
          cr[p][q] = mfma16(a[p].x, vb[q].x, cr[p][q]);
          if constexpr (!is_real_elem<CT>) {
            cr[p][q] = mfma16(a[p].y, vb[q].y, cr[p][q]);
            ci[p][q] = mfma16(a[p].y, vb[q].x, ci[p][q]);
            ci[p][q] = mfma16(-a[p].x, vb[q].y, ci[p][q]);
          }
        }
      if (kk + 4 < kend) {
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          va[p] = na[p];
          vb[p] = nb[p];
        }
        d = nd;
      }
    }
  }
  // Epilogue.
  // * First trailing pass, FP64: the accumulators hold the updated entries; stored.
  //   FP32: the four entries of a 16 x 16 block pulled together (pre_load), combined,
  //   and stored less the products.
  // * Later passes, 8-byte entries (the real factor, the complex FP32 factor): every
  //   old value loaded before the first store (the stores cannot alias the other
  //   entries' loads, but the compiler cannot know that and would otherwise wait on
  //   each load in turn): 5 % faster on the real factorization of the bladed sector.
  // * Later passes, complex FP64: one load, wait and store per entry; the batched
  //   form measured 3.5 % slower there (a burst of 16 loads per lane at the end of
  //   every tile).
  if (kAccInit && pull) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int gj = t.j0 + q0 + 16 * q + l16;
      if (gj >= jlim) continue;
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int gi = t.i0 + r0 + 16 * p + mfma_row(RT(0), j, l4);
          if (gi >= m || gi < gj) continue;
          *front_at(Pn, Un, m, w, r, gi, gj) = mk<CT>(cr[p][q][j], ci[p][q][j]);
        }
    }
  } else if (pull) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int gj = t.j0 + q0 + 16 * q + l16;
      if (gj >= jlim) continue;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        Pre pre[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int gi = t.i0 + r0 + 16 * p + mfma_row(RT(0), j, l4);
          pre[j] = pre_load(p, q, j, gi, gj, gi < m && gi >= gj);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) pre_hold(pre[j]);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int gi = t.i0 + r0 + 16 * p + mfma_row(RT(0), j, l4);
          if (gi >= m || gi < gj) continue;
          const CT o = pre_value(pre[j], gj);
          *front_at(Pn, Un, m, w, r, gi, gj) = mk<CT>(o.x - cr[p][q][j], o.y - ci[p][q][j]);
        }
      }
    }
  } else if constexpr (sizeof(CT) == 8) {
    CT old[2][2][4];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int gj = t.j0 + q0 + 16 * q + l16;
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int gi = t.i0 + r0 + 16 * p + mfma_row(RT(0), j, l4);
          old[p][q][j] = gj < jlim && gi < m && gi >= gj ? *front_at(Pn, Un, m, w, r, gi, gj) : mk<CT>(0.0, 0.0);
        }
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int gj = t.j0 + q0 + 16 * q + l16;
      if (gj >= jlim) continue;
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int gi = t.i0 + r0 + 16 * p + mfma_row(RT(0), j, l4);
          if (gi >= m || gi < gj) continue;
          *front_at(Pn, Un, m, w, r, gi, gj) = mk<CT>(old[p][q][j].x - cr[p][q][j], old[p][q][j].y - ci[p][q][j]);
        }
    }
  } else {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int gj = t.j0 + q0 + 16 * q + l16;
      if (gj >= jlim) continue;
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int gi = t.i0 + r0 + 16 * p + mfma_row(RT(0), j, l4);
          if (gi >= m || gi < gj) continue;
          CT* dst = front_at(Pn, Un, m, w, r, gi, gj);
          *dst = mk<CT>(dst->x - cr[p][q][j], dst->y - ci[p][q][j]);
        }
    }
  }
}

// ---------------------------------------------------------------------------
// in-place inversion of the unit lower L11 of every front (after its
// factorization), 64-column blocks from the last: X_JJ = L_JJ^-1 (row by row
// in LDS), X_IJ = -(sum_{J<K<=I} X_IK L_KJ) X_JJ.  The diagonal keeps D.  Block
// column J is overwritten in place once its L is no longer read: X_JJ by
// k_inv_diag (k_inv_z then reads L_IJ, I > J, and X_JJ from the panel), X_IJ by
// k_inv_tile (which reads only columns K > J of the panel, already inverted, and
// the scratch Z).
// ---------------------------------------------------------------------------
struct InvArgs {
  const DevSn* sn;
  void* L;  // double2 or float2 (the kernels' CT)
  void* S;
  long long PE, SE;
  const int* active;  // per problem (null: all)
};

// One level of the recursive inversion of a unit lower triangle held in A (strict
// lower part; the unit diagonal is implicit and the diagonal entries hold 0): for
// each of the nblk diagonal blocks of size 2h at 2h P, whose two h x h diagonal
// blocks already hold their inverses X11, X22, the off-diagonal block L21 becomes
//   X21 = -X22 (L21 X11)
// in two passes (T = L21 X11 into L21's place, then X21), 256 threads, outputs
// with the column fastest (row reads broadcast, column reads contiguous).
// the 64 x 64 block in LDS, rows padded by one entry so that a column is read
// without bank conflicts
constexpr int kInvLd = 65;
template <int h, int nblk, class CT>
__device__ __forceinline__ void inv_combine(CT (*A)[kInvLd]) {
  constexpr int nout = nblk * h * h, per = (nout + 255) / 256;
  static_assert(nout % 256 == 0, "whole passes of 256 outputs");
  const int tid = threadIdx.x;
  CT v[per];
  // T[i][j] = L21[i][j] + sum_{k>j} L21[i][k] X11[k][j]
#pragma unroll
  for (int u = 0; u < per; ++u) {
    const int e = tid + 256 * u;
    const int j = e % h, i = (e / h) % h, o = 2 * h * (e / (h * h));
    CT t = A[o + h + i][o + j];
    for (int k = j + 1; k < h; ++k) cmac(t, A[o + h + i][o + k], A[o + k][o + j]);
    v[u] = t;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < per; ++u) {
    const int e = tid + 256 * u;
    const int j = e % h, i = (e / h) % h, o = 2 * h * (e / (h * h));
    A[o + h + i][o + j] = v[u];
  }
  __syncthreads();
  // X21[i][j] = -(T[i][j] + sum_{k<i} X22[i][k] T[k][j])
#pragma unroll
  for (int u = 0; u < per; ++u) {
    const int e = tid + 256 * u;
    const int j = e % h, i = (e / h) % h, o = 2 * h * (e / (h * h));
    CT t = A[o + h + i][o + j];
    for (int k = 0; k < i; ++k) cmac(t, A[o + h + i][o + h + k], A[o + h + k][o + j]);
    v[u] = mk<CT>(-t.x, -t.y);
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < per; ++u) {
    const int e = tid + 256 * u;
    const int j = e % h, i = (e / h) % h, o = 2 * h * (e / (h * h));
    A[o + h + i][o + j] = v[u];
  }
  __syncthreads();
}

// Inverse of the unit lower 64-block J of a front's pivot triangle (nb <= 64
// rows; the padding is an identity block), in place in LDS: the four 16 x 16
// diagonal blocks by the row recurrence, one wave each (lanes 0..15 own the
// columns), then two levels of inv_combine; written back to the panel.
template <class CT>
__global__ __launch_bounds__(256) void k_inv_diag(InvArgs ia, const int* __restrict__ tasks, int t0, int J) {
  const int id = tasks[t0 + blockIdx.x];
  const int b = blockIdx.y;
  if (ia.active && !ia.active[b]) return;
  const DevSn s = ia.sn[id];
  const int w = 3 * s.ncol, m = w + 3 * s.nstr;
  const int j0 = 64 * J, nb = min(64, w - j0);
  CT* Pn = static_cast<CT*>(ia.L) + b * ia.PE + s.poff;
  __shared__ CT A[64][kInvLd];
  const int tid = threadIdx.x;
  // rows fastest over the lanes: each column is a contiguous piece of the panel.  All
  // sixteen loads of a thread are issued before the first is stored, so their
  // latencies overlap (the LDS block already limits the kernel to two blocks per CU,
  // so the registers cost no occupancy)
  constexpr int kPer = 64 * 64 / 256;
  CT ra[kPer];
#pragma unroll
  for (int e = 0; e < kPer; ++e) {
    const int idx = tid + 256 * e, i = idx % 64, t = idx / 64;
    ra[e] = (t < i && i < nb) ? Pn[(j0 + i) + (long long)(j0 + t) * m] : mk<CT>(0.0, 0.0);
  }
#pragma unroll
  for (int e = 0; e < kPer; ++e) {
    const int idx = tid + 256 * e;
    A[idx % 64][idx / 64] = ra[e];
  }
  __syncthreads();
  {
    // row i of the 16-block o: X[i][t] = -L[i][t] - sum_{t<k<i} L[i][k] X[k][t]; rows
    // k < i already hold X.  Each wave works in its own block, and a lane writes its
    // entry of row i after its reads of that row (the stored value depends on them), so
    // the wave's own LDS order orders the steps: the waits keep the compiler from moving
    // the next step's reads ahead of this step's stores.
    const int o = 16 * (tid >> 6), t = tid & 63;
    for (int i = 1; i < 16; ++i) {
      if (t < i) {
        CT v = mk<CT>(-A[o + i][o + t].x, -A[o + i][o + t].y);
        for (int k = t + 1; k < i; ++k) v = csub(v, cmul(A[o + i][o + k], A[o + k][o + t]));
        A[o + i][o + t] = v;
      }
      __builtin_amdgcn_s_waitcnt(0xc07f);
      __builtin_amdgcn_wave_barrier();
    }
  }
  __syncthreads();
  inv_combine<16, 2, CT>(A);
  inv_combine<32, 1, CT>(A);
  for (int idx = tid; idx < 64 * 64; idx += 256) {
    const int i = idx % 64, t = idx / 64;
    if (t < i && i < nb) Pn[(j0 + i) + (long long)(j0 + t) * m] = A[i][t];
  }
}

// 64x64 tile product acc += A[rows, k] B[k, cols] over k in [k0, k1); A and B
// are read through callbacks so the same loop serves both inversion products.
template <class CT, class FA, class FB>
__device__ __forceinline__ void tile_gemm(CT (&acc)[4][4], int k0, int k1, FA fa, FB fb) {
  __shared__ CT As[64][17];
  __shared__ CT Bs[16][65];
  const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
  for (int kc = k0; kc < k1; kc += 16) {
    const int kn = min(16, k1 - kc);
    __syncthreads();
    // A's rows and B's k fastest over the lanes: both operands are column-major, so
    // a wave's loads are contiguous pieces of their columns
    for (int idx = tid; idx < 64 * 16; idx += 256) {
      const int r = idx % 64, kk = idx / 64;
      As[r][kk] = kk < kn ? fa(r, kc + kk) : mk<CT>(0.0, 0.0);
      const int kr = idx % 16, c = idx / 16;
      Bs[kr][c] = kr < kn ? fb(kc + kr, c) : mk<CT>(0.0, 0.0);
    }
    __syncthreads();
    for (int kk = 0; kk < kn; ++kk) {
      CT a[4], bb[4];
      for (int p = 0; p < 4; ++p) a[p] = As[tx + 16 * p][kk];
      for (int q = 0; q < 4; ++q) bb[q] = Bs[kk][ty + 16 * q];
      for (int p = 0; p < 4; ++p)
        for (int q = 0; q < 4; ++q) cmac(acc[p][q], a[p], bb[q]);
    }
  }
}

// Z_IJ = L_IJ X_JJ (the scratch, w x 64 per front, column-major), I > J
template <class CT>
__global__ __launch_bounds__(256) void k_inv_z(InvArgs ia, const TileTask* __restrict__ tasks, int t0) {
  const TileTask tk = tasks[t0 + blockIdx.x];
  const int b = blockIdx.y;
  if (ia.active && !ia.active[b]) return;
  const DevSn s = ia.sn[tk.sn];
  const int w = 3 * s.ncol, m = w + 3 * s.nstr;
  const int i0 = 64 * tk.i0, ni = min(64, w - i0), jj0 = 64 * tk.j0, nj = min(64, w - jj0);
  const CT* Pn = static_cast<const CT*>(ia.L) + b * ia.PE + s.poff;
  CT* Sc = static_cast<CT*>(ia.S) + b * ia.SE + s.soff;
  CT acc[4][4];
  for (int p = 0; p < 4; ++p)
    for (int q = 0; q < 4; ++q) acc[p][q] = mk<CT>(0.0, 0.0);
  tile_gemm(acc, 0, nj,
            [&](int r, int k) { return r < ni ? Pn[(i0 + r) + (long long)(jj0 + k) * m] : mk<CT>(0.0, 0.0); },
            [&](int k, int c) {  // X_JJ: unit lower, strictly lower part in the panel
              if (c >= nj || k < c) return mk<CT>(0.0, 0.0);
              return k == c ? mk<CT>(1.0, 0.0) : Pn[(jj0 + k) + (long long)(jj0 + c) * m];
            });
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  for (int p = 0; p < 4; ++p)
    for (int q = 0; q < 4; ++q) {
      const int r = tx + 16 * p, c = ty + 16 * q;
      if (r < ni && c < nj) Sc[(i0 + r) + (long long)c * w] = acc[p][q];
    }
}

// X_IJ = -sum_{J<K<=I} X_IK Z_KJ into the panel, I > J
template <class CT>
__global__ __launch_bounds__(256) void k_inv_tile(InvArgs ia, const TileTask* __restrict__ tasks, int t0) {
  const TileTask tk = tasks[t0 + blockIdx.x];
  const int b = blockIdx.y;
  if (ia.active && !ia.active[b]) return;
  const DevSn s = ia.sn[tk.sn];
  const int w = 3 * s.ncol, m = w + 3 * s.nstr;
  const int i0 = 64 * tk.i0, ni = min(64, w - i0), jj0 = 64 * tk.j0, nj = min(64, w - jj0);
  CT* Pn = static_cast<CT*>(ia.L) + b * ia.PE + s.poff;
  CT* Sc = static_cast<CT*>(ia.S) + b * ia.SE + s.soff;
  CT acc[4][4];
  for (int p = 0; p < 4; ++p)
    for (int q = 0; q < 4; ++q) acc[p][q] = mk<CT>(0.0, 0.0);
  tile_gemm(acc, jj0 + 64, i0 + ni,
            [&](int r, int k) {
              if (r >= ni) return mk<CT>(0.0, 0.0);
              const int gi = i0 + r;
              if (k < gi) return Pn[gi + (long long)k * m];  // X_IK, already inverted
              return mk<CT>(k == gi ? 1.0 : 0.0, 0.0);
            },
            [&](int k, int c) { return c < nj ? Sc[k + (long long)c * w] : mk<CT>(0.0, 0.0); });
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  for (int p = 0; p < 4; ++p)
    for (int q = 0; q < 4; ++q) {
      const int r = tx + 16 * p, c = ty + 16 * q;
      if (r < ni && c < nj) Pn[(i0 + r) + (long long)(jj0 + c) * m] = mk<CT>(-acc[p][q].x, -acc[p][q].y);
    }
}

// ---------------------------------------------------------------------------
// batched triangular solves with p right-hand sides.  Column c of problem b is
// X + b*ldx + c*n; the update vector of front s and its pivot temp live at
// T + b*TE*p + c*TE + toff (r entries, then w entries).
// ---------------------------------------------------------------------------
constexpr int PMAX = 16;

struct SolveArgs {
  const DevSn* sn;
  const int* struct_nodes;
  const int* relmap;
  const int* invmap;  // per front and child: front node position -> child update node, or -1
  const void* L;      // the factor: double2 or float2 (the kernels' LT); vectors are double2
  double2* T;
  long long PE, TE;
  double2* X;
  long long ldx, n;
  const int* active;
  int p;
  const int* lead;    // per problem: whose panels hold the shared fronts (null: its own)
};

// the factor panel of front s for problem b: a shared front's in its class leader's storage
template <class LT>
__device__ __forceinline__ const LT* panel(const SolveArgs& sa, const DevSn& s, int b) {
  const long long pb = s.shared && sa.lead ? sa.lead[b] : b;
  return static_cast<const LT*>(sa.L) + pb * sa.PE + s.poff;
}

__device__ __forceinline__ double2* tcol(const SolveArgs& sa, int b, int c, long long off) {
  return sa.T + ((long long)b * sa.p + c) * sa.TE + off;
}

// the children's update vectors of front s, pulled through the inverse relmaps:
// for each child (child0 first), its map and its update-vector offset
struct ChildPull {
  const int* inv[2];
  long long toff[2];
};
__device__ __forceinline__ ChildPull child_pull(const SolveArgs& sa, const DevSn& s) {
  ChildPull cp{{nullptr, nullptr}, {0, 0}};
  int o = s.inv0;
  for (int k = 0; k < 2; ++k) {
    const int c = k == 0 ? s.child0 : s.child1;
    if (c < 0) continue;
    cp.inv[k] = sa.invmap + o;
    o += s.ncol + s.nstr;
    cp.toff[k] = sa.sn[c].toff;
  }
  return cp;
}

// pivot rows of front s: x += the children's update entries that land on them
// (child0 then child1).  The update rows are written by k_fwd_upd.  LT (the factor's
// element type) is not read here: it names the launches of the real and the complex
// sweeps apart in kernel traces.
template <class LT>
__global__ __launch_bounds__(256) void k_fwd_asm(SolveArgs sa, const int* __restrict__ level_sn, int s0) {
  const int b = blockIdx.y;
  if (sa.active && !sa.active[b]) return;
  const DevSn s = sa.sn[level_sn[s0 + blockIdx.x]];
  const int w = 3 * s.ncol, p = sa.p;
  if (s.child0 < 0 && s.child1 < 0) return;
  const ChildPull cp = child_pull(sa, s);
  for (int idx = threadIdx.x; idx < w * p; idx += 256) {
    const int i = idx % w, c = idx / w;
    double2* x = sa.X + b * sa.ldx + (long long)c * sa.n + 3LL * s.col0 + i;
    double2 v = *x;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (!cp.inv[k]) continue;
      const int q = cp.inv[k][i / 3];
      if (q >= 0) v = cadd(v, *tcol(sa, b, c, cp.toff[k] + 3 * q + i % 3));
    }
    *x = v;
  }
}

// y[i] = x[i] + sum_{j<i} Linv[i, j] x[j]  -> pivot temp (rows [r0, r1) of the pivot block)
template <int PM, class LT>
__global__ __launch_bounds__(64) void k_fwd_inv(SolveArgs sa, const RowTask* __restrict__ tasks, int t0) {
  const int b = blockIdx.y;
  if (sa.active && !sa.active[b]) return;
  const RowTask t = tasks[t0 + blockIdx.x];
  const DevSn s = sa.sn[t.sn];
  const int w = 3 * s.ncol, m = w + 3 * s.nstr, p = sa.p;
  const LT* Pn = panel<LT>(sa, s, b);
  using E = se_t<LT>;
  const double2* x = sa.X + b * sa.ldx + 3LL * s.col0;
  __shared__ double2 xs[64][PM];
  const int i = t.r0 + threadIdx.x;
  double2 acc[PM];
  for (int c = 0; c < PM; ++c) acc[c] = make_double2(0.0, 0.0);
  for (int j0 = 0; j0 < t.r1; j0 += 64) {
    const int nj = min(64, t.r1 - j0);
    __syncthreads();
    for (int idx = threadIdx.x; idx < 64 * PM; idx += 64) {
      const int jj = idx % 64, c = idx / 64;
      xs[jj][c] = (jj < nj && c < p) ? x[(long long)c * sa.n + j0 + jj] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    if (i < t.r1) {
      if (j0 + nj <= t.r0) {  // strictly below this task's diagonal: every lane takes all nj
        const LT* lp = Pn + i + (long long)j0 * m;
        int jj = 0;
        for (; jj + 4 <= nj; jj += 4) {  // four loads in flight; entries past p are zero
          E l[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) l[u] = lde(lp[(long long)(jj + u) * m]);
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int c = 0; c < PM; ++c) emac(acc[c], l[u], xs[jj + u][c]);
        }
        for (; jj < nj; ++jj) {
          const E l = lde(lp[(long long)jj * m]);
#pragma unroll
          for (int c = 0; c < PM; ++c) emac(acc[c], l, xs[jj][c]);
        }
      } else {  // the diagonal chunk: the same pass, entries at or above the diagonal not loaded
        const LT* lp = Pn + i + (long long)j0 * m;
        int jj = 0;
        for (; jj + 4 <= nj; jj += 4) {
          E l[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int j = j0 + jj + u;
            l[u] = j < i ? lde(lp[(long long)(jj + u) * m]) : eunit<E>(j == i);
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int c = 0; c < PM; ++c) emac(acc[c], l[u], xs[jj + u][c]);
        }
        for (; jj < nj; ++jj) {
          const int j = j0 + jj;
          const E l = j < i ? lde(lp[(long long)jj * m]) : eunit<E>(j == i);
#pragma unroll
          for (int c = 0; c < PM; ++c) emac(acc[c], l, xs[jj][c]);
        }
      }
    }
  }
  if (i < t.r1)
    for (int c = 0; c < PM; ++c)
      if (c < p) *tcol(sa, b, c, s.toff + 3 * s.nstr + i) = acc[c];
}

// k_fwd_inv with the column chunks of a task dealt round-robin to four waves
// (their partial sums added in wave order): for the upper tree levels, whose
// few tasks with long rows leave most SIMDs idle under one wave per task
template <int PM, class LT>
__global__ __launch_bounds__(256) void k_fwd_inv4(SolveArgs sa, const RowTask* __restrict__ tasks, int t0) {
  const int b = blockIdx.y;
  if (sa.active && !sa.active[b]) return;
  const RowTask t = tasks[t0 + blockIdx.x];
  const DevSn s = sa.sn[t.sn];
  const int w = 3 * s.ncol, m = w + 3 * s.nstr, p = sa.p;
  const LT* Pn = panel<LT>(sa, s, b);
  using E = se_t<LT>;
  const double2* x = sa.X + b * sa.ldx + 3LL * s.col0;
  __shared__ double2 xs[4][64][PM];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = t.r0 + lane;
  double2 acc[PM];
  for (int c = 0; c < PM; ++c) acc[c] = make_double2(0.0, 0.0);
  const int nch = (t.r1 + 63) / 64;
  for (int q = 0; q < nch; q += 4) {  // the same number of rounds for every wave
    const int j0 = 64 * (q + wv);
    const int nj = q + wv < nch ? min(64, t.r1 - j0) : 0;
    __syncthreads();
    for (int idx = lane; idx < 64 * PM; idx += 64) {
      const int jj = idx % 64, c = idx / 64;
      xs[wv][jj][c] = (jj < nj && c < p) ? x[(long long)c * sa.n + j0 + jj] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    if (i < t.r1 && nj > 0) {
      const LT* lp = Pn + i + (long long)j0 * m;
      if (j0 + nj <= t.r0) {  // strictly below this task's diagonal
        int jj = 0;
        for (; jj + 4 <= nj; jj += 4) {
          E l[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) l[u] = lde(lp[(long long)(jj + u) * m]);
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int c = 0; c < PM; ++c) emac(acc[c], l[u], xs[wv][jj + u][c]);
        }
        for (; jj < nj; ++jj) {
          const E l = lde(lp[(long long)jj * m]);
#pragma unroll
          for (int c = 0; c < PM; ++c) emac(acc[c], l, xs[wv][jj][c]);
        }
      } else {  // the diagonal chunk: entries at or above the diagonal not loaded
        int jj = 0;
        for (; jj + 4 <= nj; jj += 4) {
          E l[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int j = j0 + jj + u;
            l[u] = j < i ? lde(lp[(long long)(jj + u) * m]) : eunit<E>(j == i);
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int c = 0; c < PM; ++c) emac(acc[c], l[u], xs[wv][jj + u][c]);
        }
        for (; jj < nj; ++jj) {
          const int j = j0 + jj;
          const E l = j < i ? lde(lp[(long long)jj * m]) : eunit<E>(j == i);
#pragma unroll
          for (int c = 0; c < PM; ++c) emac(acc[c], l, xs[wv][jj][c]);
        }
      }
    }
  }
  __syncthreads();  // every wave is done with its tile: the tiles now hold the partial sums
  if (wv > 0)
    for (int c = 0; c < PM; ++c) xs[wv][lane][c] = acc[c];
  __syncthreads();
  if (wv == 0 && i < t.r1) {
    for (int v = 1; v < 4; ++v)
      for (int c = 0; c < PM; ++c) acc[c] = cadd(acc[c], xs[v][lane][c]);
    for (int c = 0; c < PM; ++c)
      if (c < p) *tcol(sa, b, c, s.toff + 3 * s.nstr + i) = acc[c];
  }
}

// rows < w: write the pivot solution back into x; rows >= w: t -= L21 y
template <int PM, class LT>
__global__ __launch_bounds__(256) void k_fwd_upd(SolveArgs sa, const RowTask* __restrict__ tasks, int t0) {
  const int b = blockIdx.y;
  if (sa.active && !sa.active[b]) return;
  const RowTask t = tasks[t0 + blockIdx.x];
  const DevSn s = sa.sn[t.sn];
  const int w = 3 * s.ncol, r = 3 * s.nstr, m = w + r, p = sa.p;
  const LT* Pn = panel<LT>(sa, s, b);
  using E = se_t<LT>;
  constexpr int CH = 2048 / PM;
  __shared__ double2 ys[CH][PM];
  const int i = t.r0 + threadIdx.x;
  const bool upd = i < t.r1 && i >= w;
  double2 acc[PM];
  for (int c = 0; c < PM; ++c) acc[c] = make_double2(0.0, 0.0);
  if (t.r1 > w) {
    for (int j0 = 0; j0 < w; j0 += CH) {
      const int nj = min(CH, w - j0);
      __syncthreads();
      for (int idx = threadIdx.x; idx < CH * PM; idx += 256) {
        const int jj = idx % CH, c = idx / CH;
        ys[jj][c] = (jj < nj && c < p) ? *tcol(sa, b, c, s.toff + r + j0 + jj) : make_double2(0.0, 0.0);
      }
      __syncthreads();
      if (upd) {
        const LT* lp = Pn + i + (long long)j0 * m;
        int jj = 0;
        for (; jj + 8 <= nj; jj += 8) {  // eight loads in flight; entries past p are zero
          E l[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) l[u] = lde(lp[(long long)(jj + u) * m]);
#pragma unroll
          for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int c = 0; c < PM; ++c) emac(acc[c], l[u], ys[jj + u][c]);
        }
        for (; jj < nj; ++jj) {
          const E l = lde(lp[(long long)jj * m]);
#pragma unroll
          for (int c = 0; c < PM; ++c) emac(acc[c], l, ys[jj][c]);
        }
      }
    }
  }
  if (i < t.r1 && i < w) {
    for (int c = 0; c < PM; ++c)
      if (c < p) sa.X[b * sa.ldx + (long long)c * sa.n + 3LL * s.col0 + i] = *tcol(sa, b, c, s.toff + r + i);
  } else if (upd) {
    // the update entry: the children's entries that land on row i, less L21 y
    const ChildPull cp = child_pull(sa, s);
    int q[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) q[k] = cp.inv[k] ? cp.inv[k][i / 3] : -1;
    for (int c = 0; c < PM; ++c)
      if (c < p) {
        double2 v = make_double2(0.0, 0.0);
#pragma unroll
        for (int k = 0; k < 2; ++k)
          if (q[k] >= 0) v = cadd(v, *tcol(sa, b, c, cp.toff[k] + 3 * q[k] + i % 3));
        *tcol(sa, b, c, s.toff + i - w) = csub(v, acc[c]);
      }
  }
}

// k_fwd_upd with the pivot columns dealt in 64-column chunks round-robin to NG
// groups of 256 threads (their partial sums added in group order), for the upper
// tree levels, whose few long-row tasks leave most SIMDs idle
template <int PM, int NG, class LT>
__global__ __launch_bounds__(256 * NG) void k_fwd_upd_split(SolveArgs sa, const RowTask* __restrict__ tasks, int t0) {
  const int b = blockIdx.y;
  if (sa.active && !sa.active[b]) return;
  const RowTask t = tasks[t0 + blockIdx.x];
  const DevSn s = sa.sn[t.sn];
  const int w = 3 * s.ncol, r = 3 * s.nstr, m = w + r, p = sa.p;
  const LT* Pn = panel<LT>(sa, s, b);
  using E = se_t<LT>;
  constexpr int CH = 64;
  constexpr int SH = NG * CH * PM > 256 * PM ? NG * CH * PM : 256 * PM;
  __shared__ double2 sh[SH];  // NG tiles of CH x PM solution entries, then the partial sums
  const int g = threadIdx.x >> 8, tr = threadIdx.x & 255;
  double2* ys = sh + g * CH * PM;
  const int i = t.r0 + tr;
  const bool upd = i < t.r1 && i >= w;
  double2 acc[PM];
  for (int c = 0; c < PM; ++c) acc[c] = make_double2(0.0, 0.0);
  if (t.r1 > w) {
    const int nch = (w + CH - 1) / CH;
    for (int q = 0; q < nch; q += NG) {  // the same number of rounds for every group
      const int j0 = CH * (q + g);
      const int nj = q + g < nch ? min(CH, w - j0) : 0;
      __syncthreads();
      for (int idx = tr; idx < CH * PM; idx += 256) {
        const int jj = idx % CH, c = idx / CH;
        ys[jj * PM + c] = (jj < nj && c < p) ? *tcol(sa, b, c, s.toff + r + j0 + jj) : make_double2(0.0, 0.0);
      }
      __syncthreads();
      if (upd && nj > 0) {
        const LT* lp = Pn + i + (long long)j0 * m;
        int jj = 0;
        for (; jj + 8 <= nj; jj += 8) {
          E l[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) l[u] = lde(lp[(long long)(jj + u) * m]);
#pragma unroll
          for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int c = 0; c < PM; ++c) emac(acc[c], l[u], ys[(jj + u) * PM + c]);
        }
        for (; jj < nj; ++jj) {
          const E l = lde(lp[(long long)jj * m]);
#pragma unroll
          for (int c = 0; c < PM; ++c) emac(acc[c], l, ys[jj * PM + c]);
        }
      }
    }
  }
  for (int v = 1; v < NG; ++v) {  // group v's partial sums into group 0, in group order
    __syncthreads();
    if (g == v)
      for (int c = 0; c < PM; ++c) sh[tr * PM + c] = acc[c];
    __syncthreads();
    if (g == 0)
      for (int c = 0; c < PM; ++c) acc[c] = cadd(acc[c], sh[tr * PM + c]);
  }
  if (g != 0) return;
  if (i < t.r1 && i < w) {
    for (int c = 0; c < PM; ++c)
      if (c < p) sa.X[b * sa.ldx + (long long)c * sa.n + 3LL * s.col0 + i] = *tcol(sa, b, c, s.toff + r + i);
  } else if (upd) {
    const ChildPull cp = child_pull(sa, s);
    int q[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) q[k] = cp.inv[k] ? cp.inv[k][i / 3] : -1;
    for (int c = 0; c < PM; ++c)
      if (c < p) {
        double2 v = make_double2(0.0, 0.0);
#pragma unroll
        for (int k = 0; k < 2; ++k)
          if (q[k] >= 0) v = cadd(v, *tcol(sa, b, c, cp.toff[k] + 3 * q[k] + i % 3));
        *tcol(sa, b, c, s.toff + i - w) = csub(v, acc[c]);
      }
  }
}

// backward: the solved ancestor values x[struct(q)] of every front of the level,
// gathered once into the front's update-vector slots (free after the forward
// sweep), so the column dot products below read them contiguously
template <class LT>  // as k_fwd_asm: for the traces only
__global__ __launch_bounds__(256) void k_bwd_gather(SolveArgs sa, const int* __restrict__ level_sn, int s0) {
  const int b = blockIdx.y;
  if (sa.active && !sa.active[b]) return;
  const DevSn s = sa.sn[level_sn[s0 + blockIdx.x]];
  const int r = 3 * s.nstr, p = sa.p;
  const double2* xb = sa.X + b * sa.ldx;
  for (int idx = threadIdx.x; idx < r * p; idx += 256) {
    const int q = idx % r, c = idx / r;
    const long long g = 3LL * sa.struct_nodes[s.str0 + q / 3] + q % 3;
    *tcol(sa, b, c, s.toff + q) = xb[(long long)c * sa.n + g];
  }
}

// Backward sweep of one front, 64 pivot columns [j0, j1) per block:
//   stage 0: z_j = y_j / d_j - sum_q conj(L[w+q, j]) x[struct(q)]   -> pivot temp
//   stage 1: x_j = z_j + sum_{i>j} conj(Linv[i, j]) z_i             -> x
// 32-row tiles of the column-major panel are staged in LDS with coalesced
// loads; thread (j, part) owns column j and right-hand sides part, part + NP,
// ... (PM >= 4), or rows part, part + 4, ... of every tile with a reduction of
// the four partial sums at the end (PM = 1).  Four waves per block.
template <int PM, class LT>
__global__ __launch_bounds__(256) void k_bwd_tile(SolveArgs sa, const RowTask* __restrict__ tasks, int t0,
                                                  int stage) {
  constexpr int NP = 4;
  constexpr int CPT = PM >= 4 ? PM / NP : 1;  // right-hand sides per thread
  constexpr bool ROWSPLIT = PM < 4;
  constexpr int TR = 32;        // tile rows
  const int b = blockIdx.y;
  if (sa.active && !sa.active[b]) return;
  const RowTask t = tasks[t0 + blockIdx.x];
  const DevSn s = sa.sn[t.sn];
  const int w = 3 * s.ncol, r = 3 * s.nstr, m = w + r, p = sa.p;
  const int j0 = t.r0, nj = t.r1 - t.r0;
  const LT* Pn = panel<LT>(sa, s, b);
  using E = se_t<LT>;
  __shared__ E At[TR][64 + 1];
  __shared__ double2 vs[TR][PM];
  const int tid = threadIdx.x, jj = tid % 64, part = tid / 64;
  const int j = j0 + jj;
  double2 acc[CPT];
#pragma unroll
  for (int u = 0; u < CPT; ++u) acc[u] = make_double2(0.0, 0.0);
  const int ib = stage == 0 ? w : j0, ie = stage == 0 ? m : w;
  // a tile's loads are issued one tile ahead (registers), so they overlap the
  // products of the previous tile
  constexpr int APER = TR * 64 / 256, VPER = (TR * PM + 255) / 256;
  E ra[APER];
  double2 rv[VPER];
  auto fetch = [&](int i0) {
    const int ni = min(TR, ie - i0);
#pragma unroll
    for (int e = 0; e < APER; ++e) {
      const int idx = tid + 256 * e, ii = idx % TR, jc = idx / TR;
      ra[e] = (ii < ni && jc < nj) ? lde(Pn[(i0 + ii) + (long long)(j0 + jc) * m]) : E();
    }
#pragma unroll
    for (int e = 0; e < VPER; ++e) {
      const int idx = tid + 256 * e, ii = idx % TR, c = idx / TR;
      const long long off = stage == 0 ? s.toff + (i0 + ii - w) : s.toff + r + (i0 + ii);
      rv[e] = (idx < TR * PM && ii < ni && c < p) ? *tcol(sa, b, c, off) : make_double2(0.0, 0.0);
    }
  };
  if (ib < ie) fetch(ib);
  for (int i0 = ib; i0 < ie; i0 += TR) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < APER; ++e) {
      const int idx = tid + 256 * e;
      At[idx % TR][idx / TR] = ra[e];
    }
#pragma unroll
    for (int e = 0; e < VPER; ++e) {
      const int idx = tid + 256 * e;
      if (idx < TR * PM) vs[idx % TR][idx / TR] = rv[e];
    }
    __syncthreads();
    if (i0 + TR < ie) fetch(i0 + TR);
    // rows past ni are zero in At and vs; in stage 1 only rows strictly below
    // the diagonal (i0 + k > j) count
    const int kmin = stage == 1 ? max(0, min(TR, j - i0 + 1)) : 0;
    if (ROWSPLIT) {
      for (int k = part; k < TR; k += NP) {
        const E a = k >= kmin ? At[k][jj] : E();
#pragma unroll
        for (int u = 0; u < CPT; ++u) emacc(acc[u], a, vs[k][u]);
      }
    } else {
#pragma unroll 8
      for (int k = 0; k < TR; ++k) {
        const E a = k >= kmin ? At[k][jj] : E();
#pragma unroll
        for (int u = 0; u < CPT; ++u) emacc(acc[u], a, vs[k][part + NP * u]);
      }
    }
  }
  if constexpr (ROWSPLIT) {  // fixed-order sum of the four row partials (PM < 4: CPT = 1, columns 0..p-1 ... p = 1)
    __shared__ double2 red[NP][64];
    red[part][jj] = acc[0];
    __syncthreads();
    if (part != 0) return;
    acc[0] = cadd(cadd(red[0][jj], red[1][jj]), cadd(red[2][jj], red[3][jj]));
  }
  if (jj >= nj) return;
  const double2* x = sa.X + b * sa.ldx + 3LL * s.col0;
  if (stage == 0) {
    const double d = ldc(Pn[j + (long long)j * m]).x;
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
      const int c = part + NP * u;
      if (c < p) {
        const double2 y = x[(long long)c * sa.n + j];
        *tcol(sa, b, c, s.toff + r + j) = make_double2(y.x / d - acc[u].x, y.y / d - acc[u].y);
      }
    }
  } else {
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
      const int c = part + NP * u;
      if (c < p) {
        const double2 z = *tcol(sa, b, c, s.toff + r + j);
        sa.X[b * sa.ldx + (long long)c * sa.n + 3LL * s.col0 + j] = make_double2(z.x + acc[u].x, z.y + acc[u].y);
      }
    }
  }
}

}  // namespace

static FactorArgs factor_args(const Mesh& M, const Plan& P, const OpArrays& ops, const BatchBufs& bb) {
  FactorArgs fa;
  fa.sn = P.d_sn.as<DevSn>();
  fa.struct_nodes = M.d_struct.as<int>();
  fa.relmap = M.d_relmap.as<int>();
  fa.invmap = P.d_invmap.as<int>();
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
  fa.pivinfo = bb.pivinfo;
  fa.diag_slot = M.d_diag_slot.as<int>();
  fa.params = bb.params;
  fa.active = nullptr;
  fa.ops_b = nullptr;
  fa.lead = nullptr;
  return fa;
}

// Staged depths of the Schur tiles that the autotuner chooses from, per tree level
// (bit-identical results, different occupancy and barrier counts: depth 8 keeps 18 KB
// of LDS per block, 16 keeps 35 KB, 32 keeps 68 KB with two blocks per CU).
constexpr int kSchurDepths[] = {8, 16, 32};
constexpr int kSchurDefault = 1;  // depth 16

template <class CT, bool PULL>
static void launch_schur(hipStream_t stream, int kc, const FactorStep& st, int B, const FactorArgs& fa,
                         const TileTask* tasks) {
  if (kc == 8)
    hipLaunchKernelGGL((k_schur_tile<CT, 8, PULL>), dim3(st.ntile, B), dim3(256), 0, stream, fa, tasks, st.tile0,
                       st.kb);
  else if (kc == 32)
    hipLaunchKernelGGL((k_schur_tile<CT, 32, PULL>), dim3(st.ntile, B), dim3(256), 0, stream, fa, tasks, st.tile0,
                       st.kb);
  else
    hipLaunchKernelGGL((k_schur_tile<CT, 16, PULL>), dim3(st.ntile, B), dim3(256), 0, stream, fa, tasks, st.tile0,
                       st.kb);
}

// The launches of one tree level's factorization tasks (L: all fronts of the level, or
// the shared or the other fronts of it) on `stream`, Schur tiles at staged depth
// kSchurDepths[dv].
template <class CT>
static void invert_level(hipStream_t stream, const Plan& P, const LevelPlan& L, const InvArgs& ia, int B) {
  for (const InvStep& iv : L.inv) {
    hipLaunchKernelGGL(k_inv_diag<CT>, dim3(iv.ndiag, B), dim3(256), 0, stream, ia, P.d_inv_diag.as<int>(),
                       iv.diag0, iv.J);
    if (iv.ntile) {
      hipLaunchKernelGGL(k_inv_z<CT>, dim3(iv.ntile, B), dim3(256), 0, stream, ia, P.d_inv_tiles.as<TileTask>(),
                         iv.tile0);
      hipLaunchKernelGGL(k_inv_tile<CT>, dim3(iv.ntile, B), dim3(256), 0, stream, ia, P.d_inv_tiles.as<TileTask>(),
                         iv.tile0);
    }
  }
}

// (inv: the level's pivot-block inversion too)
template <class CT>
static void factor_level(hipStream_t stream, const Plan& P, const LevelPlan& L, const FactorArgs& fa,
                         const InvArgs& ia, int B, int dv, bool inv = true) {
  if (L.nasm)
    hipLaunchKernelGGL(k_front_assemble<CT>, dim3(L.nasm, B), dim3(256), 0, stream, fa, P.d_asm.as<AsmTask>(),
                       L.asm0);
  for (const FactorStep& st : L.steps) {
    if (st.ndiag)
      hipLaunchKernelGGL(k_diag_ldl<CT>, dim3(st.ndiag, B), dim3(256), 0, stream, fa, P.d_diag.as<int>(), st.diag0,
                         st.kb);
    if (st.ntrsm)
      hipLaunchKernelGGL(k_panel_trsm<CT>, dim3(st.ntrsm, B), dim3(256), 0, stream, fa, P.d_row.as<RowTask>(),
                         st.trsm0, st.kb);
    if (st.ntile) {
      if (st.pull)
        launch_schur<CT, true>(stream, kSchurDepths[dv], st, B, fa, P.d_tile.as<TileTask>());
      else
        launch_schur<CT, false>(stream, kSchurDepths[dv], st, B, fa, P.d_tile.as<TileTask>());
    }
  }
  if (inv) invert_level<CT>(stream, P, L, ia, B);
}

// One factorization's launches; `depth` gives per tree level an index into
// kSchurDepths (null: the default), and `marks` (when given, levels + 1 events) are
// recorded at the level boundaries.  With `sfa` (shared-front arguments: the class
// leaders' masks and statistics) each level runs its shared fronts with sfa / sia and
// its other fronts with fa / ia.  With ctx->fork, the shared fronts run on the context's
// side stream and the other fronts on its stream, level h of each after
// level h - 1 of both: a front reads only lower levels' update matrices, and the
// level-scheduled update pool is reused only from one level to a later one, so the two
// parts of one level run side by side (the leaders' shared fronts are small launches,
// latency rather than throughput).  The other fronts' pivot-block inversion of level h
// runs on the side stream too, before its shared fronts of level h + 1: it touches only
// level h's panels, which no later level reads while factoring, and the inversion
// scratch, which only the side stream uses.  The side stream forks from the stream first
// and joins it at the end, so the sequence records into one graph.
template <class CT>
static void factor_levels(Context* ctx, const Plan& P, const FactorArgs& fa, const InvArgs& ia, int B,
                          const std::vector<int8_t>* depth = nullptr, hipEvent_t* marks = nullptr,
                          const FactorArgs* sfa = nullptr, const InvArgs* sia = nullptr) {
  const size_t H = P.levels.size();
  if (!sfa || !ctx->fork) {  // one stream: a level's shared fronts, then its other fronts
    if (marks) TM_HIP(hipEventRecord(marks[0], ctx->stream));
    for (size_t h = 0; h < H; ++h) {
      const int dv = depth ? (*depth)[h] : kSchurDefault;
      if (sfa) {
        factor_level<CT>(ctx->stream, P, P.slevels[h], *sfa, *sia, B, dv);
        factor_level<CT>(ctx->stream, P, P.clevels[h], fa, ia, B, dv);
      } else {
        factor_level<CT>(ctx->stream, P, P.levels[h], fa, ia, B, dv);
      }
      if (marks) TM_HIP(hipEventRecord(marks[h + 1], ctx->stream));
    }
    return;
  }
  if (H == 0) return;
  hipStream_t ms = ctx->stream, ss = ctx->side_stream();
  TM_HIP(hipEventRecord(ctx->fork_event(0), ms));
  TM_HIP(hipStreamWaitEvent(ss, ctx->fork_event(0), 0));
  for (size_t h = 0; h < H; ++h) {
    const int dv = depth ? (*depth)[h] : kSchurDefault;
    hipEvent_t sdone = ctx->fork_event(1 + 2 * h), cdone = ctx->fork_event(2 + 2 * h);
    if (h > 0) {
      TM_HIP(hipStreamWaitEvent(ss, ctx->fork_event(2 + 2 * (h - 1)), 0));
      invert_level<CT>(ss, P, P.clevels[h - 1], ia, B);
      TM_HIP(hipStreamWaitEvent(ms, ctx->fork_event(1 + 2 * (h - 1)), 0));
    }
    factor_level<CT>(ss, P, P.slevels[h], *sfa, *sia, B, dv);
    TM_HIP(hipEventRecord(sdone, ss));
    factor_level<CT>(ms, P, P.clevels[h], fa, ia, B, dv, false);
    TM_HIP(hipEventRecord(cdone, ms));
  }
  TM_HIP(hipStreamWaitEvent(ss, ctx->fork_event(2 + 2 * (H - 1)), 0));
  invert_level<CT>(ss, P, P.clevels[H - 1], ia, B);
  TM_HIP(hipEventRecord(ctx->fork_event(1 + 2 * H), ss));
  TM_HIP(hipStreamWaitEvent(ms, ctx->fork_event(1 + 2 * H), 0));
}

// ---- the frontal-update autotuner ------------------------------------------
// The Schur tiles' staged depth is chosen per tree level by timing: the first batched
// factorization of a (plan, element type, batch size class) runs once with each
// depth at every level, timed per level with events, and every level keeps its
// fastest depth (the default unless another is more than 1 % faster).  The results do
// not depend on the choice.  While it times, the tuner holds g_launch_mu alone and the
// device is idle otherwise (every launch sequence of the engine holds it shared), so
// other slot groups' kernels do not enter the timings.  TMHIP_AUTOTUNE=0: the default
// depth everywhere; TMHIP_TRACE: the timings and choices on stderr.
namespace {
std::shared_mutex g_launch_mu;
std::mutex g_tune_mu;
std::map<std::tuple<uint64_t, int, int>, std::vector<int8_t>> g_tune;

template <class CT>
std::vector<int8_t> tune_levels(Context* ctx, const Plan& P, const FactorArgs& fa, const InvArgs& ia, int B) {
  const size_t nl = P.levels.size();
  constexpr int nv = sizeof(kSchurDepths) / sizeof(kSchurDepths[0]);
  std::vector<hipEvent_t> ev((nl + 1) * nv, nullptr);
  for (auto& e : ev) TM_HIP(hipEventCreate(&e));
  std::vector<std::vector<float>> ms(nv, std::vector<float>(nl, 0.0f));
  try {
    for (int v = 0; v < nv; ++v) {
      const std::vector<int8_t> all(nl, (int8_t)v);
      factor_levels<CT>(ctx, P, fa, ia, B, &all, &ev[(nl + 1) * v]);
    }
    TM_HIP(hipGetLastError());
    TM_HIP(hipStreamSynchronize(ctx->stream));
    for (int v = 0; v < nv; ++v)
      for (size_t h = 0; h < nl; ++h)
        TM_HIP(hipEventElapsedTime(&ms[v][h], ev[(nl + 1) * v + h], ev[(nl + 1) * v + h + 1]));
  } catch (...) {
    for (auto& e : ev) (void)hipEventDestroy(e);
    throw;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  std::vector<int8_t> pick(nl, (int8_t)kSchurDefault);
  const bool trace = std::getenv("TMHIP_TRACE") != nullptr;
  for (size_t h = 0; h < nl; ++h) {
    int best = kSchurDefault;
    for (int v = 0; v < nv; ++v)
      if (ms[v][h] < ms[best][h]) best = v;
    if (ms[best][h] < 0.99f * ms[kSchurDefault][h]) pick[h] = (int8_t)best;
    if (trace) {
      std::fprintf(stderr, "[tmhip] tune B=%d level %zu:", B, h);
      for (int v = 0; v < nv; ++v) std::fprintf(stderr, " depth %d %.3f ms", kSchurDepths[v], ms[v][h]);
      std::fprintf(stderr, " -> %d\n", kSchurDepths[pick[h]]);
    }
  }
  return pick;
}
}  // namespace

// The per-level depths for this factorization: tuned on first use of its key when
// `tune` (a batched sweep, which can afford three extra factorizations and a moment of
// the device alone); otherwise the table of the nearest batch class tuned so far for
// this plan and element type, or the default (a serving loop never stops the device).
template <class CT>
static std::vector<int8_t> schur_depths(Context* ctx, const Plan& P, const FactorArgs& fa, const InvArgs& ia,
                                        int B, FactorElem elem, bool tune) {
  const std::vector<int8_t> dflt(P.levels.size(), (int8_t)kSchurDefault);
  const char* env = std::getenv("TMHIP_AUTOTUNE");
  if (env && std::strcmp(env, "0") == 0) return dflt;
  int bucket = 1;
  while (bucket < B) bucket <<= 1;
  const auto key = std::make_tuple(P.serial, (int)elem, bucket);
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) return it->second;
    if (!tune) {
      const std::vector<int8_t>* best = &dflt;
      int gap = 1 << 30;
      for (const auto& kv : g_tune)
        if (std::get<0>(kv.first) == P.serial && std::get<1>(kv.first) == (int)elem) {
          const int b2 = std::get<2>(kv.first), d = b2 > bucket ? b2 / bucket : bucket / b2;
          if (d < gap) {
            gap = d;
            best = &kv.second;
          }
        }
      return *best;
    }
  }
  std::unique_lock<std::shared_mutex> alone(g_launch_mu);
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) return it->second;
  }
  TM_HIP(hipDeviceSynchronize());  // no other group's work in the timings
  std::vector<int8_t> pick = tune_levels<CT>(ctx, P, fa, ia, B);
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tune[key] = pick;
  return pick;
}

void run_captured(Context* ctx, const std::string& key, const std::function<void()>& record) {
  const char* env = std::getenv("TMHIP_GRAPHS");
  if (env && std::strcmp(env, "0") == 0) {
    record();
    return;
  }
  auto it = ctx->graphs.find(key);
  if (it == ctx->graphs.end()) {
    if (ctx->graphs.size() >= 32) {  // bounded: a new batch shape starts afresh
      // graphs launched earlier on this stream may still run: an executable graph is
      // destroyed only once its launches have finished
      TM_HIP(hipStreamSynchronize(ctx->stream));
      ctx->clear_graphs();
      if (std::getenv("TMHIP_TRACE")) std::fprintf(stderr, "[tmhip] graph cache full: cleared\n");
    }
    hipGraph_t g = nullptr;
    TM_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    try {
      record();
    } catch (...) {
      (void)hipStreamEndCapture(ctx->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    const hipError_t le = hipGetLastError();
    TM_HIP(hipStreamEndCapture(ctx->stream, &g));
    if (le != hipSuccess) {
      (void)hipGraphDestroy(g);
      throw HipError(std::string("captured launch: ") + hipGetErrorString(le));
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    TM_HIP(e);
    it = ctx->graphs.emplace(key, ex).first;
  }
  TM_HIP(hipGraphLaunch(it->second, ctx->stream));
}

void factor_batch(Context* ctx, const Mesh& M, const Plan& P, const OpArrays& ops,
                  const BatchBufs& bb, const int* active, int nactive, const OpArrays* ops_b, bool graph,
                  bool tune) {
  FactorArgs fa = factor_args(M, P, ops, bb);
  fa.active = active;
  fa.ops_b = ops_b;
  InvArgs ia{P.d_sn.as<DevSn>(), bb.L, bb.S, P.panel_entries, P.inv_scratch_entries, active};
  const int B = bb.B;
  // algorithmic flops: 4 sum c_j^2 real flops for a complex factor, a quarter of that
  // for a real one
  const bool real = bb.elem == FactorElem::R64;
  // per-level Schur depths: tuned on first use of the key (tune: the batched sweeps,
  // before the timed region), else a table tuned earlier for this plan and element type,
  // else the default
  std::vector<int8_t> depth;
  switch (bb.elem) {
    case FactorElem::C64: depth = schur_depths<double2>(ctx, P, fa, ia, B, bb.elem, tune); break;
    case FactorElem::C32: depth = schur_depths<float2>(ctx, P, fa, ia, B, bb.elem, tune); break;
    case FactorElem::R64: depth = schur_depths<rdouble>(ctx, P, fa, ia, B, bb.elem, tune); break;
  }
  // shared fronts once per class: by the class leaders, with their own masks and
  // statistics, into the leaders' storage, where the members' factorizations and
  // triangular sweeps read them (TMHIP_SHARE=0: every problem factors every front into
  // its own storage; the results are the same)
  const char* share_env = std::getenv("TMHIP_SHARE");
  const bool share = bb.lead && bb.share_ws && P.has_shared && !(share_env && std::strcmp(share_env, "0") == 0);
  FactorArgs sfa = fa;
  InvArgs sia = ia;
  if (share) {
    if (bb.nlead < 1) throw std::logic_error("factor: shared fronts without a class leader");
    fa.lead = bb.lead;
    sfa = fa;
    sfa.active = bb.share_ws;
    sfa.status = bb.share_ws + B;
    sfa.pivinfo = bb.share_ws + 2 * B;
    sia.active = bb.share_ws;
  }
  const int nf = nactive >= 0 ? nactive : B;
  const double flops = share ? P.flops_other * nf + P.flops_shared * bb.nlead : M.S.flops * nf;
  KTimer t(ctx, bb.elem == FactorElem::C32 ? "factor_fp32" : (real ? "factor_real" : "factor"), 0,
           flops * (real ? 0.25 : 1.0));
  std::shared_lock<std::shared_mutex> shared(g_launch_mu);
  auto record = [&] {
    if (bb.pivinfo)
      hipLaunchKernelGGL(k_pivinfo_init, dim3((B + 63) / 64), dim3(64), 0, ctx->stream, bb.pivinfo, active, B);
    if (share)
      hipLaunchKernelGGL(k_share_prep, dim3((B + 255) / 256), dim3(256), 0, ctx->stream, bb.lead, active, B,
                         bb.share_ws);
    const FactorArgs* sf = share ? &sfa : nullptr;
    const InvArgs* si = share ? &sia : nullptr;
    switch (bb.elem) {
      case FactorElem::C64: factor_levels<double2>(ctx, P, fa, ia, B, &depth, nullptr, sf, si); break;
      case FactorElem::C32: factor_levels<float2>(ctx, P, fa, ia, B, &depth, nullptr, sf, si); break;
      case FactorElem::R64: factor_levels<rdouble>(ctx, P, fa, ia, B, &depth, nullptr, sf, si); break;
    }
    if (share)
      hipLaunchKernelGGL(k_share_finish, dim3((B + 255) / 256), dim3(256), 0, ctx->stream, bb.lead, active, B,
                         bb.status, bb.pivinfo, bb.share_ws);
  };
  if (share && ctx->fork) {  // the side stream and its events exist before any capture
    ctx->side_stream();
    ctx->fork_event(2 * P.levels.size() + 1);
  }
  if (graph) {
    // every launch argument: the plan (its task tables and level structure), the
    // argument blocks (all 8-byte fields, no padding), the batch and the element type
    static_assert(sizeof(FactorArgs) % 8 == 0 && sizeof(InvArgs) % 8 == 0, "unpadded argument blocks");
    std::string key("factor");
    key_add(key, P.serial);
    key_add(key, fa);
    key_add(key, ia);
    key_add(key, B);
    key_add(key, bb.elem);
    key_add(key, bb.pivinfo);
    key_add(key, share);
    key_add(key, ctx->fork);
    key_add(key, bb.share_ws);
    key.append(reinterpret_cast<const char*>(depth.data()), depth.size());
    run_captured(ctx, key, record);
  } else {
    record();
  }
  TM_HIP(hipGetLastError());
}

// tasks per level at or below which k_fwd_inv4 runs (measured on the bladed sector:
// 11 to 52 tasks 242-169 -> 99-147 us, but 61 to 76 tasks slower, 46 -> 102 us at 76)
constexpr int kFwdInvSplit = 56;
// k_fwd_upd_split runs on levels of kFwdUpdSplitLo..kFwdUpdSplitHi tasks (measured on
// the bladed sector: 10 tasks 287 -> 151 us, 17 tasks 355 -> 306 us, but 59 to 121
// tasks slower, up to 383 -> 617 us, and 3 tasks 6 -> 10 us)
constexpr int kFwdUpdSplitLo = 8, kFwdUpdSplitHi = 24;

template <int PM, class LT>
static void trisolve_levels(Context* ctx, const Plan& P, const SolveArgs& sa, int B) {
  const int* lsn = P.d_level_sn.as<int>();
  for (const LevelPlan& L : P.levels) {
    hipLaunchKernelGGL(k_fwd_asm<LT>, dim3(L.nsn, B), dim3(256), 0, ctx->stream, sa, lsn, L.sn0);
    // levels with few pivot-row tasks (the upper tree) split each task's columns
    // over four waves; the choice depends on the level only, never on the batch
    if (L.nfinv > kFwdInvSplit)
      hipLaunchKernelGGL((k_fwd_inv<PM, LT>), dim3(L.nfinv, B), dim3(64), 0, ctx->stream, sa, P.d_finv.as<RowTask>(),
                         L.finv0);
    else if (L.nfinv)
      hipLaunchKernelGGL((k_fwd_inv4<PM, LT>), dim3(L.nfinv, B), dim3(256), 0, ctx->stream, sa, P.d_finv.as<RowTask>(),
                         L.finv0);
    constexpr int NG = PM <= 8 ? 4 : 2;
    if (L.nfup > kFwdUpdSplitHi || (L.nfup && L.nfup < kFwdUpdSplitLo))
      hipLaunchKernelGGL((k_fwd_upd<PM, LT>), dim3(L.nfup, B), dim3(256), 0, ctx->stream, sa, P.d_fup.as<RowTask>(),
                         L.fup0);
    else if (L.nfup)
      hipLaunchKernelGGL((k_fwd_upd_split<PM, NG, LT>), dim3(L.nfup, B), dim3(256 * NG), 0, ctx->stream, sa,
                         P.d_fup.as<RowTask>(), L.fup0);
  }
  for (int h = static_cast<int>(P.levels.size()) - 1; h >= 0; --h) {
    const LevelPlan& L = P.levels[h];
    if (!L.nbcol) continue;
    hipLaunchKernelGGL(k_bwd_gather<LT>, dim3(L.nsn, B), dim3(256), 0, ctx->stream, sa, lsn, L.sn0);
    hipLaunchKernelGGL((k_bwd_tile<PM, LT>), dim3(L.nbcol, B), dim3(256), 0, ctx->stream, sa, P.d_bcol.as<RowTask>(),
                       L.bcol0, 0);
    hipLaunchKernelGGL((k_bwd_tile<PM, LT>), dim3(L.nbcol, B), dim3(256), 0, ctx->stream, sa, P.d_bcol.as<RowTask>(),
                       L.bcol0, 1);
  }
}

void trisolve_batch(Context* ctx, const Mesh& M, const Plan& P, const BatchBufs& bb, double2* X,
                    long long ldx, const int* active, int p, long long colstride, int nactive, bool graph,
                    int nclass) {
  if (p < 1 || p > PMAX) throw ArgError("trisolve: 1 <= p <= 16 right-hand sides");
  SolveArgs sa;
  sa.sn = P.d_sn.as<DevSn>();
  sa.struct_nodes = M.d_struct.as<int>();
  sa.relmap = M.d_relmap.as<int>();
  sa.invmap = P.d_invmap.as<int>();
  sa.L = bb.L;
  sa.T = bb.T;
  sa.PE = P.panel_entries;
  sa.TE = P.tvec_entries;
  sa.X = X;
  sa.ldx = ldx;
  sa.n = colstride > 0 ? colstride : 3LL * M.S.nf;  // column stride of X
  sa.active = active;
  sa.p = p;
  sa.lead = bb.lead;
  const int B = bb.B;
  // algorithmic bytes: the factor once (L21 + inverted pivot triangles; the shared
  // fronts' part once per class) and ~5 n-vectors per right-hand side
  const double vec_bytes = 5.0 * 3 * M.S.nf * 16.0;
  const double l_bytes = P.trisolve_bytes_c16 - vec_bytes;
  const int nb = nactive >= 0 ? std::min(nactive, B) : B;  // problems that do the work
  const double l_read = bb.lead && nclass >= 0
                            ? (l_bytes - P.trisolve_shared_bytes_c16) * nb + P.trisolve_shared_bytes_c16 * nclass
                            : l_bytes * nb;
  // a complex-FP32 or real-FP64 factor entry is 8 + 4 B instead of 16 + 4
  const bool half = bb.elem != FactorElem::C64;
  const double lb = half ? l_read * 12.0 / 20.0 : l_read;
  const char* name = bb.elem == FactorElem::C32 ? "trisolve_fp32" : (bb.elem == FactorElem::R64 ? "trisolve_real" : "trisolve");
  KTimer t(ctx, name, lb + p * vec_bytes * nb, 8.0 * p * nb * l_bytes / 20.0);
  std::shared_lock<std::shared_mutex> shared(g_launch_mu);  // not while the autotuner times
  auto levels = [&](auto lt) {
    using LT = decltype(lt);
    if (p == 1) trisolve_levels<1, LT>(ctx, P, sa, B);
    else if (p <= 4) trisolve_levels<4, LT>(ctx, P, sa, B);
    else if (p <= 8) trisolve_levels<8, LT>(ctx, P, sa, B);
    else trisolve_levels<16, LT>(ctx, P, sa, B);
  };
  auto record = [&] {
    switch (bb.elem) {
      case FactorElem::C64: levels(double2{}); break;
      case FactorElem::C32: levels(float2{}); break;
      case FactorElem::R64: levels(rdouble{}); break;
    }
  };
  if (graph) {
    std::string key("trisolve");
    key_add(key, P.serial);
    key_add(key, sa.sn);  // the argument block field by field (it ends in padding)
    key_add(key, sa.struct_nodes);
    key_add(key, sa.relmap);
    key_add(key, sa.invmap);
    key_add(key, sa.L);
    key_add(key, sa.lead);
    key_add(key, sa.T);
    key_add(key, sa.PE);
    key_add(key, sa.TE);
    key_add(key, sa.X);
    key_add(key, sa.ldx);
    key_add(key, sa.n);
    key_add(key, sa.active);
    key_add(key, sa.p);
    key_add(key, B);
    key_add(key, bb.elem);
    run_captured(ctx, key, record);
  } else {
    record();
  }
  TM_HIP(hipGetLastError());
}

}  // namespace tme
