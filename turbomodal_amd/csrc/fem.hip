// TET10 element kernels and fixed-order segmented assembly (gfx950).
//
// Element integrals follow src/element.cpp:292-371 (Ke, Me: 14-point Keast,
// |detJ|), src/rotating_effects.cpp:5-228 (centrifugal load: 4-point rule;
// spin softening, gyroscopic, stress stiffening: Keast-14 with the signed
// detJ) and src/assembler.cpp:71-155 (prestress sigma = D B(gp) u_e).
//
// One 64-lane wavefront integrates one element: the 14 Gauss-point geometries
// are formed cooperatively in LDS, then lane p < 55 owns node pair p (a <= b)
// and accumulates its 3x3 stiffness block and scalar masses in registers.
// Me, Kw and G are all rho*N_a*N_b*detJ*w times a constant 3x3 (I, I - aa^T,
// [a x]), so the kernel stores only the scalar mass per node pair.
//
// Assembly never forms the sector-size matrices: every 3x3 element block is
// summed straight into the phase-split cyclic blocks (K_c / K_p etc.) of the
// free-node pattern, rotated by R(alpha) on the right-node side(s) (T0 of
// src/cyclic_solver.cpp:291-327).  Each output slot owns the ordered list of
// its contributions, so the sum order is fixed (elements ascending, then a,
// then b) and the result is bit-reproducible run to run.
#include "engine.hpp"
#include "kernels.hpp"

namespace tme {

__constant__ double c_keast_pts[14][3] = {
    {0.5, 0.5, 0.0}, {0.5, 0.0, 0.5}, {0.5, 0.0, 0.0}, {0.0, 0.5, 0.5}, {0.0, 0.5, 0.0},
    {0.0, 0.0, 0.5},
    {0.100526765225204467, 0.100526765225204467, 0.100526765225204467},
    {0.698419704324386603, 0.100526765225204467, 0.100526765225204467},
    {0.100526765225204467, 0.698419704324386603, 0.100526765225204467},
    {0.100526765225204467, 0.100526765225204467, 0.698419704324386603},
    {0.314372873493192195, 0.314372873493192195, 0.314372873493192195},
    {0.0568813795204234229, 0.314372873493192195, 0.314372873493192195},
    {0.314372873493192195, 0.0568813795204234229, 0.314372873493192195},
    {0.314372873493192195, 0.314372873493192195, 0.0568813795204234229}};
__constant__ double c_keast_w[14] = {
    0.00317460317460317503, 0.00317460317460317503, 0.00317460317460317503,
    0.00317460317460317503, 0.00317460317460317503, 0.00317460317460317503,
    0.01476497079049678,    0.01476497079049678,    0.01476497079049678,
    0.01476497079049678,    0.02213979111426512,    0.02213979111426512,
    0.02213979111426512,    0.02213979111426512};
__constant__ double c_g4_pts[4][3] = {
    {0.1381966011250105, 0.1381966011250105, 0.1381966011250105},
    {0.5854101966249685, 0.1381966011250105, 0.1381966011250105},
    {0.1381966011250105, 0.5854101966249685, 0.1381966011250105},
    {0.1381966011250105, 0.1381966011250105, 0.5854101966249685}};

// node pair p (0..54) -> (a, b), a <= b
__device__ __forceinline__ void pair_of(int p, int& a, int& b) {
  int aa = 0, rem = p;
  while (rem >= 10 - aa) { rem -= 10 - aa; ++aa; }
  a = aa;
  b = aa + rem;
}

__device__ __forceinline__ void shape_fn(const double* x, double* N) {
  double L1 = 1.0 - x[0] - x[1] - x[2], L2 = x[0], L3 = x[1], L4 = x[2];
  N[0] = L1 * (2.0 * L1 - 1.0);
  N[1] = L2 * (2.0 * L2 - 1.0);
  N[2] = L3 * (2.0 * L3 - 1.0);
  N[3] = L4 * (2.0 * L4 - 1.0);
  N[4] = 4.0 * L1 * L2;
  N[5] = 4.0 * L2 * L3;
  N[6] = 4.0 * L1 * L3;
  N[7] = 4.0 * L1 * L4;
  N[8] = 4.0 * L2 * L4;
  N[9] = 4.0 * L3 * L4;
}

// dN (10 x 3) at a natural point (src/element.cpp:114-201)
__device__ __forceinline__ void shape_deriv(const double* x, double dN[10][3]) {
  double L1 = 1.0 - x[0] - x[1] - x[2], L2 = x[0], L3 = x[1], L4 = x[2];
  double d0 = 1.0 - 4.0 * L1;
  dN[0][0] = d0; dN[0][1] = d0; dN[0][2] = d0;
  dN[1][0] = 4.0 * L2 - 1.0; dN[1][1] = 0.0; dN[1][2] = 0.0;
  dN[2][0] = 0.0; dN[2][1] = 4.0 * L3 - 1.0; dN[2][2] = 0.0;
  dN[3][0] = 0.0; dN[3][1] = 0.0; dN[3][2] = 4.0 * L4 - 1.0;
  dN[4][0] = 4.0 * (L1 - L2); dN[4][1] = -4.0 * L2; dN[4][2] = -4.0 * L2;
  dN[5][0] = 4.0 * L3; dN[5][1] = 4.0 * L2; dN[5][2] = 0.0;
  dN[6][0] = -4.0 * L3; dN[6][1] = 4.0 * (L1 - L3); dN[6][2] = -4.0 * L3;
  dN[7][0] = -4.0 * L4; dN[7][1] = -4.0 * L4; dN[7][2] = 4.0 * (L1 - L4);
  dN[8][0] = 4.0 * L4; dN[8][1] = 0.0; dN[8][2] = 4.0 * L2;
  dN[9][0] = 0.0; dN[9][1] = 4.0 * L4; dN[9][2] = 4.0 * L3;
}

// Geometry of one Gauss point: detJ and physical gradients g[a][3].
// Computed by every lane from LDS coordinates (a few hundred flops).
__device__ __forceinline__ double gp_geometry(const double* xg, const double (*X)[3],
                                              double g[10][3]) {
  double dN[10][3];
  shape_deriv(xg, dN);
  double J[3][3] = {{0}};
  for (int a = 0; a < 10; ++a)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) J[i][j] += dN[a][i] * X[a][j];
  double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  double id = 1.0 / det;
  double Ji[3][3];
  Ji[0][0] = c00 * id;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  Ji[1][0] = c01 * id;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  Ji[2][0] = c02 * id;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  // dNdx = dNdxi * Jinv^T
  for (int a = 0; a < 10; ++a)
    for (int j = 0; j < 3; ++j)
      g[a][j] = dN[a][0] * Ji[j][0] + dN[a][1] * Ji[j][1] + dN[a][2] * Ji[j][2];
  return det;
}

// B_a (6x3) as in src/element.cpp:240-255
__device__ __forceinline__ void b_block(const double* g, double B[6][3]) {
  B[0][0] = g[0]; B[0][1] = 0.0;  B[0][2] = 0.0;
  B[1][0] = 0.0;  B[1][1] = g[1]; B[1][2] = 0.0;
  B[2][0] = 0.0;  B[2][1] = 0.0;  B[2][2] = g[2];
  B[3][0] = g[1]; B[3][1] = g[0]; B[3][2] = 0.0;
  B[4][0] = 0.0;  B[4][1] = g[2]; B[4][2] = g[1];
  B[5][0] = g[2]; B[5][1] = 0.0;  B[5][2] = g[0];
}

struct DMat { double d[36]; };

// out: ke [E][10][10][9] (row-major 3x3 per node pair), m_abs / m_raw [E][10][10],
//      fe [E][30] (unit omega, 4-point rule)
__global__ __launch_bounds__(64) void k_element_base(
    const double* __restrict__ coords, const int* __restrict__ elements, int ne, DMat D,
    double rho, double ax0, double ax1, double ax2,
    double* __restrict__ ke, double* __restrict__ m_abs, double* __restrict__ m_raw,
    double* __restrict__ fe) {
  const int e = blockIdx.x;
  if (e >= ne) return;
  const int lane = threadIdx.x;
  __shared__ double X[10][3];
  __shared__ double G[14][10][3];
  __shared__ double DJ[14];
  if (lane < 30) {
    int a = lane / 3, c = lane % 3;
    X[a][c] = coords[3 * (long long)elements[10 * (long long)e + a] + c];
  }
  __syncthreads();
  if (lane < 14) {
    double g[10][3];
    double xg[3] = {c_keast_pts[lane][0], c_keast_pts[lane][1], c_keast_pts[lane][2]};
    DJ[lane] = gp_geometry(xg, X, g);
    for (int a = 0; a < 10; ++a)
      for (int j = 0; j < 3; ++j) G[lane][a][j] = g[a][j];
  }
  __syncthreads();
  if (lane < 55) {
    int a, b;
    pair_of(lane, a, b);
    double K[3][3] = {{0}};
    double ma = 0.0, mr = 0.0;
    for (int gp = 0; gp < 14; ++gp) {
      double Ba[6][3], Bb[6][3];
      b_block(G[gp][a], Ba);
      b_block(G[gp][b], Bb);
      double det = DJ[gp];
      double wa = fabs(det) * c_keast_w[gp];
      double DB[6][3];
      for (int p = 0; p < 6; ++p)
        for (int d = 0; d < 3; ++d) {
          double s = 0.0;
          for (int q = 0; q < 6; ++q) s += D.d[6 * p + q] * Bb[q][d];
          DB[p][d] = s;
        }
      for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) {
          double s = 0.0;
          for (int p = 0; p < 6; ++p) s += Ba[p][c] * DB[p][d];
          K[c][d] += s * wa;
        }
      double xg[3] = {c_keast_pts[gp][0], c_keast_pts[gp][1], c_keast_pts[gp][2]};
      double N[10];
      shape_fn(xg, N);
      double nn = rho * N[a] * N[b];
      ma += nn * wa;
      mr += nn * det * c_keast_w[gp];
    }
    long long base = (long long)e * 100;
    double* kab = ke + (base + a * 10 + b) * 9;
    double* kba = ke + (base + b * 10 + a) * 9;
    for (int c = 0; c < 3; ++c)
      for (int d = 0; d < 3; ++d) {
        kab[3 * c + d] = K[c][d];
        kba[3 * d + c] = K[c][d];
      }
    m_abs[base + a * 10 + b] = ma;
    m_abs[base + b * 10 + a] = ma;
    m_raw[base + a * 10 + b] = mr;
    m_raw[base + b * 10 + a] = mr;
  }
  // centrifugal load at unit omega, 4-point rule, signed detJ
  if (lane < 30) {
    int a = lane / 3, c = lane % 3;
    double ax[3] = {ax0, ax1, ax2};
    double f = 0.0;
    for (int gp = 0; gp < 4; ++gp) {
      double xg[3] = {c_g4_pts[gp][0], c_g4_pts[gp][1], c_g4_pts[gp][2]};
      double g[10][3];
      double det = gp_geometry(xg, X, g);
      double N[10];
      shape_fn(xg, N);
      double pos[3] = {0, 0, 0};
      for (int q = 0; q < 10; ++q)
        for (int j = 0; j < 3; ++j) pos[j] += N[q] * X[q][j];
      double pa = pos[0] * ax[0] + pos[1] * ax[1] + pos[2] * ax[2];
      double rp = pos[c] - pa * ax[c];
      f += rho * N[a] * rp * (det * (1.0 / 24.0));
    }
    fe[(long long)e * 30 + lane] = f;
  }
}

// Stress-stiffening scalar per node pair: s_ab = sum_14 g_a^T sigma g_b detJ w,
// sigma = D B(gp) u_e (src/assembler.cpp:102-113, src/rotating_effects.cpp:53-138).
__global__ __launch_bounds__(64) void k_element_sigma(
    const double* __restrict__ coords, const int* __restrict__ elements, int ne, DMat D,
    const double* __restrict__ u_full, double* __restrict__ s_out) {
  const int e = blockIdx.x;
  if (e >= ne) return;
  const int lane = threadIdx.x;
  __shared__ double X[10][3];
  __shared__ double U[30];
  __shared__ double G[14][10][3];
  __shared__ double DJ[14];
  __shared__ double SIG[14][6];
  if (lane < 30) {
    int a = lane / 3, c = lane % 3;
    int node = elements[10 * (long long)e + a];
    X[a][c] = coords[3 * (long long)node + c];
    U[lane] = u_full[3 * (long long)node + c];
  }
  __syncthreads();
  if (lane < 14) {
    double g[10][3];
    double xg[3] = {c_keast_pts[lane][0], c_keast_pts[lane][1], c_keast_pts[lane][2]};
    DJ[lane] = gp_geometry(xg, X, g);
    double eps[6] = {0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 10; ++a) {
      double B[6][3];
      b_block(g[a], B);
      for (int p = 0; p < 6; ++p)
        eps[p] += B[p][0] * U[3 * a] + B[p][1] * U[3 * a + 1] + B[p][2] * U[3 * a + 2];
      for (int j = 0; j < 3; ++j) G[lane][a][j] = g[a][j];
    }
    for (int p = 0; p < 6; ++p) {
      double s = 0.0;
      for (int q = 0; q < 6; ++q) s += D.d[6 * p + q] * eps[q];
      SIG[lane][p] = s;
    }
  }
  __syncthreads();
  if (lane < 55) {
    int a, b;
    pair_of(lane, a, b);
    double acc = 0.0;
    for (int gp = 0; gp < 14; ++gp) {
      const double* s = SIG[gp];
      double sm[3][3] = {{s[0], s[3], s[5]}, {s[3], s[1], s[4]}, {s[5], s[4], s[2]}};
      const double* ga = G[gp][a];
      const double* gb = G[gp][b];
      double v = 0.0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) v += ga[i] * sm[i][j] * gb[j];
      acc += v * DJ[gp] * c_keast_w[gp];
    }
    long long base = (long long)e * 100;
    s_out[base + a * 10 + b] = acc;
    s_out[base + b * 10 + a] = acc;
  }
}

// ---------------------------------------------------------------------------
// segmented assembly into phase-split cyclic blocks
// ---------------------------------------------------------------------------
struct RotInfo {
  double c, s;     // cos(alpha), sin(alpha)
  int c1, c2, cax; // rotation plane / axial component
};

// R (3x3) for the T0 right-node map: u_r = R q_l (src/cyclic_solver.cpp:316-326)
__device__ __forceinline__ void rot_matrix(const RotInfo& ri, double R[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i][j] = 0.0;
  R[ri.c1][ri.c1] = ri.c;
  R[ri.c1][ri.c2] = -ri.s;
  R[ri.c2][ri.c1] = ri.s;
  R[ri.c2][ri.c2] = ri.c;
  R[ri.cax][ri.cax] = 1.0;
}

enum SrcKind { SRC_K = 0, SRC_M = 1, SRC_DK = 2, SRC_G = 3 };

// One thread per pattern slot: out[slot] (9 doubles) = sum over its ordered
// contribution list of R_r^T X R_c.
template <int KIND>
__global__ void k_assemble_blocks(int nblk, const int* __restrict__ ptr,
                                  const int* __restrict__ src, const double* __restrict__ ke,
                                  const double* __restrict__ m_abs,
                                  const double* __restrict__ m_raw,
                                  const double* __restrict__ s_ab, RotInfo ri,
                                  double* __restrict__ out) {
  int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nblk) return;
  double R[3][3];
  rot_matrix(ri, R);
  double acc[3][3] = {{0}};
  for (int q = ptr[slot]; q < ptr[slot + 1]; ++q) {
    int code = src[q];
    int flags = code & 3;
    long long eb = code >> 2;
    double X[3][3];
    if (KIND == SRC_K) {
      const double* k = ke + eb * 9;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) X[i][j] = k[3 * i + j];
    } else {
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) X[i][j] = 0.0;
      if (KIND == SRC_M) {
        double m = m_abs[eb];
        X[0][0] = m; X[1][1] = m; X[2][2] = m;
      } else if (KIND == SRC_DK) {
        // Ks1 - Kw1: s_ab I - m_ab (I - a a^T), a = unit rotation axis
        double s = s_ab[eb], m = m_raw[eb];
        for (int i = 0; i < 3; ++i) X[i][i] = s - (i == ri.cax ? 0.0 : m);
      } else {  // SRC_G: m_ab [a x]
        double m = m_raw[eb];
        double av[3] = {0, 0, 0};
        av[ri.cax] = 1.0;
        X[0][1] = -av[2] * m; X[0][2] = av[1] * m;
        X[1][0] = av[2] * m;  X[1][2] = -av[0] * m;
        X[2][0] = -av[1] * m; X[2][1] = av[0] * m;
      }
    }
    double Y[3][3];
    if (flags & 2) {  // column side rotated: X R
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          Y[i][j] = X[i][0] * R[0][j] + X[i][1] * R[1][j] + X[i][2] * R[2][j];
    } else {
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Y[i][j] = X[i][j];
    }
    if (flags & 1) {  // row side rotated: R^T Y
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          acc[i][j] += R[0][i] * Y[0][j] + R[1][i] * Y[1][j] + R[2][i] * Y[2][j];
    } else {
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) acc[i][j] += Y[i][j];
    }
  }
  double* o = out + (long long)slot * 9;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o[3 * i + j] = acc[i][j];
}

// F_cyc = T0^T F per free node, pinned DOFs zeroed (static right-hand side).
__global__ void k_assemble_load(int nf, const int* __restrict__ ptr, const int* __restrict__ src,
                                const double* __restrict__ fe, const unsigned char* __restrict__ pinned,
                                RotInfo ri, double* __restrict__ f) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nf) return;
  double R[3][3];
  rot_matrix(ri, R);
  double acc[3] = {0, 0, 0};
  for (int q = ptr[p]; q < ptr[p + 1]; ++q) {
    int code = src[q];
    long long ea = code >> 1;
    const double* v = fe + ea * 3;
    if (code & 1) {
      for (int i = 0; i < 3; ++i) acc[i] += R[0][i] * v[0] + R[1][i] * v[1] + R[2][i] * v[2];
    } else {
      for (int i = 0; i < 3; ++i) acc[i] += v[i];
    }
  }
  for (int i = 0; i < 3; ++i) f[3 * p + i] = pinned[3 * p + i] ? 0.0 : acc[i];
}

// u_full = T0 u (real, unit phase): the static displacement on every sector node.
__global__ void k_expand_real(int nn, const int* __restrict__ free_of_full,
                              const int* __restrict__ node_class, const double2* __restrict__ u,
                              RotInfo ri, double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  int p = free_of_full[i];
  double v[3] = {0, 0, 0};
  if (p >= 0)
    for (int c = 0; c < 3; ++c) v[c] = u[3 * p + c].x;
  if (p >= 0 && node_class[i] == TM_NODE_RIGHT) {
    double R[3][3];
    rot_matrix(ri, R);
    double w[3];
    for (int a = 0; a < 3; ++a) w[a] = R[a][0] * v[0] + R[a][1] * v[1] + R[a][2] * v[2];
    for (int a = 0; a < 3; ++a) v[a] = w[a];
  }
  for (int c = 0; c < 3; ++c) out[3 * (long long)i + c] = v[c];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static RotInfo rot_info(const Symbolic& S) {
  RotInfo r;
  double alpha = 2.0 * M_PI / S.N;
  r.c = std::cos(alpha);
  r.s = std::sin(alpha);
  r.c1 = S.c1;
  r.c2 = S.c2;
  r.cax = S.cax;
  return r;
}

static DMat dmat(double E, double nu) {
  DMat D{};
  double c = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) D.d[6 * i + j] = c * nu;
  for (int i = 0; i < 3; ++i) D.d[6 * i + i] = c * (1.0 - nu);
  for (int i = 3; i < 6; ++i) D.d[6 * i + i] = c * (1.0 - 2.0 * nu) / 2.0;
  return D;
}

void launch_element_base(Context* ctx, Mesh& M, double E, double nu, double rho, DevBuf& ke,
                         DevBuf& m_abs, DevBuf& m_raw, DevBuf& fe) {
  const Symbolic& S = M.S;
  double ax[3] = {0, 0, 0};
  ax[S.cax] = 1.0;
  double bytes = S.ne * (10.0 * 4 + 30 * 8) + S.ne * (900.0 + 200 + 30) * 8;
  KTimer t(ctx, "element_base", bytes, 0);
  hipLaunchKernelGGL(k_element_base, dim3(S.ne), dim3(64), 0, ctx->stream,
                     M.d_coords.as<double>(), M.d_elements.as<int>(), S.ne, dmat(E, nu), rho,
                     ax[0], ax[1], ax[2], ke.as<double>(), m_abs.as<double>(), m_raw.as<double>(),
                     fe.as<double>());
  TM_HIP(hipGetLastError());
}

void launch_element_sigma(Context* ctx, Mesh& M, double E, double nu, const DevBuf& u_full,
                          DevBuf& s_ab) {
  const Symbolic& S = M.S;
  KTimer t(ctx, "element_sigma", S.ne * (10.0 * 4 + 60 * 8 + 100 * 8), 0);
  hipLaunchKernelGGL(k_element_sigma, dim3(S.ne), dim3(64), 0, ctx->stream,
                     M.d_coords.as<double>(), M.d_elements.as<int>(), S.ne, dmat(E, nu),
                     u_full.as<double>(), s_ab.as<double>());
  TM_HIP(hipGetLastError());
}

void launch_assemble(Context* ctx, Mesh& M, int kind, bool phase, const DevBuf& ke,
                     const DevBuf& m_abs, const DevBuf& m_raw, const DevBuf* s_ab, DevBuf& out) {
  const Symbolic& S = M.S;
  int nblk = static_cast<int>(S.colidx.size());
  const int* ptr = phase ? M.d_kp_ptr.as<int>() : M.d_kc_ptr.as<int>();
  const int* src = phase ? M.d_kp_src.as<int>() : M.d_kc_src.as<int>();
  size_t ncontrib = phase ? S.kp_src.size() : S.kc_src.size();
  double per = kind == SRC_K ? 72.0 : (kind == SRC_DK ? 16.0 : 8.0);
  KTimer t(ctx, "assemble", ncontrib * (4.0 + per) + nblk * 80.0, 0);
  RotInfo ri = rot_info(S);
  dim3 grid((nblk + 255) / 256), blk(256);
  const double* sab = s_ab ? s_ab->as<double>() : nullptr;
  switch (kind) {
    case SRC_K:
      hipLaunchKernelGGL(k_assemble_blocks<SRC_K>, grid, blk, 0, ctx->stream, nblk, ptr, src,
                         ke.as<double>(), m_abs.as<double>(), m_raw.as<double>(), sab, ri,
                         out.as<double>());
      break;
    case SRC_M:
      hipLaunchKernelGGL(k_assemble_blocks<SRC_M>, grid, blk, 0, ctx->stream, nblk, ptr, src,
                         ke.as<double>(), m_abs.as<double>(), m_raw.as<double>(), sab, ri,
                         out.as<double>());
      break;
    case SRC_DK:
      hipLaunchKernelGGL(k_assemble_blocks<SRC_DK>, grid, blk, 0, ctx->stream, nblk, ptr, src,
                         ke.as<double>(), m_abs.as<double>(), m_raw.as<double>(), sab, ri,
                         out.as<double>());
      break;
    default:
      hipLaunchKernelGGL(k_assemble_blocks<SRC_G>, grid, blk, 0, ctx->stream, nblk, ptr, src,
                         ke.as<double>(), m_abs.as<double>(), m_raw.as<double>(), sab, ri,
                         out.as<double>());
  }
  TM_HIP(hipGetLastError());
}

void launch_assemble_load(Context* ctx, Mesh& M, const DevBuf& fe, DevBuf& f) {
  const Symbolic& S = M.S;
  RotInfo ri = rot_info(S);
  hipLaunchKernelGGL(k_assemble_load, dim3((S.nf + 255) / 256), dim3(256), 0, ctx->stream, S.nf,
                     M.d_f_ptr.as<int>(), M.d_f_src.as<int>(), fe.as<double>(),
                     M.d_pinned.as<unsigned char>(), ri, f.as<double>());
  TM_HIP(hipGetLastError());
}

void launch_expand_real(Context* ctx, Mesh& M, const double2* u, DevBuf& out) {
  const Symbolic& S = M.S;
  RotInfo ri = rot_info(S);
  hipLaunchKernelGGL(k_expand_real, dim3((S.nn + 255) / 256), dim3(256), 0, ctx->stream, S.nn,
                     M.d_free_of_full.as<int>(), M.d_node_class.as<int>(), u, ri,
                     out.as<double>());
  TM_HIP(hipGetLastError());
}

}  // namespace tme
