// Host driver of the engine: context timing, operator assembly with the
// static prestress solve, and the batched shift-invert Lanczos.
//
// The eigen-solver restates HermitianLanczosEigenSolver::solve and
// implicit_restart (src/hermitian_lanczos.cpp:114-460) with the policy of
// CyclicSymmetrySolver::solve_harmonic (src/cyclic_solver.cpp:937-1224):
// sigma = 1, nev = ceil(M/2) for degenerate harmonics, ncv = min(max(4 nev + 1,
// 40), n_free), full M-orthogonal reorthogonalisation (classical Gram-Schmidt
// applied twice on the device instead of one modified pass), exact-shift
// implicit QR restarts, Ritz vectors V S.  Systems with n_free <= 200 use a
// full Krylov space (ncv = n_free), which is the exact dense solution the
// reference obtains from its dense fallback.
#include <algorithm>
#include <chrono>
#include <complex>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>

#include "vec.hpp"

namespace tme {

// ---------------------------------------------------------------------------
// timing
// ---------------------------------------------------------------------------
hipEvent_t Context::get_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  TM_HIP(hipEventCreate(&e));
  return e;
}

void Context::flush_timing() {
  if (pending.empty()) return;
  TM_HIP(hipStreamSynchronize(stream));
  for (auto& p : pending) {
    float ms = 0.f;
    TM_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    KernelStat& s = stats[p.name];
    s.launches += 1;
    s.total_ms += ms;
    s.bytes += p.bytes;
    s.flops += p.flops;
    event_pool.push_back(p.a);
    event_pool.push_back(p.b);
  }
  pending.clear();
}

KTimer::KTimer(Context* c, const char* n, double by, double fl) : ctx(c), name(n), bytes(by), flops(fl) {
  if (ctx && ctx->profiling) {
    a = ctx->get_event();
    TM_HIP(hipEventRecord(a, ctx->stream));
  }
}

KTimer::~KTimer() {
  if (ctx && ctx->profiling && a) {
    hipEvent_t b = ctx->get_event();
    if (hipEventRecord(b, ctx->stream) == hipSuccess) ctx->pending.push_back({name, a, b, bytes, flops});
    if (ctx->pending.size() > 4096) {
      try { ctx->flush_timing(); } catch (...) {}
    }
  }
}

// ---------------------------------------------------------------------------
// tridiagonal eigen-solver (implicit QL with Wilkinson shifts)
// ---------------------------------------------------------------------------
void tridiag_eigen(int n, std::vector<double>& d, std::vector<double> e, std::vector<double>& z) {
  z.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) z[i + (size_t)i * n] = 1.0;
  e.resize(n);
  if (n > 0) e[n - 1] = 0.0;
  const double eps = std::numeric_limits<double>::epsilon();
  for (int l = 0; l < n; ++l) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; ++m) {
        double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd) break;
      }
      if (m != l) {
        if (++iter > 200) throw NumericError("tridiagonal eigensolver did not converge");
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        bool early = false;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i], b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            early = true;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          double* zi = &z[(size_t)i * n];
          double* zi1 = &z[(size_t)(i + 1) * n];
          for (int k = 0; k < n; ++k) {
            double fz = zi1[k];
            zi1[k] = s * zi[k] + c * fz;
            zi[k] = c * zi[k] - s * fz;
          }
        }
        if (early) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
  std::vector<double> d2(n), z2((size_t)n * n);
  for (int q = 0; q < n; ++q) {
    d2[q] = d[idx[q]];
    std::memcpy(&z2[(size_t)q * n], &z[(size_t)idx[q] * n], sizeof(double) * n);
  }
  d.swap(d2);
  z.swap(z2);
}

// ---------------------------------------------------------------------------
// Hermitian eigenproblem of the projected block-Lanczos matrix: Householder
// reduction to complex tridiagonal form, a diagonal phase change to a real
// tridiagonal, implicit QL (tridiag_eigen), back-transformation.
// A: m x m row-major complex (re, im); theta ascending; S column q = S[k*m+q].
// ---------------------------------------------------------------------------
void herm_eig(int m, const std::vector<double>& A, std::vector<double>& theta, std::vector<double>& S) {
  using cd = std::complex<double>;
  std::vector<cd> a(m * m), Q(m * m, cd(0, 0));
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) {
      const cd x(A[2 * (i * m + j)], A[2 * (i * m + j) + 1]);
      const cd y(A[2 * (j * m + i)], A[2 * (j * m + i) + 1]);
      a[i * m + j] = 0.5 * (x + std::conj(y));
    }
  for (int i = 0; i < m; ++i) Q[i * m + i] = 1.0;
  std::vector<cd> v(m), p(m), wv(m);
  for (int k = 0; k + 2 < m; ++k) {
    double xn = 0;
    for (int i = k + 1; i < m; ++i) xn += std::norm(a[i * m + k]);
    xn = std::sqrt(xn);
    if (xn == 0.0) continue;
    const cd x0 = a[(k + 1) * m + k];
    const cd ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : cd(1, 0);
    const cd alpha = -ph * xn;
    std::fill(v.begin(), v.end(), cd(0, 0));
    for (int i = k + 1; i < m; ++i) v[i] = a[i * m + k];
    v[k + 1] -= alpha;
    double vn = 0;
    for (int i = k + 1; i < m; ++i) vn += std::norm(v[i]);
    vn = std::sqrt(vn);
    if (vn == 0.0) continue;
    for (int i = k + 1; i < m; ++i) v[i] /= vn;
    // p = A v (rows/cols k..m-1 touched by the reflector act on k+1..m-1)
    for (int i = k; i < m; ++i) {
      cd s(0, 0);
      for (int j = k + 1; j < m; ++j) s += a[i * m + j] * v[j];
      p[i] = s;
    }
    cd K(0, 0);
    for (int i = k + 1; i < m; ++i) K += std::conj(v[i]) * p[i];
    for (int i = k; i < m; ++i) wv[i] = 2.0 * p[i] - 2.0 * K * (i > k ? v[i] : cd(0, 0));
    // A <- A - v w^H - w v^H (v is zero at index k)
    for (int i = k; i < m; ++i)
      for (int j = k; j < m; ++j) {
        const cd vi = i > k ? v[i] : cd(0, 0), vj = j > k ? v[j] : cd(0, 0);
        a[i * m + j] -= vi * std::conj(wv[j]) + wv[i] * std::conj(vj);
      }
    // Q <- Q H
    for (int i = 0; i < m; ++i) {
      cd s(0, 0);
      for (int j = k + 1; j < m; ++j) s += Q[i * m + j] * v[j];
      for (int j = k + 1; j < m; ++j) Q[i * m + j] -= 2.0 * s * std::conj(v[j]);
    }
  }
  std::vector<double> d(m), e(m, 0.0);
  std::vector<cd> phase(m, cd(1, 0));
  for (int k = 0; k < m; ++k) d[k] = a[k * m + k].real();
  for (int k = 0; k + 1 < m; ++k) {
    const cd ek = a[(k + 1) * m + k];
    const double r = std::abs(ek);
    e[k] = r;
    phase[k + 1] = r > 0 ? phase[k] * ek / r : phase[k];
  }
  std::vector<double> z;
  tridiag_eigen(m, d, e, z);
  theta = d;
  S.assign(2 * (size_t)m * m, 0.0);
  for (int i = 0; i < m; ++i)
    for (int q = 0; q < m; ++q) {
      cd s(0, 0);
      for (int k = 0; k < m; ++k) s += Q[i * m + k] * phase[k] * z[(size_t)q * m + k];
      S[2 * ((size_t)i * m + q)] = s.real();
      S[2 * ((size_t)i * m + q) + 1] = s.imag();
    }
}

// ---------------------------------------------------------------------------
// small generalised Hermitian eigenproblem (Cholesky + cyclic complex Jacobi)
// ---------------------------------------------------------------------------
void herm_gen_eig(int n, const std::vector<double>& A, const std::vector<double>& B, std::vector<double>& mu,
                  std::vector<double>& Z) {
  using cd = std::complex<double>;
  auto at = [&](const std::vector<double>& M, int i, int j) { return cd(M[2 * (i * n + j)], M[2 * (i * n + j) + 1]); };
  std::vector<cd> a(n * n), bm(n * n), L(n * n, cd(0, 0));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      a[i * n + j] = 0.5 * (at(A, i, j) + std::conj(at(A, j, i)));
      bm[i * n + j] = 0.5 * (at(B, i, j) + std::conj(at(B, j, i)));
    }
  for (int j = 0; j < n; ++j) {
    double dj = bm[j * n + j].real();
    for (int k = 0; k < j; ++k) dj -= std::norm(L[j * n + k]);
    if (!(dj > 0)) throw NumericError("Rayleigh-Ritz mass Gram matrix is not positive definite");
    L[j * n + j] = std::sqrt(dj);
    for (int i = j + 1; i < n; ++i) {
      cd s = bm[i * n + j];
      for (int k = 0; k < j; ++k) s -= L[i * n + k] * std::conj(L[j * n + k]);
      L[i * n + j] = s / L[j * n + j];
    }
  }
  // W = L^-1 A (column by column), C = W L^-H = (L^-1 W^H)^H
  auto fwd = [&](std::vector<cd>& X) {  // X <- L^-1 X, X n x n row-major
    for (int c = 0; c < n; ++c)
      for (int i = 0; i < n; ++i) {
        cd s = X[i * n + c];
        for (int k = 0; k < i; ++k) s -= L[i * n + k] * X[k * n + c];
        X[i * n + c] = s / L[i * n + i];
      }
  };
  std::vector<cd> W = a;
  fwd(W);
  std::vector<cd> Wh(n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) Wh[i * n + j] = std::conj(W[j * n + i]);
  fwd(Wh);
  std::vector<cd> C(n * n), V(n * n, cd(0, 0));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) C[i * n + j] = std::conj(Wh[j * n + i]);
  for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
  double fro = 0;
  for (auto& x : C) fro += std::norm(x);
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += std::norm(C[p * n + q]);
    if (off <= 1e-32 * fro) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double g = std::abs(C[p * n + q]);
        if (g <= 1e-300) continue;
        const cd e = C[p * n + q] / g;  // phase; D = diag(1, conj(e)) makes (p,q) real
        for (int k = 0; k < n; ++k) C[k * n + q] *= std::conj(e);
        for (int k = 0; k < n; ++k) C[q * n + k] *= e;
        for (int k = 0; k < n; ++k) V[k * n + q] *= std::conj(e);
        const double app = C[p * n + p].real(), aqq = C[q * n + q].real();
        const double zeta = (aqq - app) / (2.0 * g);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(zeta * zeta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // columns
          const cd kp = C[k * n + p], kq = C[k * n + q];
          C[k * n + p] = c * kp - s * kq;
          C[k * n + q] = s * kp + c * kq;
          const cd vp = V[k * n + p], vq = V[k * n + q];
          V[k * n + p] = c * vp - s * vq;
          V[k * n + q] = s * vp + c * vq;
        }
        for (int k = 0; k < n; ++k) {  // rows
          const cd pk = C[p * n + k], qk = C[q * n + k];
          C[p * n + k] = c * pk - s * qk;
          C[q * n + k] = s * pk + c * qk;
        }
        C[p * n + q] = C[q * n + p] = 0.0;
      }
  }
  // Z = L^-H V (back substitution with L^H)
  std::vector<cd> Zc = V;
  for (int c = 0; c < n; ++c)
    for (int i = n - 1; i >= 0; --i) {
      cd s = Zc[i * n + c];
      for (int k = i + 1; k < n; ++k) s -= std::conj(L[k * n + i]) * Zc[k * n + c];
      Zc[i * n + c] = s / std::conj(L[i * n + i]);
    }
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return C[x * n + x].real() < C[y * n + y].real(); });
  mu.resize(n);
  Z.assign(2 * (size_t)n * n, 0.0);
  for (int q = 0; q < n; ++q) {
    mu[q] = C[idx[q] * n + idx[q]].real();
    for (int k = 0; k < n; ++k) {
      Z[2 * (k * n + q)] = Zc[k * n + idx[q]].real();
      Z[2 * (k * n + q) + 1] = Zc[k * n + idx[q]].imag();
    }
  }
}

// ---------------------------------------------------------------------------
// operator: assembly at unit omega + static prestress
// ---------------------------------------------------------------------------
struct BatchMem {
  DevBuf L, U, T, S, status, params;
  int pcols = 1;
  BatchBufs bufs(const Plan&, int B) {
    BatchBufs bb;
    bb.L = L.as<double2>();
    bb.U = U.as<double2>();
    bb.T = T.as<double2>();
    bb.S = S.as<double2>();
    bb.status = status.as<int>();
    bb.params = params.as<ProbParam>();
    bb.B = B;
    return bb;
  }
  void alloc(const Plan& P, int B, int p = 1) {
    pcols = p;
    L.alloc(sizeof(double2) * P.panel_entries * B);
    U.alloc(sizeof(double2) * P.upool_entries * B);
    T.alloc(sizeof(double2) * P.tvec_entries * B * p);
    S.alloc(sizeof(double2) * P.inv_scratch_entries * B);
    status.alloc(sizeof(int) * B);
    params.alloc(sizeof(ProbParam) * B);
  }
};

OpArrays op_arrays(const Operator& op) {
  OpArrays a;
  a.Kc = op.Kc.as<double>();
  a.Kp = op.Kp.as<double>();
  a.dKc = op.dKc.as<double>();
  a.dKp = op.dKp.as<double>();
  a.Mc = op.Mc.as<double>();
  a.Mp = op.Mp.as<double>();
  return a;
}

void create_operator(Context* ctx, Mesh& M, Operator& op) {
  const Symbolic& S = M.S;
  const size_t nblk = S.colidx.size();
  op.nblk = nblk;
  hipStream_t st = ctx->stream;
  DevBuf ke(sizeof(double) * 900 * (size_t)S.ne), m_abs(sizeof(double) * 100 * (size_t)S.ne),
      m_raw(sizeof(double) * 100 * (size_t)S.ne), fe(sizeof(double) * 30 * (size_t)S.ne),
      s_ab(sizeof(double) * 100 * (size_t)S.ne);
  launch_element_base(ctx, M, op.E, op.nu, op.rho, ke, m_abs, m_raw, fe);
  for (DevBuf* b : {&op.Kc, &op.Kp, &op.dKc, &op.dKp, &op.Mc, &op.Mp, &op.Gc, &op.Gp})
    b->alloc(sizeof(double) * 9 * nblk);
  launch_assemble(ctx, M, 0, false, ke, m_abs, m_raw, nullptr, op.Kc);
  launch_assemble(ctx, M, 0, true, ke, m_abs, m_raw, nullptr, op.Kp);
  launch_assemble(ctx, M, 1, false, ke, m_abs, m_raw, nullptr, op.Mc);
  launch_assemble(ctx, M, 1, true, ke, m_abs, m_raw, nullptr, op.Mp);
  launch_assemble(ctx, M, 3, false, ke, m_abs, m_raw, nullptr, op.Gc);
  launch_assemble(ctx, M, 3, true, ke, m_abs, m_raw, nullptr, op.Gp);
  TM_HIP(hipMemsetAsync(s_ab.p, 0, s_ab.bytes, st));
  op.u_static.assign(3 * (size_t)S.nn, 0.0);
  if (op.prestress) {
    // K u = F at unit omega with k = 0 cyclic BCs (base K only)
    TM_HIP(hipMemsetAsync(op.dKc.p, 0, op.dKc.bytes, st));
    TM_HIP(hipMemsetAsync(op.dKp.p, 0, op.dKp.bytes, st));
    BatchMem bm;
    bm.alloc(M.plan, 1);
    ProbParam pp{1.0, 0.0, 0.0, 0.0};
    TM_HIP(hipMemcpyAsync(bm.params.p, &pp, sizeof(pp), hipMemcpyHostToDevice, st));
    TM_HIP(hipMemsetAsync(bm.status.p, 0, sizeof(int), st));
    BatchBufs bb = bm.bufs(M.plan, 1);
    factor_batch(ctx, M, M.plan, op_arrays(op), bb);
    const long long n = 3LL * S.nf;
    DevBuf f(sizeof(double) * n), x(sizeof(double2) * n), ufull(sizeof(double) * 3 * S.nn);
    launch_assemble_load(ctx, M, fe, f);
    hipLaunchKernelGGL(k_real_to_complex, dim3((n + 255) / 256), dim3(256), 0, st, f.as<double>(),
                       x.as<double2>(), n);
    trisolve_batch(ctx, M, M.plan, bb, x.as<double2>(), n, nullptr);
    int status = 0;
    TM_HIP(hipMemcpyAsync(&status, bm.status.p, sizeof(int), hipMemcpyDeviceToHost, st));
    TM_HIP(hipStreamSynchronize(st));
    if (status) throw NumericError("zero pivot in the static centrifugal solve");
    launch_expand_real(ctx, M, x.as<double2>(), ufull);
    TM_HIP(hipMemcpyAsync(op.u_static.data(), ufull.p, ufull.bytes, hipMemcpyDeviceToHost, st));
    launch_element_sigma(ctx, M, op.E, op.nu, ufull, s_ab);
  }
  launch_assemble(ctx, M, 2, false, ke, m_abs, m_raw, &s_ab, op.dKc);
  launch_assemble(ctx, M, 2, true, ke, m_abs, m_raw, &s_ab, op.dKp);
  TM_HIP(hipStreamSynchronize(st));
}

const double2* harmonic_mass(Context* ctx, Operator& op, int k) {
  auto it = op.Mk_cache.find(k);
  if (it != op.Mk_cache.end()) return it->second.as<double2>();
  Mesh& M = *op.mesh;
  const Symbolic& S = M.S;
  DevBuf b(sizeof(double2) * 9 * op.nblk);
  double alpha = 2.0 * M_PI / S.N;
  hipLaunchKernelGGL(k_build_phase_blocks, dim3((S.nf + 127) / 128), dim3(128), 0, ctx->stream, S.nf,
                     M.d_rowptr.as<int>(), M.d_colidx.as<int>(), M.d_tslot.as<int>(),
                     M.d_pinned.as<unsigned char>(), op.Mc.as<double>(), op.Mp.as<double>(),
                     (const double*)nullptr, (const double*)nullptr, 0.0, std::cos(k * alpha),
                     std::sin(k * alpha), b.as<double2>());
  TM_HIP(hipGetLastError());
  const double2* p = b.as<double2>();
  op.Mk_cache.emplace(k, std::move(b));
  return p;
}

// ---------------------------------------------------------------------------
// batched Lanczos
// ---------------------------------------------------------------------------
struct HostProblem {
  tm_problem in;
  tm_result* out = nullptr;
  int k = 0;
  double omega = 0, s = 0, pre = 1, pim = 0;
  bool degenerate = false;
  int nev_target = 0, nev = 0, ncv = 0;
  int restarts = 0, max_restarts = 1000;
  bool done = false, converged = false, failed = false;
  int n_extract = 0;
  std::vector<double> theta_sorted_lambda;  // extracted eigenvalues
  std::vector<double> S_extract;            // ncv x n_extract (row-major k, c)
  std::vector<double> residual, xmx;        // per extracted pair
  int refinements = 0;
  int ncv_final = 0;
  std::string message;
};

static void fill_policy(const Symbolic& S, HostProblem& hp) {
  const int N = S.N, max_k = N / 2;
  hp.k = hp.in.harmonic;
  const double alpha = 2.0 * M_PI / N;
  hp.pre = std::cos(hp.k * alpha);
  hp.pim = std::sin(hp.k * alpha);
  hp.omega = hp.in.rpm * 2.0 * M_PI / 60.0;
  hp.s = hp.omega * hp.omega;
  hp.degenerate = hp.k > 0 && !(N % 2 == 0 && hp.k == max_k);
  const int M = hp.in.num_modes;
  hp.nev_target = hp.degenerate ? (M + 1) / 2 : M;
  const int n_free = S.n_free_dofs;
  const int extra = hp.in.min_frequency > 0.0 ? 6 : 0;
  hp.nev = std::min(hp.nev_target + extra, n_free - 1);
  if (n_free <= 200) hp.ncv = n_free;  // full Krylov space == the reference's dense solve
  else hp.ncv = std::min(std::max(4 * hp.nev + 1, 40), n_free);
  hp.max_restarts = hp.in.max_restarts > 0 ? hp.in.max_restarts : 1000;
}

class BatchLanczos {
 public:
  BatchLanczos(Context* ctx, Operator& op, std::vector<HostProblem*>& probs)
      : ctx_(ctx), op_(op), M_(*op.mesh), S_(M_.S), P_(M_.plan), probs_(probs) {}

  void run() {
    auto t0 = std::chrono::steady_clock::now();
    setup();
    factor();
    start_vectors();
    iterate();
    extract();
    double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (auto* hp : probs_)
      if (hp->out) hp->out->seconds = sec;
  }

 private:
  Context* ctx_;
  Operator& op_;
  Mesh& M_;
  const Symbolic& S_;
  const Plan& P_;
  std::vector<HostProblem*>& probs_;
  int B_ = 0, ncvmax_ = 0, kmax_ = 0, nevmax_ = 0, nchunk_ = 0;
  long long n_ = 0, ldb_ = 0;
  BatchMem bm_;
  DevBuf V_, MV_, W_, MW_, h_, part_, alpha_, beta_, bnew_, jcur_, ncv_, live_, frozen_, act_;
  DevBuf mk_ptrs_, tmp_, coef_, nk_, nc_;
  LzState st_{};
  std::vector<const double2*> mk_;
  std::vector<int> h_jcur_, h_ncv_, h_live_;

  hipStream_t stream() const { return ctx_->stream; }

  void setup() {
    B_ = static_cast<int>(probs_.size());
    n_ = 3LL * S_.nf;
    ncvmax_ = 0;
    nevmax_ = 0;
    for (auto* hp : probs_) {
      ncvmax_ = std::max(ncvmax_, hp->ncv);
      nevmax_ = std::max(nevmax_, hp->nev);
    }
    kmax_ = ncvmax_ + 1;
    ldb_ = (long long)(ncvmax_ + 1) * n_;
    nchunk_ = static_cast<int>((n_ + kDotChunk - 1) / kDotChunk);
    bm_.alloc(P_, B_);
    V_.alloc(sizeof(double2) * ldb_ * B_);
    MV_.alloc(sizeof(double2) * ldb_ * B_);
    W_.alloc(sizeof(double2) * n_ * B_);
    MW_.alloc(sizeof(double2) * n_ * B_);
    h_.alloc(sizeof(double2) * kmax_ * B_);
    part_.alloc(sizeof(double2) * (size_t)kmax_ * nchunk_ * B_);
    alpha_.alloc(sizeof(double) * ncvmax_ * B_);
    beta_.alloc(sizeof(double) * ncvmax_ * B_);
    bnew_.alloc(sizeof(double) * B_);
    for (DevBuf* b : {&jcur_, &ncv_, &live_, &frozen_, &act_}) b->alloc(sizeof(int) * B_);
    nk_.alloc(sizeof(int) * B_);
    nc_.alloc(sizeof(int) * B_);
    // params and M(k)
    std::vector<ProbParam> pp(B_);
    mk_.resize(B_);
    h_ncv_.resize(B_);
    for (int b = 0; b < B_; ++b) {
      HostProblem* hp = probs_[b];
      pp[b] = {hp->pre, hp->pim, hp->s, hp->in.shift};
      mk_[b] = harmonic_mass(ctx_, op_, hp->k);
      h_ncv_[b] = hp->ncv;
    }
    TM_HIP(hipMemcpyAsync(bm_.params.p, pp.data(), sizeof(ProbParam) * B_, hipMemcpyHostToDevice, stream()));
    TM_HIP(hipMemsetAsync(bm_.status.p, 0, sizeof(int) * B_, stream()));
    mk_ptrs_ = upload(mk_, stream());
    TM_HIP(hipMemcpyAsync(ncv_.p, h_ncv_.data(), sizeof(int) * B_, hipMemcpyHostToDevice, stream()));
    TM_HIP(hipMemsetAsync(alpha_.p, 0, alpha_.bytes, stream()));
    TM_HIP(hipMemsetAsync(beta_.p, 0, beta_.bytes, stream()));
    st_.V = V_.as<double2>();
    st_.MV = MV_.as<double2>();
    st_.W = W_.as<double2>();
    st_.MW = MW_.as<double2>();
    st_.h = h_.as<double2>();
    st_.part = part_.as<double2>();
    st_.alpha = alpha_.as<double>();
    st_.beta = beta_.as<double>();
    st_.bnew = bnew_.as<double>();
    st_.jcur = jcur_.as<int>();
    st_.ncv = ncv_.as<int>();
    st_.live = live_.as<int>();
    st_.frozen = frozen_.as<int>();
    st_.act = act_.as<int>();
    st_.ldb = ldb_;
    st_.ncvmax = ncvmax_;
    st_.nchunk = nchunk_;
  }

  void factor() {
    BatchBufs bb = bm_.bufs(P_, B_);
    factor_batch(ctx_, M_, P_, op_arrays(op_), bb);
    std::vector<int> status(B_);
    TM_HIP(hipMemcpyAsync(status.data(), bm_.status.p, sizeof(int) * B_, hipMemcpyDeviceToHost, stream()));
    TM_HIP(hipStreamSynchronize(stream()));
    for (int b = 0; b < B_; ++b)
      if (status[b]) {
        probs_[b]->failed = true;
        probs_[b]->message = "zero pivot in the shift-invert factorization";
      }
  }

  void set_host_state(const std::vector<int>& jcur, const std::vector<int>& live,
                      const std::vector<int>& frozen) {
    TM_HIP(hipMemcpyAsync(jcur_.p, jcur.data(), sizeof(int) * B_, hipMemcpyHostToDevice, stream()));
    TM_HIP(hipMemcpyAsync(live_.p, live.data(), sizeof(int) * B_, hipMemcpyHostToDevice, stream()));
    TM_HIP(hipMemcpyAsync(frozen_.p, frozen.data(), sizeof(int) * B_, hipMemcpyHostToDevice, stream()));
    TM_HIP(hipStreamSynchronize(stream()));
    hipLaunchKernelGGL(k_update_act, dim3((B_ + 63) / 64), dim3(64), 0, stream(), st_, B_);
  }

  // y = M(k) x for one column per problem
  void spmv_m(const double2* X, long long ldx, double2* Y, long long ldy, const int* act) {
    KTimer t(ctx_, "spmv_m", 0, 0);
    hipLaunchKernelGGL(k_spmv_bsr, dim3((3 * S_.nf + 255) / 256, B_), dim3(256), 0, stream(), S_.nf,
                       M_.d_rowptr.as<int>(), M_.d_colidx.as<int>(), mk_ptrs_.as<const double2*>(), X,
                       ldx, Y, ldy, act);
  }

  void mdot(int mode) {
    KTimer t(ctx_, "mdot", 0, 0);
    hipLaunchKernelGGL(k_mdot, dim3(nchunk_, B_), dim3(256), 0, stream(), st_, n_, mode, kmax_);
    hipLaunchKernelGGL(k_mdot_reduce, dim3(B_), dim3(64), 0, stream(), st_, mode, kmax_);
  }

  std::vector<double> random_vector(uint32_t seed) {
    std::mt19937 gen(seed);
    std::normal_distribution<double> dist(0.0, 1.0);
    std::vector<double> v(2 * n_);
    for (long long i = 0; i < n_; ++i) {
      double re = dist(gen);
      double im = dist(gen);
      const bool pin = S_.pinned[i];
      v[2 * i] = pin ? 0.0 : re;
      v[2 * i + 1] = pin ? 0.0 : im;
    }
    return v;
  }

  void start_vectors() {
    std::vector<double> v0 = random_vector(42);
    for (int b = 0; b < B_; ++b)
      TM_HIP(hipMemcpyAsync(V_.as<double2>() + b * ldb_, v0.data(), sizeof(double2) * n_,
                            hipMemcpyHostToDevice, stream()));
    TM_HIP(hipStreamSynchronize(stream()));
    spmv_m(V_.as<double2>(), ldb_, MV_.as<double2>(), ldb_, nullptr);
    // norms via the beta path: W = V0, MW = MV0
    for (int b = 0; b < B_; ++b) {
      TM_HIP(hipMemcpyAsync(W_.as<double2>() + b * n_, V_.as<double2>() + b * ldb_, sizeof(double2) * n_,
                            hipMemcpyDeviceToDevice, stream()));
      TM_HIP(hipMemcpyAsync(MW_.as<double2>() + b * n_, MV_.as<double2>() + b * ldb_,
                            sizeof(double2) * n_, hipMemcpyDeviceToDevice, stream()));
    }
    std::vector<int> zero(B_, 0), one(B_, 1), mone(B_, -1);
    // jcur = -1 so that k_normalize_next writes column 0
    set_host_state(mone, one, zero);
    mdot(2);
    hipLaunchKernelGGL(k_finish_beta_start, dim3((B_ + 63) / 64), dim3(64), 0, stream(), st_, B_, kmax_);
    hipLaunchKernelGGL(k_normalize_next, dim3((n_ + 255) / 256, B_), dim3(256), 0, stream(), st_, n_);
    h_jcur_.assign(B_, 0);
    h_live_.assign(B_, 1);
    for (int b = 0; b < B_; ++b)
      if (probs_[b]->failed) h_live_[b] = 0;
    set_host_state(h_jcur_, h_live_, zero);
  }

  void step() {
    hipLaunchKernelGGL(k_update_act, dim3((B_ + 63) / 64), dim3(64), 0, stream(), st_, B_);
    {
      KTimer t(ctx_, "vec", 0, 0);
      hipLaunchKernelGGL(k_copy_mv, dim3((n_ + 255) / 256, B_), dim3(256), 0, stream(), st_, n_);
    }
    BatchBufs bb = bm_.bufs(P_, B_);
    trisolve_batch(ctx_, M_, P_, bb, W_.as<double2>(), n_, act_.as<int>());
    mdot(1);  // alpha_j
    {
      KTimer t(ctx_, "vec", 0, 0);
      hipLaunchKernelGGL(k_three_term, dim3((n_ + 255) / 256, B_), dim3(256), 0, stream(), st_, n_);
    }
    for (int pass = 0; pass < 2; ++pass) {
      mdot(0);
      KTimer t(ctx_, "reorth", 0, 0);
      hipLaunchKernelGGL(k_axpy_multi, dim3((n_ + 255) / 256, B_), dim3(256), 0, stream(), st_, n_, kmax_);
    }
    spmv_m(W_.as<double2>(), n_, MW_.as<double2>(), n_, act_.as<int>());
    mdot(2);
    hipLaunchKernelGGL(k_finish_beta, dim3((B_ + 63) / 64), dim3(64), 0, stream(), st_, B_, kmax_);
    {
      KTimer t(ctx_, "vec", 0, 0);
      hipLaunchKernelGGL(k_normalize_next, dim3((n_ + 255) / 256, B_), dim3(256), 0, stream(), st_, n_);
    }
    hipLaunchKernelGGL(k_advance, dim3((B_ + 63) / 64), dim3(64), 0, stream(), st_, B_);
  }

  // Breakdown at step j: a fresh random vector, orthogonalised (twice)
  // against V[:, 0..j], starts a new Krylov block.  Its coupling to the old
  // block is zero, so beta_j = 0 in T (the reference stores the new vector's
  // norm there, src/hermitian_lanczos.cpp:350-369).  If nothing survives the
  // orthogonalisation the basis spans an invariant subspace: truncate.
  bool recover_breakdown(int b, int j, int restart) {
    std::vector<double> v = random_vector(42 + restart * 100 + j);
    TM_HIP(hipMemcpyAsync(W_.as<double2>() + b * n_, v.data(), sizeof(double2) * n_,
                          hipMemcpyHostToDevice, stream()));
    std::vector<int> jc(B_, 0), lv(B_, 0), fz(B_, 0);
    jc[b] = j;
    lv[b] = 1;
    set_host_state(jc, lv, fz);
    spmv_m(W_.as<double2>(), n_, MW_.as<double2>(), n_, act_.as<int>());
    mdot(2);
    double2 n0;
    TM_HIP(hipMemcpyAsync(&n0, h_.as<double2>() + (size_t)b * kmax_, sizeof(double2), hipMemcpyDeviceToHost, stream()));
    TM_HIP(hipStreamSynchronize(stream()));
    const double norm0 = std::sqrt(std::fabs(n0.x));
    for (int pass = 0; pass < 2; ++pass) {
      mdot(0);
      hipLaunchKernelGGL(k_axpy_multi, dim3((n_ + 255) / 256, B_), dim3(256), 0, stream(), st_, n_, kmax_);
    }
    spmv_m(W_.as<double2>(), n_, MW_.as<double2>(), n_, act_.as<int>());
    mdot(2);
    hipLaunchKernelGGL(k_finish_beta_start, dim3((B_ + 63) / 64), dim3(64), 0, stream(), st_, B_, kmax_);
    double bn = 0.0;
    TM_HIP(hipMemcpyAsync(&bn, bnew_.as<double>() + b, sizeof(double), hipMemcpyDeviceToHost, stream()));
    TM_HIP(hipStreamSynchronize(stream()));
    TM_HIP(hipMemsetAsync(beta_.as<double>() + (size_t)b * ncvmax_ + j, 0, sizeof(double), stream()));
    if (!(bn > 1e-10 * norm0)) return false;
    hipLaunchKernelGGL(k_normalize_next, dim3((n_ + 255) / 256, B_), dim3(256), 0, stream(), st_, n_);
    TM_HIP(hipStreamSynchronize(stream()));
    return true;
  }

  void iterate() {
    std::vector<int> frozen(B_, 0);
    std::vector<double> alpha((size_t)ncvmax_ * B_), beta((size_t)ncvmax_ * B_);
    while (true) {
      int steps = 0;
      for (int b = 0; b < B_; ++b)
        if (h_live_[b]) steps = std::max(steps, h_ncv_[b] - h_jcur_[b]);
      if (steps <= 0) break;
      set_host_state(h_jcur_, h_live_, frozen);
      for (int t = 0; t < steps; ++t) step();
      ctx_->flush_timing();
      TM_HIP(hipMemcpyAsync(h_jcur_.data(), jcur_.p, sizeof(int) * B_, hipMemcpyDeviceToHost, stream()));
      TM_HIP(hipMemcpyAsync(frozen.data(), frozen_.p, sizeof(int) * B_, hipMemcpyDeviceToHost, stream()));
      TM_HIP(hipStreamSynchronize(stream()));
      // breakdowns
      for (int b = 0; b < B_; ++b) {
        if (!h_live_[b] || !frozen[b]) continue;
        HostProblem* hp = probs_[b];
        const int j = h_jcur_[b];
        frozen[b] = 0;
        TM_HIP(hipMemsetAsync(frozen_.as<int>() + b, 0, sizeof(int), stream()));
        if (recover_breakdown(b, j, hp->restarts)) {
          h_jcur_[b] = j + 1;
        } else {
          // truly degenerate: truncate the pass at j + 1 (src/hermitian_lanczos.cpp:362-366)
          h_ncv_[b] = j + 1;
          h_jcur_[b] = j + 1;
          TM_HIP(hipMemsetAsync(frozen_.as<int>() + b, 0, sizeof(int), stream()));
          TM_HIP(hipMemsetAsync(beta_.as<double>() + (size_t)b * ncvmax_ + j, 0, sizeof(double), stream()));
        }
      }
      TM_HIP(hipMemcpyAsync(ncv_.p, h_ncv_.data(), sizeof(int) * B_, hipMemcpyHostToDevice, stream()));
      TM_HIP(hipMemcpyAsync(alpha.data(), alpha_.p, sizeof(double) * alpha.size(), hipMemcpyDeviceToHost, stream()));
      TM_HIP(hipMemcpyAsync(beta.data(), beta_.p, sizeof(double) * beta.size(), hipMemcpyDeviceToHost, stream()));
      TM_HIP(hipStreamSynchronize(stream()));
      std::vector<int> restart_list;
      for (int b = 0; b < B_; ++b) {
        if (!h_live_[b] || h_jcur_[b] < h_ncv_[b]) continue;
        if (check_pass(b, &alpha[(size_t)b * ncvmax_], &beta[(size_t)b * ncvmax_])) {
          h_live_[b] = 0;
        } else {
          restart_list.push_back(b);
        }
      }
      if (!restart_list.empty()) implicit_restart(restart_list, alpha, beta);
    }
  }

  // convergence check of one completed pass; true when the problem is finished
  bool check_pass(int b, const double* a, const double* bt) {
    HostProblem* hp = probs_[b];
    const int ncv = h_ncv_[b];
    std::vector<double> d(a, a + ncv), e(bt, bt + ncv), z;
    tridiag_eigen(ncv, d, e, z);
    const double sigma = hp->in.shift;
    std::vector<double> lam(ncv);
    for (int i = 0; i < ncv; ++i) lam[i] = std::fabs(d[i]) > 1e-15 ? sigma + 1.0 / d[i] : 1e30;
    std::vector<int> idx(ncv);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return lam[x] < lam[y]; });
    const double bl = bt[ncv - 1];
    const double tol_rel = hp->in.tol_rel > 0 ? hp->in.tol_rel : 1e-12;
    int nconv = 0;
    const int nev = std::min(hp->nev, ncv);
    for (int i = 0; i < nev; ++i) {
      const int ii = idx[i];
      const double res = std::fabs(bl * z[(size_t)ii * ncv + (ncv - 1)]);
      bool ok;
      if (hp->in.conv_policy == TM_CONV_REFERENCE) ok = res < 1e-8 * std::max(1.0, std::fabs(d[ii]));
      else ok = res <= tol_rel * std::fabs(d[ii]);
      nconv += ok ? 1 : 0;
    }
    if (nconv >= hp->nev || hp->restarts >= hp->max_restarts) {
      int ne = std::min(nconv > 0 ? nconv : hp->nev, ncv);
      ne = std::min(ne, hp->nev);
      hp->n_extract = ne;
      hp->converged = nconv >= hp->nev;
      hp->ncv_final = ncv;
      hp->theta_sorted_lambda.resize(ne);
      hp->S_extract.assign((size_t)ncv * ne, 0.0);
      for (int c = 0; c < ne; ++c) {
        hp->theta_sorted_lambda[c] = lam[idx[c]];
        for (int k = 0; k < ncv; ++k) hp->S_extract[(size_t)k * ne + c] = z[(size_t)idx[c] * ncv + k];
      }
      hp->done = true;
      return true;
    }
    return false;
  }

  // exact-shift implicit QR restart (src/hermitian_lanczos.cpp:114-260)
  void implicit_restart(const std::vector<int>& list, std::vector<double>& alpha, std::vector<double>& beta) {
    const int ldc = nevmax_ + 1;
    std::vector<double> coef((size_t)B_ * (ncvmax_ + 1) * ldc, 0.0);
    std::vector<int> nk(B_, 0), nc(B_, 0);
    std::vector<double> bt_tri(B_, 0.0);
    for (int b : list) {
      HostProblem* hp = probs_[b];
      hp->restarts += 1;
      const int ncv = h_ncv_[b], nev = hp->nev;
      double* a = &alpha[(size_t)b * ncvmax_];
      double* bt = &beta[(size_t)b * ncvmax_];
      std::vector<double> d(a, a + ncv), e(bt, bt + ncv), z;
      tridiag_eigen(ncv, d, e, z);
      std::vector<double> Q((size_t)ncv * ncv, 0.0);  // column-major
      for (int i = 0; i < ncv; ++i) Q[i + (size_t)i * ncv] = 1.0;
      for (int sh = 0; sh < ncv - nev; ++sh) {
        const double mu = d[sh];  // ascending: the ncv - nev smallest Ritz values
        double x = a[0] - mu, zz = bt[0];
        for (int i = 0; i < ncv - 1; ++i) {
          double r = std::hypot(x, zz), c, s;
          if (r < 1e-300) { c = 1.0; s = 0.0; }
          else { c = x / r; s = zz / r; }
          if (i > 0) bt[i - 1] = r;
          const double a0 = a[i], a1 = a[i + 1], b0 = bt[i];
          a[i] = c * c * a0 + 2.0 * c * s * b0 + s * s * a1;
          a[i + 1] = s * s * a0 - 2.0 * c * s * b0 + c * c * a1;
          bt[i] = c * s * (a1 - a0) + (c * c - s * s) * b0;
          if (i + 2 < ncv) {
            const double bulge = s * bt[i + 1];
            bt[i + 1] = c * bt[i + 1];
            x = bt[i];
            zz = bulge;
          }
          for (int row = 0; row < ncv; ++row) {
            double q0 = Q[row + (size_t)i * ncv], q1 = Q[row + (size_t)(i + 1) * ncv];
            Q[row + (size_t)i * ncv] = c * q0 + s * q1;
            Q[row + (size_t)(i + 1) * ncv] = -s * q0 + c * q1;
          }
        }
      }
      const double beta_tri = bt[nev - 1];
      const double beta_dangling = bt[ncv - 1];
      double* cb = &coef[(size_t)b * (ncvmax_ + 1) * ldc];
      for (int k = 0; k < ncv; ++k) {
        for (int c = 0; c < nev; ++c) cb[(size_t)k * ldc + c] = Q[k + (size_t)c * ncv];
        if (nev < ncv) cb[(size_t)k * ldc + nev] = beta_tri * Q[k + (size_t)nev * ncv];
      }
      if (nev < ncv) cb[(size_t)ncv * ldc + nev] = beta_dangling * Q[(ncv - 1) + (size_t)(nev - 1) * ncv];
      nk[b] = ncv + 1;
      nc[b] = nev < ncv ? nev + 1 : nev;
    }
    if (tmp_.bytes < sizeof(double2) * (size_t)(nevmax_ + 1) * n_ * B_)
      tmp_.alloc(sizeof(double2) * (size_t)(nevmax_ + 1) * n_ * B_);
    coef_ = upload(coef, stream());
    TM_HIP(hipMemcpyAsync(nk_.p, nk.data(), sizeof(int) * B_, hipMemcpyHostToDevice, stream()));
    TM_HIP(hipMemcpyAsync(nc_.p, nc.data(), sizeof(int) * B_, hipMemcpyHostToDevice, stream()));
    const long long ldo = (long long)(nevmax_ + 1) * n_;
    hipLaunchKernelGGL(k_combine, dim3((n_ + 255) / 256, B_), dim3(256), 0, stream(), V_.as<double2>(), ldb_,
                       n_, coef_.as<double>(), ldc, (long long)(ncvmax_ + 1) * ldc, nk_.as<int>(),
                       nc_.as<int>(), tmp_.as<double2>(), ldo);
    std::vector<int> ncolv(B_, 0), jc(h_jcur_), lv(B_, 0), fz(B_, 0);
    for (int b : list) {
      HostProblem* hp = probs_[b];
      const int nev = hp->nev, ncv = h_ncv_[b];
      TM_HIP(hipMemcpyAsync(V_.as<double2>() + b * ldb_, tmp_.as<double2>() + b * ldo,
                            sizeof(double2) * n_ * nev, hipMemcpyDeviceToDevice, stream()));
      ncolv[b] = nev;
      if (nev < ncv) {
        TM_HIP(hipMemcpyAsync(W_.as<double2>() + b * n_, tmp_.as<double2>() + b * ldo + (long long)nev * n_,
                              sizeof(double2) * n_, hipMemcpyDeviceToDevice, stream()));
        lv[b] = 1;
        jc[b] = nev - 1;  // normalize_next writes column nev
      }
    }
    // MV[:, 0:nev] = M V[:, 0:nev]
    DevBuf ncol = upload(ncolv, stream());
    hipLaunchKernelGGL(k_spmv_multi, dim3((3 * S_.nf + 255) / 256, B_, nevmax_), dim3(256), 0, stream(), S_.nf,
                       M_.d_rowptr.as<int>(), M_.d_colidx.as<int>(), mk_ptrs_.as<const double2*>(),
                       V_.as<double2>(), ldb_, n_, ncol.as<int>(), MV_.as<double2>());
    // continuation vector f: beta_new = ||f||_M
    set_host_state(jc, lv, fz);
    spmv_m(W_.as<double2>(), n_, MW_.as<double2>(), n_, act_.as<int>());
    mdot(2);
    hipLaunchKernelGGL(k_finish_beta_start, dim3((B_ + 63) / 64), dim3(64), 0, stream(), st_, B_, kmax_);
    std::vector<double> bn(B_, 0.0);
    TM_HIP(hipMemcpyAsync(bn.data(), bnew_.p, sizeof(double) * B_, hipMemcpyDeviceToHost, stream()));
    TM_HIP(hipStreamSynchronize(stream()));
    std::vector<int> lv2(B_, 0);
    for (int b : list)
      if (lv[b] && bn[b] > 1e-14) lv2[b] = 1;
    set_host_state(jc, lv2, fz);
    hipLaunchKernelGGL(k_normalize_next, dim3((n_ + 255) / 256, B_), dim3(256), 0, stream(), st_, n_);
    for (int b : list) {
      HostProblem* hp = probs_[b];
      const int nev = hp->nev;
      beta[(size_t)b * ncvmax_ + nev - 1] = lv2[b] ? bn[b] : 0.0;
      h_jcur_[b] = nev;
    }
    TM_HIP(hipMemcpyAsync(alpha_.p, alpha.data(), sizeof(double) * alpha.size(), hipMemcpyHostToDevice, stream()));
    TM_HIP(hipMemcpyAsync(beta_.p, beta.data(), sizeof(double) * beta.size(), hipMemcpyHostToDevice, stream()));
    TM_HIP(hipStreamSynchronize(stream()));
  }

  static constexpr double kRefineTol = 5e-9;  // keeps a 2x margin under the 1e-8 bar

  // residual, ||K x||, x^H M x per Ritz vector (fixed-order chunk reduction)
  void residuals(DevBuf& X, DevBuf& KX, DevBuf& MX, long long ldx, int nemax, const std::vector<int>& nc,
                 const DevBuf& kp) {
    DevBuf dnc = upload(nc, stream());
    hipLaunchKernelGGL(k_spmv_multi, dim3((3 * S_.nf + 255) / 256, B_, nemax), dim3(256), 0, stream(), S_.nf,
                       M_.d_rowptr.as<int>(), M_.d_colidx.as<int>(), kp.as<const double2*>(),
                       X.as<double2>(), ldx, n_, dnc.as<int>(), KX.as<double2>());
    hipLaunchKernelGGL(k_spmv_multi, dim3((3 * S_.nf + 255) / 256, B_, nemax), dim3(256), 0, stream(), S_.nf,
                       M_.d_rowptr.as<int>(), M_.d_colidx.as<int>(), mk_ptrs_.as<const double2*>(),
                       X.as<double2>(), ldx, n_, dnc.as<int>(), MX.as<double2>());
    std::vector<double> lam((size_t)B_ * nemax, 0.0);
    for (int b = 0; b < B_; ++b)
      for (int c = 0; c < nc[b]; ++c) lam[(size_t)b * nemax + c] = probs_[b]->theta_sorted_lambda[c];
    DevBuf dl = upload(lam, stream());
    DevBuf part(sizeof(double) * 3 * (size_t)B_ * nemax * nchunk_);
    hipLaunchKernelGGL(k_residual_partial, dim3(nchunk_, B_, nemax), dim3(256), 0, stream(), X.as<double2>(),
                       KX.as<double2>(), MX.as<double2>(), ldx, n_, dl.as<double>(), nemax, dnc.as<int>(),
                       part.as<double>(), nchunk_);
    std::vector<double> hp_part(3 * (size_t)B_ * nemax * nchunk_);
    TM_HIP(hipMemcpyAsync(hp_part.data(), part.p, part.bytes, hipMemcpyDeviceToHost, stream()));
    TM_HIP(hipStreamSynchronize(stream()));
    for (int b = 0; b < B_; ++b) {
      HostProblem* hp = probs_[b];
      hp->residual.assign(nc[b], 0.0);
      hp->xmx.assign(nc[b], 0.0);
      for (int c = 0; c < nc[b]; ++c) {
        double num = 0, den = 0, xm = 0;
        for (int q = 0; q < nchunk_; ++q) {
          const double* pq = &hp_part[(((size_t)b * nemax + c) * nchunk_ + q) * 3];
          num += pq[0];
          den += pq[1];
          xm += pq[2];
        }
        hp->residual[c] = den > 0 ? std::sqrt(num / den) : std::sqrt(num);
        hp->xmx[c] = xm;
      }
    }
  }

  // block inverse iteration + Rayleigh-Ritz for the problems with nr[b] > 0
  void refine(DevBuf& X, DevBuf& Y, DevBuf& KY, DevBuf& MY, long long ldx, int nemax, const std::vector<int>& nr,
              const DevBuf& kp) {
    DevBuf dnr = upload(nr, stream());
    // Y = M X, then solve in place column by column
    hipLaunchKernelGGL(k_spmv_multi, dim3((3 * S_.nf + 255) / 256, B_, nemax), dim3(256), 0, stream(), S_.nf,
                       M_.d_rowptr.as<int>(), M_.d_colidx.as<int>(), mk_ptrs_.as<const double2*>(),
                       X.as<double2>(), ldx, n_, dnr.as<int>(), Y.as<double2>());
    BatchBufs bb = bm_.bufs(P_, B_);
    for (int q = 0; q < nemax; ++q) {
      hipLaunchKernelGGL(k_flag_cols, dim3((B_ + 63) / 64), dim3(64), 0, stream(), dnr.as<int>(), q, B_,
                         act_.as<int>());
      trisolve_batch(ctx_, M_, P_, bb, Y.as<double2>() + (long long)q * n_, ldx, act_.as<int>());
    }
    hipLaunchKernelGGL(k_spmv_multi, dim3((3 * S_.nf + 255) / 256, B_, nemax), dim3(256), 0, stream(), S_.nf,
                       M_.d_rowptr.as<int>(), M_.d_colidx.as<int>(), kp.as<const double2*>(),
                       Y.as<double2>(), ldx, n_, dnr.as<int>(), KY.as<double2>());
    hipLaunchKernelGGL(k_spmv_multi, dim3((3 * S_.nf + 255) / 256, B_, nemax), dim3(256), 0, stream(), S_.nf,
                       M_.d_rowptr.as<int>(), M_.d_colidx.as<int>(), mk_ptrs_.as<const double2*>(),
                       Y.as<double2>(), ldx, n_, dnr.as<int>(), MY.as<double2>());
    const size_t gsz = (size_t)B_ * nemax * nemax * nchunk_;
    DevBuf pk(sizeof(double2) * gsz), pm(sizeof(double2) * gsz);
    hipLaunchKernelGGL(k_gram_partial, dim3(nchunk_, B_, nemax * nemax), dim3(256), 0, stream(), Y.as<double2>(),
                       KY.as<double2>(), ldx, n_, dnr.as<int>(), nemax, pk.as<double2>(), nchunk_);
    hipLaunchKernelGGL(k_gram_partial, dim3(nchunk_, B_, nemax * nemax), dim3(256), 0, stream(), Y.as<double2>(),
                       MY.as<double2>(), ldx, n_, dnr.as<int>(), nemax, pm.as<double2>(), nchunk_);
    std::vector<double2> hk(gsz), hm(gsz);
    TM_HIP(hipMemcpyAsync(hk.data(), pk.p, pk.bytes, hipMemcpyDeviceToHost, stream()));
    TM_HIP(hipMemcpyAsync(hm.data(), pm.p, pm.bytes, hipMemcpyDeviceToHost, stream()));
    TM_HIP(hipStreamSynchronize(stream()));
    std::vector<double2> zc((size_t)B_ * nemax * nemax, make_double2(0.0, 0.0));
    for (int b = 0; b < B_; ++b) {
      const int ne = nr[b];
      if (!ne) continue;
      std::vector<double> A(2 * ne * ne), Bm(2 * ne * ne), mu, Z;
      for (int p = 0; p < ne; ++p)
        for (int q = 0; q < ne; ++q) {
          double ar = 0, ai = 0, br = 0, bi = 0;
          for (int c = 0; c < nchunk_; ++c) {
            const size_t o = (((size_t)b * nemax + p) * nemax + q) * nchunk_ + c;
            ar += hk[o].x; ai += hk[o].y;
            br += hm[o].x; bi += hm[o].y;
          }
          A[2 * (p * ne + q)] = ar; A[2 * (p * ne + q) + 1] = ai;
          Bm[2 * (p * ne + q)] = br; Bm[2 * (p * ne + q) + 1] = bi;
        }
      herm_gen_eig(ne, A, Bm, mu, Z);
      HostProblem* hp = probs_[b];
      for (int q = 0; q < ne; ++q) hp->theta_sorted_lambda[q] = mu[q];
      for (int k = 0; k < ne; ++k)
        for (int q = 0; q < ne; ++q)
          zc[((size_t)b * nemax + k) * nemax + q] = make_double2(Z[2 * (k * ne + q)], Z[2 * (k * ne + q) + 1]);
    }
    DevBuf dz = upload(zc, stream());
    hipLaunchKernelGGL(k_combine_c, dim3((n_ + 255) / 256, B_), dim3(256), 0, stream(), Y.as<double2>(), ldx, n_,
                       dz.as<double2>(), nemax, (long long)nemax * nemax, dnr.as<int>(), X.as<double2>(), ldx);
  }

  void extract() {
    // Ritz vectors X = V S for every finished problem
    int nemax = 0;
    for (auto* hp : probs_) nemax = std::max(nemax, hp->n_extract);
    if (nemax == 0) {
      finish_failed();
      return;
    }
    const int ldc = nemax;
    std::vector<double> coef((size_t)B_ * (ncvmax_ + 1) * ldc, 0.0);
    std::vector<int> nk(B_, 0), nc(B_, 0);
    for (int b = 0; b < B_; ++b) {
      HostProblem* hp = probs_[b];
      if (!hp->done || hp->failed) continue;
      const int ne = hp->n_extract, ncv = hp->ncv_final;
      for (int k = 0; k < ncv; ++k)
        for (int c = 0; c < ne; ++c)
          coef[(size_t)b * (ncvmax_ + 1) * ldc + (size_t)k * ldc + c] = hp->S_extract[(size_t)k * ne + c];
      nk[b] = ncv;
      nc[b] = ne;
    }
    const long long ldx = (long long)nemax * n_;
    DevBuf X(sizeof(double2) * ldx * B_), KX(sizeof(double2) * ldx * B_), MX(sizeof(double2) * ldx * B_);
    coef_ = upload(coef, stream());
    TM_HIP(hipMemcpyAsync(nk_.p, nk.data(), sizeof(int) * B_, hipMemcpyHostToDevice, stream()));
    TM_HIP(hipMemcpyAsync(nc_.p, nc.data(), sizeof(int) * B_, hipMemcpyHostToDevice, stream()));
    hipLaunchKernelGGL(k_combine, dim3((n_ + 255) / 256, B_), dim3(256), 0, stream(), V_.as<double2>(), ldb_,
                       n_, coef_.as<double>(), ldc, (long long)(ncvmax_ + 1) * ldc, nk_.as<int>(),
                       nc_.as<int>(), X.as<double2>(), ldx);
    // FP64 K(k) per problem for residuals and refinement
    std::vector<const double2*> kptr(B_);
    std::vector<DevBuf> kb(B_);
    for (int b = 0; b < B_; ++b) {
      HostProblem* hp = probs_[b];
      if (nc[b] == 0) { kptr[b] = mk_[b]; continue; }
      kb[b].alloc(sizeof(double2) * 9 * op_.nblk);
      hipLaunchKernelGGL(k_build_phase_blocks, dim3((S_.nf + 127) / 128), dim3(128), 0, stream(), S_.nf,
                         M_.d_rowptr.as<int>(), M_.d_colidx.as<int>(), M_.d_tslot.as<int>(),
                         M_.d_pinned.as<unsigned char>(), op_.Kc.as<double>(), op_.Kp.as<double>(),
                         op_.dKc.as<double>(), op_.dKp.as<double>(), hp->s, hp->pre, hp->pim,
                         kb[b].as<double2>());
      kptr[b] = kb[b].as<double2>();
    }
    DevBuf kp = upload(kptr, stream());
    residuals(X, KX, MX, ldx, nemax, nc, kp);
    // Refinement of Ritz vectors whose residual ||K x - lambda M x|| / ||K x|| is
    // above kRefineTol: one block inverse iteration Y = (K - sigma M)^-1 M X
    // followed by an FP64 Rayleigh-Ritz step on span(Y).  Inverse iteration damps
    // the roundoff-level high-mode content of x by lambda_i / lambda_high.
    DevBuf Y;
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<int> nr(B_, 0);
      bool any = false;
      for (int b = 0; b < B_; ++b) {
        HostProblem* hp = probs_[b];
        if (nc[b] == 0) continue;
        double worst = 0;
        for (int c = 0; c < nc[b]; ++c) worst = std::max(worst, hp->residual[c]);
        if (worst > kRefineTol || !(worst == worst)) {
          nr[b] = nc[b];
          any = true;
        }
      }
      if (!any) break;
      if (!Y.p) Y.alloc(sizeof(double2) * ldx * B_);
      refine(X, Y, KX, MX, ldx, nemax, nr, kp);
      residuals(X, KX, MX, ldx, nemax, nc, kp);
      for (int b = 0; b < B_; ++b)
        if (nr[b]) probs_[b]->refinements += 1;
    }
    // results: degenerate duplication + expansion (src/cyclic_solver.cpp:1153-1211)
    for (int b = 0; b < B_; ++b) {
      HostProblem* hp = probs_[b];
      tm_result* out = hp->out;
      if (!out) continue;
      if (hp->failed || !hp->done || hp->n_extract == 0) {
        out->status = TM_ERR_NUMERIC;
        out->n_modes = 0;
        continue;
      }
      const int nmodes_req = hp->in.num_modes;
      // min_frequency filter, then truncation to nev_target (src/cyclic_solver.cpp:1058-1096)
      std::vector<int> keep;
      for (int u = 0; u < hp->n_extract; ++u) {
        const double lam = hp->theta_sorted_lambda[u];
        const double f = lam > 0 ? std::sqrt(lam) / (2.0 * M_PI) : 0.0;
        if (hp->in.min_frequency > 0.0 && !(f > hp->in.min_frequency)) continue;
        keep.push_back(u);
      }
      if (keep.empty()) {
        out->status = TM_ERR_NUMERIC;
        out->n_modes = 0;
        continue;
      }
      if ((int)keep.size() > hp->nev_target) keep.resize(hp->nev_target);
      const int nu = static_cast<int>(keep.size());
      std::vector<double> freq, mass;
      std::vector<int> src, conjq, whirl;
      const double Nsec = static_cast<double>(S_.N);
      for (int q = 0; q < nu && q < nmodes_req; ++q) {
        const int u = keep[q];
        if (out->eigenvalues) out->eigenvalues[q] = hp->theta_sorted_lambda[u];
        if (out->residuals) out->residuals[q] = hp->residual[u];
      }
      auto freq_of = [&](int u) {
        const double lam = hp->theta_sorted_lambda[u];
        return lam > 0 ? std::sqrt(lam) / (2.0 * M_PI) : 0.0;
      };
      if (hp->degenerate) {
        const int nexp = std::min(2 * nu, nmodes_req);
        for (int q = 0; q < nexp; ++q) {
          const int u = keep[q / 2];
          freq.push_back(freq_of(u));
          mass.push_back(Nsec * hp->xmx[u]);
          src.push_back(u);
          conjq.push_back(q % 2);
          whirl.push_back(q % 2 == 0 ? 1 : -1);
        }
      } else {
        for (int q = 0; q < nu; ++q) {
          const int u = keep[q];
          freq.push_back(freq_of(u));
          mass.push_back(Nsec * hp->xmx[u]);
          src.push_back(u);
          conjq.push_back(0);
          whirl.push_back(0);
        }
      }
      const int nm = static_cast<int>(freq.size());
      out->n_modes = nm;
      out->n_unique = nu;
      out->converged = hp->converged ? 1 : 0;
      out->restarts = hp->restarts;
      out->status = TM_OK;
      if (out->frequencies) std::copy(freq.begin(), freq.end(), out->frequencies);
      if (out->whirl) std::copy(whirl.begin(), whirl.end(), out->whirl);
      if (out->modal_mass) std::copy(mass.begin(), mass.end(), out->modal_mass);
      if (out->mode_shapes && nm > 0) {
        DevBuf dsrc = upload(src, stream()), dconj = upload(conjq, stream());
        DevBuf dout(sizeof(double2) * 3 * (size_t)S_.nn * nm);
        const double alpha = 2.0 * M_PI / S_.N;
        const long long tot = (long long)S_.nn * nm;
        hipLaunchKernelGGL(k_expand_modes, dim3((tot + 255) / 256), dim3(256), 0, stream(), S_.nn,
                           M_.d_free_of_full.as<int>(), M_.d_node_class.as<int>(), std::cos(alpha),
                           std::sin(alpha), S_.c1, S_.c2, S_.cax, X.as<double2>() + b * ldx, n_,
                           dsrc.as<int>(), dconj.as<int>(), nm, hp->pre, hp->pim, dout.as<double2>());
        TM_HIP(hipMemcpyAsync(out->mode_shapes, dout.p, dout.bytes, hipMemcpyDeviceToHost, stream()));
        TM_HIP(hipStreamSynchronize(stream()));
      }
    }
    TM_HIP(hipStreamSynchronize(stream()));
  }

  void finish_failed() {
    for (auto* hp : probs_)
      if (hp->out) {
        hp->out->status = TM_ERR_NUMERIC;
        hp->out->n_modes = 0;
      }
  }
};

int64_t problem_footprint(const Operator& op, int num_modes, int harmonic) {
  const Symbolic& S = op.mesh->S;
  const Plan& P = op.mesh->plan;
  HostProblem hp;
  hp.in.harmonic = harmonic;
  hp.in.num_modes = num_modes;
  hp.in.rpm = 0;
  hp.in.min_frequency = 0.0;
  hp.in.max_restarts = 0;
  fill_policy(S, hp);
  const int64_t n = 3LL * S.nf;
  int64_t bytes = 16 * (P.panel_entries + P.upool_entries + P.tvec_entries + P.inv_scratch_entries);
  bytes += 2 * 16 * (int64_t)(hp.ncv + 1) * n;       // V, MV
  bytes += 2 * 16 * n;                               // W, MW
  bytes += 3 * 16 * (int64_t)(hp.nev + 1) * n;       // restart / Ritz / residual vectors
  bytes += 16 * 9 * (int64_t)op.nblk;                // K(k) for the residual check
  return bytes;
}

void solve_problems(Context* ctx, const tm_problem* problems, int n, tm_result* results) {
  std::vector<HostProblem> hps(n);
  std::map<Operator*, std::vector<HostProblem*>> groups;
  std::vector<Operator*> order;
  for (int i = 0; i < n; ++i) {
    hps[i].in = problems[i];
    hps[i].out = results ? &results[i] : nullptr;
    if (!problems[i].op) throw ArgError("problem without operator");
    if (problems[i].num_modes <= 0) throw ArgError("num_modes must be positive");
    Operator* op = static_cast<Operator*>(problems[i].op);
    const int max_k = op->mesh->S.N / 2;
    if (problems[i].harmonic < 0 || problems[i].harmonic > max_k)
      throw ArgError("harmonic index out of range 0..N/2");
    if (!groups.count(op)) order.push_back(op);
    groups[op].push_back(&hps[i]);
  }
  for (Operator* op : order) {
    auto& list = groups[op];
    int64_t per = 0;
    for (auto* hp : list) {
      fill_policy(op->mesh->S, *hp);
      per = std::max<int64_t>(per, problem_footprint(*op, hp->in.num_modes, hp->k));
    }
    size_t free_b = 0, total_b = 0;
    TM_HIP(hipMemGetInfo(&free_b, &total_b));
    int64_t budget = (int64_t)(0.85 * (double)free_b);
    size_t max_batch = (size_t)std::max<int64_t>(1, std::min<int64_t>(64, budget / std::max<int64_t>(per, 1)));
    for (size_t i = 0; i < list.size(); i += max_batch) {
      std::vector<HostProblem*> chunk(list.begin() + i, list.begin() + std::min(list.size(), i + max_batch));
      BatchLanczos bl(ctx, *op, chunk);
      bl.run();
    }
  }
}

}  // namespace tme
