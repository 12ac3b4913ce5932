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
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <set>
#include <numeric>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <random>
#include <exception>
#include <thread>
#include <type_traits>

#include "vec.hpp"

namespace tme {

// ---------------------------------------------------------------------------
// timing
// ---------------------------------------------------------------------------
Deliverer::Deliverer(int device) : device_(device) {
  TM_HIP(hipSetDevice(device));
  TM_HIP(hipStreamCreateWithFlags(&cs_, hipStreamNonBlocking));
  TM_HIP(hipStreamCreateWithFlags(&ws_, hipStreamNonBlocking));
  th_ = std::thread([this] { run(); });
}

Deliverer::~Deliverer() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  if (th_.joinable()) th_.join();
  for (auto& f : free_) (void)hipFree(f.second);
  for (auto e : events_) (void)hipEventDestroy(e);
  if (mark_) (void)hipEventDestroy(mark_);
  (void)hipStreamDestroy(cs_);
  (void)hipStreamDestroy(ws_);
}

void* Deliverer::stage(size_t n, size_t& cap) {
  n = DevPool::round(n);
  std::unique_lock<std::mutex> lk(mu_);
  cv_done_.wait(lk, [&] { return inflight_ == 0 || inflight_ + n <= kMaxInFlight || !err_.empty(); });
  if (!err_.empty()) throw HipError("result copy: " + err_);
  auto it = free_.lower_bound(n);
  void* p = nullptr;
  if (it != free_.end() && it->first <= 2 * n + (size_t(4) << 20)) {
    p = it->second;
    cap = it->first;
    free_.erase(it);
  } else {
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess && !free_.empty()) {  // give the idle blocks back and retry
      (void)hipGetLastError();
      for (auto& f : free_) (void)hipFree(f.second);
      free_.clear();
      e = hipMalloc(&p, n);
    }
    if (e != hipSuccess) throw HipError("hipMalloc(" + std::to_string(n) + " B): " + hipGetErrorString(e));
    cap = n;
  }
  inflight_ += cap;
  return p;
}

void Deliverer::unstage(void* p, size_t cap) {
  std::lock_guard<std::mutex> g(mu_);
  free_.emplace(cap, p);
  inflight_ -= cap;
  cv_done_.notify_all();
}

void Deliverer::enqueue(hipStream_t s, Job j) {
  {
    std::lock_guard<std::mutex> g(mu_);
    if (!events_.empty()) {
      j.ev = events_.back();
      events_.pop_back();
    }
  }
  if (!j.ev) TM_HIP(hipEventCreateWithFlags(&j.ev, hipEventDisableTiming));
  if (!mark_) TM_HIP(hipEventCreateWithFlags(&mark_, hipEventDisableTiming));
  // the compute stream's point, passed on to the wait stream while the compute
  // stream is not capturing (this thread is its only user); the worker then waits
  // for j.ev on the wait stream only
  TM_HIP(hipEventRecord(mark_, s));
  TM_HIP(hipStreamWaitEvent(ws_, mark_, 0));
  TM_HIP(hipEventRecord(j.ev, ws_));
  std::lock_guard<std::mutex> g(mu_);
  q_.push_back(std::move(j));
  cv_.notify_all();
}

void Deliverer::copy(hipStream_t s, void* src, size_t cap, void* dst, size_t n) {
  Job j;
  j.src = src;
  j.cap = cap;
  j.dst = dst;
  j.n = n;
  enqueue(s, std::move(j));
}

void Deliverer::then(hipStream_t s, std::function<void(const std::string&)> fn) {
  Job j;
  j.fn = std::move(fn);
  enqueue(s, std::move(j));
}

void Deliverer::run() {
  (void)hipSetDevice(device_);
  while (true) {
    Job j;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] { return stop_ || !q_.empty(); });
      if (q_.empty()) return;
      j = std::move(q_.front());
      q_.pop_front();
      ++active_;
    }
    std::string e;
    hipError_t r = hipSuccess;
    // j.ev is on the wait stream (never capturing): the compute stream may be
    // capturing a graph by now, and an event last recorded on a capturing stream can
    // be neither waited for nor queried
    r = hipEventSynchronize(j.ev);
    if (j.dst) {
      if (r == hipSuccess) r = hipMemcpyAsync(j.dst, j.src, j.n, hipMemcpyDeviceToHost, cs_);
      if (r == hipSuccess) r = hipStreamSynchronize(cs_);
    }
    if (r != hipSuccess) e = hipGetErrorString(r);
    {
      std::lock_guard<std::mutex> g(mu_);
      if (!e.empty() && err_.empty()) err_ = e;
      if (j.fn && e.empty()) e = err_;
    }
    if (j.fn) {
      try {
        j.fn(e);
      } catch (const std::exception& x) {
        std::lock_guard<std::mutex> g(mu_);
        if (err_.empty()) err_ = x.what();
      }
    }
    {
      std::lock_guard<std::mutex> g(mu_);
      if (j.src) {
        free_.emplace(j.cap, j.src);
        inflight_ -= j.cap;
      }
      events_.push_back(j.ev);
      --active_;
    }
    cv_done_.notify_all();
  }
}

void Deliverer::drain_quiet() {
  std::unique_lock<std::mutex> lk(mu_);
  cv_done_.wait(lk, [&] { return q_.empty() && active_ == 0; });
}

void Deliverer::drain() {
  drain_quiet();
  std::lock_guard<std::mutex> g(mu_);
  if (!err_.empty()) {
    const std::string e = err_;
    err_.clear();
    throw HipError("result copy: " + e);
  }
}

hipStream_t Context::side_stream() {
  if (!side) TM_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
  return side;
}

hipEvent_t Context::fork_event(size_t i) {
  while (fork_events.size() <= i) {
    hipEvent_t e = nullptr;
    TM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    fork_events.push_back(e);
  }
  return fork_events[i];
}

void Context::release_side() {
  if (side) (void)hipStreamSynchronize(side);
  for (auto e : fork_events) (void)hipEventDestroy(e);
  fork_events.clear();
  if (side) (void)hipStreamDestroy(side);
  side = nullptr;
}

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

static double wall_ms() {
  return 1e3 * std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

HostTimer::HostTimer(Context* c, const char* n) : ctx(c), name(n) {
  if (ctx && ctx->profiling) t0 = wall_ms();
}

HostTimer::~HostTimer() {
  if (ctx && ctx->profiling) {
    KernelStat& s = ctx->stats[std::string("host:") + name];
    s.launches += 1;
    s.total_ms += wall_ms() - t0;
  }
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

bool gen_eig_chol(int m, const std::vector<double>& H, const std::vector<double>& G, std::vector<double>& lam,
                  std::vector<double>& X) {
  using cd = std::complex<double>;
  auto at = [m](const std::vector<double>& A, int i, int j) {
    return cd(A[2 * ((size_t)i * m + j)], A[2 * ((size_t)i * m + j) + 1]);
  };
  std::vector<cd> L((size_t)m * m, cd(0, 0));
  double gmax = 0;
  for (int i = 0; i < m; ++i) gmax = std::max(gmax, std::fabs(G[2 * ((size_t)i * m + i)]));
  for (int j = 0; j < m; ++j) {
    double dj = 0.5 * (at(G, j, j).real() + at(G, j, j).real());
    for (int k = 0; k < j; ++k) dj -= std::norm(L[(size_t)j * m + k]);
    if (!(dj > 1e-13 * gmax)) return false;
    const double ljj = std::sqrt(dj);
    L[(size_t)j * m + j] = ljj;
    for (int i = j + 1; i < m; ++i) {
      cd s = 0.5 * (at(G, i, j) + std::conj(at(G, j, i)));
      for (int k = 0; k < j; ++k) s -= L[(size_t)i * m + k] * std::conj(L[(size_t)j * m + k]);
      L[(size_t)i * m + j] = s / ljj;
    }
  }
  // Y = L^-1 H (columns), then C = L^-1 Y^H (= L^-1 H L^-H, Hermitian)
  std::vector<cd> Y((size_t)m * m), C((size_t)m * m);
  for (int c = 0; c < m; ++c)
    for (int i = 0; i < m; ++i) {
      cd s = 0.5 * (at(H, i, c) + std::conj(at(H, c, i)));
      for (int k = 0; k < i; ++k) s -= L[(size_t)i * m + k] * Y[(size_t)k * m + c];
      Y[(size_t)i * m + c] = s / L[(size_t)i * m + i];
    }
  for (int c = 0; c < m; ++c)
    for (int i = 0; i < m; ++i) {
      cd s = std::conj(Y[(size_t)c * m + i]);
      for (int k = 0; k < i; ++k) s -= L[(size_t)i * m + k] * C[(size_t)k * m + c];
      C[(size_t)i * m + c] = s / L[(size_t)i * m + i];
    }
  std::vector<double> A(2 * (size_t)m * m), S;
  for (size_t q = 0; q < (size_t)m * m; ++q) {
    A[2 * q] = C[q].real();
    A[2 * q + 1] = C[q].imag();
  }
  herm_eig(m, A, lam, S);
  // X = L^-H S (back substitution with the upper factor L^H)
  X.assign(2 * (size_t)m * m, 0.0);
  std::vector<cd> x(m);
  for (int q = 0; q < m; ++q) {
    for (int i = m - 1; i >= 0; --i) {
      cd s(S[2 * ((size_t)i * m + q)], S[2 * ((size_t)i * m + q) + 1]);
      for (int k = i + 1; k < m; ++k) s -= std::conj(L[(size_t)k * m + i]) * x[k];
      x[i] = s / L[(size_t)i * m + i];
    }
    for (int i = 0; i < m; ++i) {
      X[2 * ((size_t)i * m + q)] = x[i].real();
      X[2 * ((size_t)i * m + q) + 1] = x[i].imag();
    }
  }
  return true;
}

// Runs f(i) for i in [0, n) on up to hardware_concurrency() threads (static split;
// the result does not depend on the split).
template <class F>
static void parallel_for(int n, F&& f) {
  const int nt = std::max(1, std::min<int>(n, (int)std::thread::hardware_concurrency()));
  if (nt <= 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> th;
  std::vector<std::exception_ptr> err(nt);
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      try {
        for (int i = t; i < n; i += nt) f(i);
      } catch (...) {
        err[t] = std::current_exception();
      }
    });
  for (auto& x : th) x.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
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
  DevBuf L, U, T, S, status, pivinfo, params, lead, share_ws;
  int pcols = 1;
  int nlead = 0;           // classes among the last factorization's problems (0: no class table)
  std::vector<int> lead_h;  // per slot: the slot holding its shared fronts (host copy of `lead`)
  FactorElem elem = FactorElem::C64;
  BatchBufs bufs(const Plan&, int B) {
    BatchBufs bb;
    bb.L = L.p;
    bb.U = U.p;
    bb.S = S.p;
    bb.elem = elem;
    bb.T = T.as<double2>();
    bb.status = status.as<int>();
    bb.pivinfo = pivinfo.as<int>();
    bb.params = params.as<ProbParam>();
    bb.B = B;
    if (nlead > 0) {
      bb.lead = lead.as<int>();
      bb.share_ws = share_ws.as<int>();
      bb.nlead = nlead;
    }
    return bb;
  }
  // Classes of the problems `fac` (slots) about to be factored: the same operator,
  // omega^2 and shift give the same shared fronts.  lead[b] = first slot of b's class
  // in `fac`; the other slots keep theirs.  Uploaded on `st`.
  void set_classes(const std::vector<int>& fac, const std::vector<ProbParam>& prm, const std::vector<OpArrays>& ops,
                   int B, hipStream_t st) {
    if ((int)lead_h.size() != B) {
      lead_h.resize(B);
      for (int b = 0; b < B; ++b) lead_h[b] = b;
    }
    std::vector<int>& ld = lead_h;
    for (int b : fac) ld[b] = b;
    nlead = 0;
    for (size_t i = 0; i < fac.size(); ++i) {
      const int b = fac[i];
      for (size_t j = 0; j < i; ++j) {
        const int c = fac[j];
        if (ld[c] == c && ops[c].Kc == ops[b].Kc && prm[c].s == prm[b].s && prm[c].sigma == prm[b].sigma) {
          ld[b] = c;
          break;
        }
      }
      if (ld[b] == b) ++nlead;
    }
    TM_HIP(hipMemcpyAsync(lead.p, ld.data(), sizeof(int) * B, hipMemcpyHostToDevice, st));
    TM_HIP(hipStreamSynchronize(st));
  }
  void alloc(DevPool* pool, const Plan& P, int B, int p = 1, FactorElem e = FactorElem::C64) {
    pcols = p;
    elem = e;
    const size_t es = factor_elem_bytes(e);
    L.alloc(pool, es * P.panel_entries * B);
    U.alloc(pool, es * P.upool_entries * B);
    T.alloc(pool, sizeof(double2) * P.tvec_entries * B * p);
    S.alloc(pool, es * P.inv_scratch_entries * B);
    status.alloc(pool, sizeof(int) * B);
    pivinfo.alloc(pool, sizeof(int) * 2 * B);
    params.alloc(pool, sizeof(ProbParam) * B);
    lead.alloc(pool, sizeof(int) * B);
    share_ws.alloc(pool, sizeof(int) * 4 * B);
    nlead = 0;
    lead_h.assign(B, 0);
    for (int b = 0; b < B; ++b) lead_h[b] = b;
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
    bm.alloc(nullptr, M.plan, 1);
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
    if (status & 1) throw NumericError("zero pivot in the static centrifugal solve");
    launch_expand_real(ctx, M, x.as<double2>(), ufull);
    TM_HIP(hipMemcpyAsync(op.u_static.data(), ufull.p, ufull.bytes, hipMemcpyDeviceToHost, st));
    launch_element_sigma(ctx, M, op.E, op.nu, ufull, s_ab);
  }
  launch_assemble(ctx, M, 2, false, ke, m_abs, m_raw, &s_ab, op.dKc);
  launch_assemble(ctx, M, 2, true, ke, m_abs, m_raw, &s_ab, op.dKp);
  op.real_slot.alloc(nblk);
  hipLaunchKernelGGL(k_real_slots, dim3((unsigned)((nblk + 255) / 256)), dim3(256), 0, st, (long long)nblk,
                     M.d_slot_row.as<int>(), M.d_colidx.as<int>(), M.d_tslot.as<int>(),
                     M.d_pinned.as<unsigned char>(), op.Kp.as<double>(), op.dKp.as<double>(), op.Mp.as<double>(),
                     op.real_slot.as<unsigned char>());
  op.real_pack.alloc(sizeof(double2) * 14 * nblk);
  hipLaunchKernelGGL(k_pack_real_slots, dim3((unsigned)((14 * nblk + 255) / 256)), dim3(256), 0, st,
                     (long long)nblk, op.Kc.as<double>(), op.dKc.as<double>(), op.Mc.as<double>(),
                     op.real_pack.as<double2>());
  TM_HIP(hipGetLastError());
  TM_HIP(hipStreamSynchronize(st));
}

// Kc[diag_slot[f]] += add[f] (9 doubles per free node)
__global__ void k_add_diag_blocks(int nf, const int* __restrict__ diag_slot, const double* __restrict__ add,
                                  double* __restrict__ Kc) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= 9LL * nf) return;
  const int f = (int)(i / 9), q = (int)(i % 9);
  Kc[(long long)diag_slot[f] * 9 + q] += add[i];
}

// Grounded springs (ELASTIC_SUPPORT) on the diagonal blocks of K_c: a node of the
// interior or the left boundary adds diag(k) to its own block; a right-boundary node
// acts on its left partner, u_right = R(alpha) u_left with R rotating the in-plane pair
// (c1, c2) (src/cyclic_solver.cpp:291-327), so it adds R^T diag(k) R, the same for every
// harmonic (|P| = 1).  Pinned components are left out (removed DOFs in the reference).
void add_springs(Context* ctx, Operator& op, const double* kfull) {
  Mesh& M = *op.mesh;
  const Symbolic& S = M.S;
  const double alpha = 2.0 * M_PI / S.N, ca = std::cos(alpha), sa = std::sin(alpha);
  std::vector<double> add(9 * (size_t)S.nf, 0.0);
  bool any = false;
  for (int i = 0; i < S.nn; ++i) {
    const double* kv = kfull + 3 * (size_t)i;
    for (int c = 0; c < 3; ++c)
      if (!(kv[c] >= 0.0) || !std::isfinite(kv[c])) throw ArgError("spring stiffness must be finite and >= 0");
    if (kv[0] == 0.0 && kv[1] == 0.0 && kv[2] == 0.0) continue;
    const int f = S.free_of_full[i];
    if (f < 0) continue;
    double blk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (S.node_class[i] == TM_NODE_RIGHT) {
      const int a = S.c1, b = S.c2, z = S.cax;
      const double k1 = kv[a], k2 = kv[b];
      blk[3 * a + a] = ca * ca * k1 + sa * sa * k2;
      blk[3 * a + b] = blk[3 * b + a] = sa * ca * (k2 - k1);
      blk[3 * b + b] = sa * sa * k1 + ca * ca * k2;
      blk[3 * z + z] = kv[z];
    } else {
      for (int c = 0; c < 3; ++c) blk[4 * c] = kv[c];
    }
    for (int ci = 0; ci < 3; ++ci)
      for (int cj = 0; cj < 3; ++cj) {
        if (S.pinned[3 * (size_t)f + ci] || S.pinned[3 * (size_t)f + cj]) continue;
        add[9 * (size_t)f + 3 * ci + cj] += blk[3 * ci + cj];
        any = any || blk[3 * ci + cj] != 0.0;
      }
  }
  if (!any) return;
  hipStream_t st = ctx->stream;
  DevBuf d = upload(add, st);
  hipLaunchKernelGGL(k_add_diag_blocks, dim3((unsigned)((9LL * S.nf + 255) / 256)), dim3(256), 0, st, S.nf,
                     M.d_diag_slot.as<int>(), d.as<double>(), op.Kc.as<double>());
  // the split K/M product reads K_c from the packed real slots
  hipLaunchKernelGGL(k_pack_real_slots, dim3((unsigned)((14 * op.nblk + 255) / 256)), dim3(256), 0, st,
                     (long long)op.nblk, op.Kc.as<double>(), op.dKc.as<double>(), op.Mc.as<double>(),
                     op.real_pack.as<double2>());
  TM_HIP(hipGetLastError());
  TM_HIP(hipStreamSynchronize(st));
}

__global__ void k_inertia_start(ProbParam* params, ProbParam pp, int* status) {
  if (threadIdx.x == 0) {
    params[0] = pp;
    status[0] = 0;
  }
}

// Factors K(k) - shift M(k) at rpm (one problem) into bb and reads back the number of
// negative pivots, the smallest relative pivot and the factor's status bits.  Launches
// and a device-to-host copy only: bb's memory is the caller's.
static void inertia_into(Context* ctx, Operator& op, double rpm, int k, double shift, const BatchBufs& bb,
                         int64_t* neg, double* ratio, int* status) {
  Mesh& M = *op.mesh;
  const Symbolic& S = M.S;
  hipStream_t st = ctx->stream;
  const double w = rpm * 2.0 * M_PI / 60.0;
  ProbParam pp{1.0, 0.0, w * w, shift};
  harmonic_phase(S.N, k, pp.pre, pp.pim);
  hipLaunchKernelGGL(k_inertia_start, dim3(1), dim3(64), 0, st, const_cast<ProbParam*>(bb.params), pp, bb.status);
  factor_batch(ctx, M, M.plan, op_arrays(op), bb);
  int info[3] = {0, 0, 0};
  TM_HIP(hipMemcpyAsync(info, bb.status, sizeof(int), hipMemcpyDeviceToHost, st));
  TM_HIP(hipMemcpyAsync(info + 1, bb.pivinfo, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
  TM_HIP(hipStreamSynchronize(st));
  float r;
  std::memcpy(&r, &info[2], sizeof(float));
  *neg = info[1];
  *ratio = r;
  *status = info[0];
}

// Negative pivots of the LDL^H of K(k) - shift M(k) at rpm, workspace from the
// context's pool.
void operator_inertia(Context* ctx, Operator& op, double rpm, int k, double shift, int64_t* neg, double* ratio,
                      int* status) {
  BatchMem bm;
  bm.alloc(&ctx->pool, op.mesh->plan, 1);
  inertia_into(ctx, op, rpm, k, shift, bm.bufs(op.mesh->plan, 1), neg, ratio, status);
}

// Bytes of one problem's factor workspace laid out by operator_inertia_in.
static size_t factor_workspace_bytes(const Plan& P) {
  auto al = [](size_t n) { return (n + 255) / 256 * 256; };
  return al(sizeof(double2) * P.panel_entries) + al(sizeof(double2) * P.upool_entries) +
         al(sizeof(double2) * P.inv_scratch_entries) + al(sizeof(int)) + al(2 * sizeof(int)) + al(sizeof(ProbParam));
}

// The same with the workspace carved out of caller-given device memory (no allocation).
void operator_inertia_in(Context* ctx, Operator& op, double rpm, int k, double shift, void* work, size_t work_bytes,
                         int64_t* neg, double* ratio, int* status) {
  const Plan& P = op.mesh->plan;
  const size_t need = factor_workspace_bytes(P);
  if (need > work_bytes)
    throw ArgError("the factor workspace of this mesh (" + std::to_string(need) + " B) exceeds the " +
                   std::to_string(work_bytes) + " B device scratch of the allocation-free entry point");
  auto al = [](size_t n) { return (n + 255) / 256 * 256; };
  char* p = static_cast<char*>(work);
  BatchBufs bb;
  bb.L = p;
  p += al(sizeof(double2) * P.panel_entries);
  bb.U = p;
  p += al(sizeof(double2) * P.upool_entries);
  bb.T = nullptr;  // the factorization uses no solve vectors
  bb.S = p;
  p += al(sizeof(double2) * P.inv_scratch_entries);
  bb.status = reinterpret_cast<int*>(p);
  p += al(sizeof(int));
  bb.pivinfo = reinterpret_cast<int*>(p);
  p += al(2 * sizeof(int));
  bb.params = reinterpret_cast<const ProbParam*>(p);
  bb.B = 1;
  inertia_into(ctx, op, rpm, k, shift, bb, neg, ratio, status);
}

// M(k) blocks of the operator, cached per harmonic.  Operators are shared by
// the slot groups (each on its own thread and stream): the first user builds the
// blocks on the operator's own stream and waits for them (a few tens of
// microseconds: nothing else runs there) before it enters them, so later users
// (on any stream) find them complete.  No event is shared between the streams: a
// stream cannot wait on, nor the host query, an event last recorded on a stream
// that is capturing a graph, which another group's may be doing.
const double2* harmonic_mass(Context* ctx, Operator& op, int k) {
  static std::mutex mu;
  std::lock_guard<std::mutex> g(mu);
  auto it = op.Mk_cache.find(k);
  if (it != op.Mk_cache.end()) return it->second.as<double2>();
  if (!op.mk_stream) {
    TM_HIP(hipSetDevice(ctx->device));
    TM_HIP(hipStreamCreateWithFlags(&op.mk_stream, hipStreamNonBlocking));
  }
  Mesh& M = *op.mesh;
  const Symbolic& S = M.S;
  DevBuf b(sizeof(double2) * 9 * op.nblk);
  double pre, pim;
  harmonic_phase(S.N, k, pre, pim);
  hipLaunchKernelGGL(k_build_phase_blocks, dim3((9 * (long long)op.nblk + 255) / 256), dim3(256), 0, op.mk_stream,
                     (long long)op.nblk, M.d_slot_row.as<int>(), M.d_colidx.as<int>(), M.d_tslot.as<int>(),
                     M.d_pinned.as<unsigned char>(), op.Mc.as<double>(), op.Mp.as<double>(),
                     (const double*)nullptr, (const double*)nullptr, 0.0, pre, pim, b.as<double2>(), 0,
                     make_double2(1.0, 0.0));
  TM_HIP(hipGetLastError());
  TM_HIP(hipStreamSynchronize(op.mk_stream));
  const double2* p = b.as<double2>();
  op.Mk_cache.emplace(k, std::move(b));
  return p;
}

// ---------------------------------------------------------------------------
// batched Lanczos
// ---------------------------------------------------------------------------
// Vectors per block of the block method: 8 for problems of at most 6 wanted
// pairs, else 16 (16 for all took as many sweeps on the C2 sweep, with more
// restarts, and each sweep costs more)
constexpr int kNarrowNev = 6;
inline int block_width(int nev, bool qep) { return (nev <= kNarrowNev && !qep) ? 8 : 16; }
// Basis length of the block method: 8 blocks of 8; 16-wide blocks 8 of them (6
// for the Coriolis pencil, whose vectors are twice as long).  With these the C4
// sweep restarts almost never and the Coriolis problems about twice on average.
// Problems wanting more pairs get a longer basis: its kept half holds the nev pairs and
// one more block.  Up to nev = 24 (48 for the Coriolis pencil) every problem of a block
// size has the same basis.
inline int basis_length(int pb, bool qep, int nev) {
  const int base = pb == 8 ? 64 : (qep ? 96 : 128);
  const int need = (2 * (nev + pb) + pb - 1) / pb * pb;
  return std::max(base, need);
}
// the most pairs the block method takes on (basis of at most 256 vectors)
constexpr int kBlockMaxNev = 112;

struct HostProblem {
  tm_problem in;
  tm_result* out = nullptr;
  int k = 0;
  double omega = 0, s = 0, pre = 1, pim = 0;
  bool degenerate = false;
  bool qep = false;  // Coriolis: the Lancaster pencil instead of (K, M)
  int nev_target = 0, nev = 0, ncv = 0;
  int restarts = 0, max_restarts = 1000;
  bool done = false, converged = false, failed = false;
  bool warm_started = false;  // its start block holds an adjacent rpm's Ritz vectors
  int n_extract = 0;
  std::vector<double> theta_sorted_lambda;  // extracted eigenvalues
  std::vector<double> S_extract;            // ncv x n_extract (row-major k, c)
  std::vector<double> residual, xmx;        // per extracted pair
  int refinements = 0;
  int ncv_final = 0;
  int steps = 0;                             // block steps taken (seeds its restarts)
  std::chrono::steady_clock::time_point t_admit;
  std::string message;
  // status bits of its shift-invert factorization: bit 0 a zero or non-finite pivot,
  // bit 1 a near-singular one (the shift on an eigenvalue); either fails the problem,
  // and its retry moves the shift (retry_failed)
  int factor_flags = 0;
};

// the real factor for real harmonics (TMHIP_REAL_FACTOR=0 in the environment: the complex one)
static bool real_factor_enabled() {
  const char* e = std::getenv("TMHIP_REAL_FACTOR");
  return !(e && std::strcmp(e, "0") == 0);
}

static const char* factor_message(int flags) {
  return (flags & 1) ? "zero or non-finite pivot in the shift-invert factorization"
                     : "near-singular pivot in the shift-invert factorization (the shift is on an eigenvalue)";
}

static void fill_policy(const Symbolic& S, HostProblem& hp) {
  const int N = S.N, max_k = N / 2;
  hp.k = hp.in.harmonic;
  const double alpha = 2.0 * M_PI / N;
  harmonic_phase(N, hp.k, hp.pre, hp.pim);
  hp.omega = hp.in.rpm * 2.0 * M_PI / 60.0;
  hp.s = hp.omega * hp.omega;
  hp.degenerate = hp.k > 0 && !(N % 2 == 0 && hp.k == max_k);
  const int M = hp.in.num_modes;
  const int n_free = S.n_free_dofs;
  const int extra = hp.in.min_frequency > 0.0 ? 6 : 0;
  // Coriolis with rotation at k > 0: the QEP, all num_modes wanted, no duplication
  // (src/cyclic_solver.cpp:975-1005); two extra pairs keep the |w| ordering safe
  hp.qep = hp.in.coriolis != 0 && hp.omega != 0.0 && hp.k > 0;
  if (hp.qep) {
    hp.degenerate = false;
    hp.nev_target = M;
    hp.nev = std::min(M + 2 + extra, 2 * n_free - 1);
    hp.ncv = 0;
    hp.max_restarts = hp.in.max_restarts > 0 ? hp.in.max_restarts : 1000;
    return;
  }
  hp.nev_target = hp.degenerate ? (M + 1) / 2 : M;
  hp.nev = std::min(hp.nev_target + extra, n_free - 1);
  if (n_free <= 200) hp.ncv = n_free;  // full Krylov space == the reference's dense solve
  else hp.ncv = std::min(std::max(4 * hp.nev + 1, 40), n_free);
  hp.max_restarts = hp.in.max_restarts > 0 ? hp.in.max_restarts : 1000;
}

class BatchLanczos {
 public:
  BatchLanczos(Context* ctx, Operator& op, const std::vector<HostProblem*>& probs, bool block)
      : ctx_(ctx), op_(op), M_(*op.mesh), S_(M_.S), P_(M_.plan), probs_(probs), block_(block) {}
  BatchLanczos(const BatchLanczos&) = delete;
  BatchLanczos& operator=(const BatchLanczos&) = delete;
  ~BatchLanczos() {
    if (pend_side_) (void)hipStreamSynchronize(pend_side_->stream);  // it writes this batch's workspaces
    for (hipEvent_t e : side_ev_)
      if (e) (void)hipEventDestroy(e);
  }

  void run() {
    auto t0 = std::chrono::steady_clock::now();
    {
      HostTimer h(ctx_, "setup");
      setup();
    }
    {
      HostTimer h(ctx_, "factor");
      factor();
    }
    {
      HostTimer h(ctx_, "iterate");
      start_vectors();
      iterate();
    }
    {
      HostTimer h(ctx_, "extract");
      extract(std::vector<int>(B_, 1));
    }
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
  std::vector<HostProblem*> probs_;  // per slot (nullptr: free)
  std::vector<int> hold_;            // per slot: problems in flight whose shared fronts it holds
  int B_ = 0, ncvmax_ = 0, kmax_ = 0, nevmax_ = 0, nchunk_ = 0;
  long long n_ = 0, ldb_ = 0;
  BatchMem bm_;
  DevBuf V_, MV_, W_, MW_, h_, part_, alpha_, beta_, bnew_, jcur_, ncv_, live_, frozen_, act_;
  DevBuf mk_ptrs_, tmp_, coef_, nk_, nc_;
  LzState st_{};
  std::vector<const double2*> mk_;
  std::vector<int> h_jcur_, h_ncv_, h_live_;

  hipStream_t stream() const { return ctx_->stream; }
  DevPool* pool() const { return &ctx_->pool; }

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
    nvec_ = n_;
    cs_ = n_;
    ldb_ = (long long)(ncvmax_ + 1) * n_;
    nchunk_ = static_cast<int>((n_ + kDotChunk - 1) / kDotChunk);
    // right-hand sides per triangular sweep: the Krylov block, and the Ritz
    // vectors of the refinement step (at most 16 per sweep)
    const int pcols = std::min(16, std::max(block_ ? block_width(nevmax_, false) : 1, nevmax_));
    bm_.alloc(pool(), P_, B_, pcols);
    act_.alloc(pool(), sizeof(int) * B_);
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
    mk_ptrs_ = upload(pool(), mk_, stream());
    if (block_) {
      block_setup();
      return;
    }
    V_.alloc(pool(), sizeof(double2) * ldb_ * B_);
    MV_.alloc(pool(), sizeof(double2) * ldb_ * B_);
    W_.alloc(pool(), sizeof(double2) * n_ * B_);
    MW_.alloc(pool(), sizeof(double2) * n_ * B_);
    h_.alloc(pool(), sizeof(double2) * kmax_ * B_);
    part_.alloc(pool(), sizeof(double2) * (size_t)kmax_ * nchunk_ * B_);
    alpha_.alloc(pool(), sizeof(double) * ncvmax_ * B_);
    beta_.alloc(pool(), sizeof(double) * ncvmax_ * B_);
    bnew_.alloc(pool(), sizeof(double) * B_);
    for (DevBuf* b : {&jcur_, &ncv_, &live_, &frozen_}) b->alloc(pool(), sizeof(int) * B_);
    nk_.alloc(pool(), sizeof(int) * B_);
    nc_.alloc(pool(), sizeof(int) * B_);
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
    {  // the batch's problems stay to the end of the run: its classes share fronts
      std::vector<int> all(B_);
      std::vector<ProbParam> pp(B_);
      for (int b = 0; b < B_; ++b) {
        all[b] = b;
        const HostProblem* hp = probs_[b];
        pp[b] = {hp->pre, hp->pim, hp->s, hp->in.shift};
      }
      bm_.set_classes(all, pp, std::vector<OpArrays>(B_, op_arrays(op_)), B_, stream());
    }
    BatchBufs bb = bm_.bufs(P_, B_);
    factor_batch(ctx_, M_, P_, op_arrays(op_), bb);
    std::vector<int> status(B_);
    TM_HIP(hipMemcpyAsync(status.data(), bm_.status.p, sizeof(int) * B_, hipMemcpyDeviceToHost, stream()));
    TM_HIP(hipStreamSynchronize(stream()));
    for (int b = 0; b < B_; ++b) {
      probs_[b]->factor_flags = status[b];
      if (status[b]) {
        probs_[b]->failed = true;
        probs_[b]->message = factor_message(status[b]);
      }
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
      tmp_.alloc(pool(), sizeof(double2) * (size_t)(nevmax_ + 1) * n_ * B_);
    coef_ = upload(pool(), coef, stream());
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
    DevBuf ncol = upload(pool(), ncolv, stream());
    spmv_cols(mk_ptrs_, V_.as<double2>(), ldb_, ncol, nevmax_, MV_.as<double2>());
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
  void residuals(double2* X, double2* KX, double2* MX, long long ldx, int nemax, const std::vector<int>& nc,
                 const DevBuf& kp) {
    (void)kp;
    DevBuf dnc = upload(pool(), nc, stream());
    apply_AB(X, ldx, dnc, nemax, KX, MX);
    std::vector<double> lam((size_t)B_ * nemax, 0.0);
    for (int b = 0; b < B_; ++b)
      for (int c = 0; c < nc[b]; ++c) {
        const double t = probs_[b]->theta_sorted_lambda[c];
        lam[(size_t)b * nemax + c] = qep_ ? (t != 0.0 ? 1.0 / t : 0.0) : t;
      }
    DevBuf dl = upload(pool(), lam, stream());
    DevBuf part(pool(), sizeof(double) * 3 * (size_t)B_ * nemax * nchunk_);
    if (qep_)  // X = [omega x; x]: ||K x + omega D x - omega^2 M x|| / ||K x|| from the b halves
      hipLaunchKernelGGL(k_residual_partial, dim3(nchunk_, B_, nemax), dim3(256), 0, stream(), X + n_, MX + n_,
                         KX + n_, ldx, n_, nvec_, dl.as<double>(), nemax, dnc.as<int>(), part.as<double>(), nchunk_);
    else
      hipLaunchKernelGGL(k_residual_partial, dim3(nchunk_, B_, nemax), dim3(256), 0, stream(), X, KX, MX, ldx,
                         nvec_, nvec_, dl.as<double>(), nemax, dnc.as<int>(), part.as<double>(), nchunk_);
    std::vector<double> hp_part(3 * (size_t)B_ * nemax * nchunk_);
    TM_HIP(hipMemcpyAsync(hp_part.data(), part.p, sizeof(double) * hp_part.size(), hipMemcpyDeviceToHost, stream()));
    TM_HIP(hipStreamSynchronize(stream()));
    for (int b = 0; b < B_; ++b) {
      HostProblem* hp = probs_[b];
      if (!nc[b]) continue;
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
    if (qep_) {  // Re(b^H M b) (the head of A z is M b); the modal mass is omega^2 times it
      DevBuf hpart(pool(), sizeof(double) * (size_t)B_ * nemax * nchunk_);
      hipLaunchKernelGGL(k_head_dot, dim3(nchunk_, B_, nemax), dim3(256), 0, stream(), X + n_, KX, ldx, nvec_, n_,
                         dnc.as<int>(), nemax, hpart.as<double>(), nchunk_);
      std::vector<double> hh((size_t)B_ * nemax * nchunk_);
      TM_HIP(hipMemcpyAsync(hh.data(), hpart.p, sizeof(double) * hh.size(), hipMemcpyDeviceToHost, stream()));
      TM_HIP(hipStreamSynchronize(stream()));
      for (int b = 0; b < B_; ++b)
        for (int c = 0; c < nc[b]; ++c) {
          double s = 0;
          for (int q = 0; q < nchunk_; ++q) s += hh[((size_t)b * nemax + c) * nchunk_ + q];
          probs_[b]->xmx[c] = s;
        }
    }
  }

  // FP64 iterative refinement of the shift-invert solves done with the complex-FP32
  // factor: W_b[:, c] (the W block's layout) approximates A_b^-1 Y_b[:, s0_b + c],
  // c < nc_b, A = K(k) - sigma M(k).  Each step forms R = Y - (K W - sigma M W) in FP64
  // (the split K/M product) and, while the largest ||R_c|| / ||Y_c|| of a problem still
  // falls by 10 % or more per step and is above kMixedTol, adds A~^-1 R (the FP32 factor
  // A~) to W; at most kMixedMaxRefine steps.  The residual's own FP64 rounding
  // (eps ||A|| ||W||) is where it stalls, as an FP64 factor's solve would.  The norms are
  // reduced in a fixed order, so the step count depends on the problem only.  KW and MW
  // serve as scratch.
  static constexpr int kMixedMaxRefine = 40;
  static constexpr double kMixedTol = 1e-14;
  void mixed_refine(double2* W, const double2* Y, long long ldy, const DevBuf* s0, const std::vector<int>& nc) {
    int ncol = 0;
    std::vector<int> act(B_, 0);
    for (int b = 0; b < B_; ++b)
      if (nc[b] > 0) {
        act[b] = 1;
        ncol = std::max(ncol, nc[b]);
      }
    std::vector<double> prev(B_, std::numeric_limits<double>::infinity());
    BatchBufs bb = bm_.bufs(P_, B_);
    const dim3 g((unsigned)((n_ + 255) / 256), B_);
    DevBuf part(pool(), sizeof(double) * 2 * (size_t)B_ * ncol * nchunk_);
    std::vector<double> hp(2 * (size_t)B_ * ncol * nchunk_);
    for (int it = 0; it <= kMixedMaxRefine; ++it) {
      int nact = 0;
      for (int b = 0; b < B_; ++b) nact += act[b];
      if (!nact) break;
      std::vector<int> ncv(B_, 0);
      for (int b = 0; b < B_; ++b) ncv[b] = act[b] ? nc[b] : 0;
      DevBuf dnc = ints(ncv), dact = ints(act);
      spmv_km(W, ldw_, dnc, ncol, KWb_.as<double2>(), MWb_.as<double2>());
      {
        KTimer t(ctx_, "vec", 0, 0);
        hipLaunchKernelGGL(k_mixed_residual, g, dim3(256), 0, stream(), Y, ldy, s0 ? s0->as<int>() : (const int*)nullptr,
                           KWb_.as<double2>(), MWb_.as<double2>(), ldw_, bm_.params.as<ProbParam>(), ncol, n_,
                           dact.as<int>());
        hipLaunchKernelGGL(k_colnorm_partial, dim3(nchunk_, B_), dim3(256), 0, stream(), KWb_.as<double2>(), ldw_, Y,
                           ldy, s0 ? s0->as<int>() : (const int*)nullptr, ncol, n_, dact.as<int>(),
                           part.as<double>(), nchunk_);
      }
      TM_HIP(hipMemcpyAsync(hp.data(), part.p, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, stream()));
      TM_HIP(hipStreamSynchronize(stream()));
      for (int b = 0; b < B_; ++b) {
        if (!act[b]) continue;
        double worst = 0.0;
        for (int c = 0; c < nc[b]; ++c) {
          double rr = 0.0, yy = 0.0;
          for (int q = 0; q < nchunk_; ++q) {
            rr += hp[(((size_t)b * ncol + c) * nchunk_ + q) * 2];
            yy += hp[(((size_t)b * ncol + c) * nchunk_ + q) * 2 + 1];
          }
          worst = std::max(worst, yy > 0.0 ? std::sqrt(rr / yy) : 0.0);
        }
        if (!(worst > kMixedTol) || !(worst < 0.9 * prev[b]) || it == kMixedMaxRefine) act[b] = 0;
        prev[b] = worst;
      }
      nact = 0;
      for (int b = 0; b < B_; ++b) nact += act[b];
      if (!nact) break;
      DevBuf dact2 = ints(act);
      trisolve_batch(ctx_, M_, P_, bb, KWb_.as<double2>(), ldw_, dact2.as<int>(), ncol, -1, nact, false, classes(act));
      KTimer t(ctx_, "vec", 0, 0);
      hipLaunchKernelGGL(k_add_cols, g, dim3(256), 0, stream(), W, KWb_.as<double2>(), ldw_, ncol, n_,
                         dact2.as<int>());
    }
  }

  // block inverse iteration + Rayleigh-Ritz for the problems with nr[b] > 0
  void refine(double2* X, double2* Y, double2* KY, double2* MY, long long ldx, int nemax, const std::vector<int>& nr,
              const DevBuf& kp) {
    DevBuf dnr = upload(pool(), nr, stream());
    const int nact = (int)std::count_if(nr.begin(), nr.end(), [](int v) { return v > 0; });
    // Y = M X, then solve in place column by column
    spmv_cols(mk_ptrs_, X, ldx, dnr, nemax, Y);
    BatchBufs bb = bm_.bufs(P_, B_);
    hipLaunchKernelGGL(k_flag_cols, dim3((B_ + 63) / 64), dim3(64), 0, stream(), dnr.as<int>(), 0, B_,
                       act_.as<int>());
    for (int q0 = 0; q0 < nemax; q0 += bm_.pcols)
      trisolve_batch(ctx_, M_, P_, bb, Y + (long long)q0 * n_, ldx, act_.as<int>(), std::min(bm_.pcols, nemax - q0),
                     -1, nact, false, classes(nr));
    if (mixed_)  // refine the columns in W-block chunks: copy in, refine against M X, copy back
      for (int q0 = 0; q0 < nemax; q0 += pb_) {
        const int kc = std::min(pb_, nemax - q0);
        std::vector<int> nc(B_, 0);
        for (int b = 0; b < B_; ++b) nc[b] = std::max(0, std::min(kc, nr[b] - q0));
        DevBuf dnc = ints(nc);
        {
          KTimer t(ctx_, "vec", 0, 0);
          hipLaunchKernelGGL(k_copy_cols, dim3((n_ + 255) / 256, B_), dim3(256), 0, stream(), Y + (long long)q0 * n_,
                             ldx, (const int*)nullptr, Wb_.as<double2>(), ldw_, (const int*)nullptr, kc, n_,
                             dnc.as<int>());
        }
        mixed_refine(Wb_.as<double2>(), MY + (long long)q0 * n_, ldx, nullptr, nc);
        KTimer t(ctx_, "vec", 0, 0);
        hipLaunchKernelGGL(k_copy_cols, dim3((n_ + 255) / 256, B_), dim3(256), 0, stream(), Wb_.as<double2>(), ldw_,
                           (const int*)nullptr, Y + (long long)q0 * n_, ldx, (const int*)nullptr, kc, n_,
                           dnc.as<int>());
      }
    if (!qep_ && &kp == &kptr_) spmv_km(Y, ldx, dnr, nemax, KY, MY);  // K Y, M Y in one pass
    else spmv_pair(kp, &mk_ptrs_, Y, ldx, dnr, nemax, KY, MY, false, false);
    const size_t gsz = (size_t)B_ * nemax * nemax * nchunk_;
    DevBuf pk(pool(), sizeof(double2) * gsz), pm(pool(), sizeof(double2) * gsz);
    hipLaunchKernelGGL(k_gram_partial, dim3(nchunk_, B_, nemax * nemax), dim3(256), 0, stream(), Y,
                       KY, ldx, n_, dnr.as<int>(), nemax, pk.as<double2>(), nchunk_);
    hipLaunchKernelGGL(k_gram_partial, dim3(nchunk_, B_, nemax * nemax), dim3(256), 0, stream(), Y,
                       MY, ldx, n_, dnr.as<int>(), nemax, pm.as<double2>(), nchunk_);
    std::vector<double2> hk(gsz), hm(gsz);
    TM_HIP(hipMemcpyAsync(hk.data(), pk.p, sizeof(double2) * hk.size(), hipMemcpyDeviceToHost, stream()));
    TM_HIP(hipMemcpyAsync(hm.data(), pm.p, sizeof(double2) * hm.size(), hipMemcpyDeviceToHost, stream()));
    TM_HIP(hipStreamSynchronize(stream()));
    std::vector<double2> zc((size_t)B_ * nemax * nemax, make_double2(0.0, 0.0));
    std::vector<int> keep = nr;  // problems whose X becomes Y Z
    bool skipped = false;
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
      try {
        herm_gen_eig(ne, A, Bm, mu, Z);
      } catch (const NumericError&) {
        // Y = (K - sigma M)^-1 M X lost rank: a shift on one of the wanted eigenvalues
        // turns every column towards that eigenvector.  X stays as it is.
        keep[b] = 0;
        skipped = true;
        continue;
      }
      HostProblem* hp = probs_[b];
      for (int q = 0; q < ne; ++q) hp->theta_sorted_lambda[q] = mu[q];
      for (int k = 0; k < ne; ++k)
        for (int q = 0; q < ne; ++q)
          zc[((size_t)b * nemax + k) * nemax + q] = make_double2(Z[2 * (k * ne + q)], Z[2 * (k * ne + q) + 1]);
    }
    DevBuf dz = upload(pool(), zc, stream());
    DevBuf dkeep = skipped ? upload(pool(), keep, stream()) : DevBuf();
    hipLaunchKernelGGL(k_combine_c, dim3((n_ + 255) / 256, B_), dim3(256), 0, stream(), Y, ldx, n_,
                       dz.as<double2>(), nemax, (long long)nemax * nemax, (skipped ? dkeep : dnr).as<int>(), X, ldx);
  }

  // Ritz vectors, residuals, refinement and results of the selected problems
  void extract(const std::vector<int>& sel) {
    int nemax = 0;
    std::vector<int> nc(B_, 0);
    for (int b = 0; b < B_; ++b) {
      HostProblem* hp = probs_[b];
      if (!sel[b] || !hp) continue;
      if (hp->done && !hp->failed) {
        nc[b] = hp->n_extract;
        nemax = std::max(nemax, hp->n_extract);
      }
    }
    if (nemax == 0) {
      finish_failed(sel);
      return;
    }
    DevBuf Xo, KXo, MXo, Yo;
    double2 *X, *KX, *MX;
    long long ldx;
    if (block_) {
      // Ritz vectors of the last Rayleigh-Ritz step of each problem
      X = Xb_.as<double2>();
      KX = KXb_.as<double2>();
      MX = MXb_.as<double2>();
      ldx = ldk_;
    } else {
      // Ritz vectors X = V S for every finished problem
      const int ldc = nemax;
      std::vector<double> coef((size_t)B_ * (ncvmax_ + 1) * ldc, 0.0);
      std::vector<int> nk(B_, 0);
      for (int b = 0; b < B_; ++b) {
        HostProblem* hp = probs_[b];
        if (!nc[b]) continue;
        const int ne = hp->n_extract, ncv = hp->ncv_final;
        for (int k = 0; k < ncv; ++k)
          for (int c = 0; c < ne; ++c)
            coef[(size_t)b * (ncvmax_ + 1) * ldc + (size_t)k * ldc + c] = hp->S_extract[(size_t)k * ne + c];
        nk[b] = ncv;
      }
      ldx = (long long)nemax * n_;
      Xo.alloc(pool(), sizeof(double2) * ldx * B_);
      KXo.alloc(pool(), sizeof(double2) * ldx * B_);
      MXo.alloc(pool(), sizeof(double2) * ldx * B_);
      X = Xo.as<double2>();
      KX = KXo.as<double2>();
      MX = MXo.as<double2>();
      coef_ = upload(pool(), coef, stream());
      TM_HIP(hipMemcpyAsync(nk_.p, nk.data(), sizeof(int) * B_, hipMemcpyHostToDevice, stream()));
      TM_HIP(hipMemcpyAsync(nc_.p, nc.data(), sizeof(int) * B_, hipMemcpyHostToDevice, stream()));
      hipLaunchKernelGGL(k_combine, dim3((n_ + 255) / 256, B_), dim3(256), 0, stream(), V_.as<double2>(), ldb_,
                         n_, coef_.as<double>(), ldc, (long long)(ncvmax_ + 1) * ldc, nk_.as<int>(),
                         nc_.as<int>(), X, ldx);
    }
    // FP64 K(k) per problem for residuals and refinement (the block method built
    // them on admission)
    if (!block_) {
      kblk_.resize(B_);
      kptr_host_.assign(B_, nullptr);
      std::vector<int> all;
      for (int b = 0; b < B_; ++b)
        if (nc[b]) all.push_back(b);
      build_stiffness_blocks(all);
    }
    const DevBuf& kp = kptr_;
    residuals(X, KX, MX, ldx, nemax, nc, kp);
    // Refinement of Ritz vectors whose residual ||K x - lambda M x|| / ||K x|| is
    // above kRefineTol: one block inverse iteration Y = (K - sigma M)^-1 M X
    // followed by an FP64 Rayleigh-Ritz step on span(Y).  Inverse iteration damps
    // the roundoff-level high-mode content of x by lambda_i / lambda_high.
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<int> nr(B_, 0);
      bool any = false;
      for (int b = 0; b < B_; ++b) {
        HostProblem* hp = probs_[b];
        if (nc[b] == 0) continue;
        double worst = 0;
        for (int c = 0; c < nc[b]; ++c) worst = std::max(worst, hp->residual[c]);
        // block results are already Rayleigh-Ritz optimal on a large basis: refine
        // only those left above the bar.  (A pair accepted within 4 x its FP64
        // floor estimate still gains: the estimate bounds the floor from above;
        // on the C2 sweep refinement takes the largest residual from 1.2e-7 to 2e-8.)
        if (!qep_ && (worst > (block_ ? kBlockStall : kRefineTol) || !(worst == worst))) {
          nr[b] = nc[b];
          any = true;
        }
      }
      if (!any) break;
      if (!Yo.p) Yo.alloc(pool(), sizeof(double2) * ldx * B_);
      refine(X, Yo.as<double2>(), KX, MX, ldx, nemax, nr, kp);
      residuals(X, KX, MX, ldx, nemax, nc, kp);
      for (int b = 0; b < B_; ++b)
        if (nr[b]) probs_[b]->refinements += 1;
    }
    // results: degenerate duplication + expansion (src/cyclic_solver.cpp:1153-1211)
    HostTimer hd(ctx_, "deliver");
    for (int b = 0; b < B_; ++b) {
      HostProblem* hp = probs_[b];
      if (!sel[b] || !hp) continue;
      tm_result* out = hp->out;
      if (!out) continue;
      if (hp->failed || !hp->done || hp->n_extract == 0) {
        out->status = TM_ERR_NUMERIC;
        out->n_modes = 0;
        continue;
      }
      const int nmodes_req = hp->in.num_modes;
      if (qep_) {
        deliver_qep(b, hp, X, ldx);
        continue;
      }
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
      if (out->mode_shapes && nm > 0) expand_and_send(b, hp, X, ldx, src, conjq, {});
    }
    TM_HIP(hipStreamSynchronize(stream()));
  }

  // Mode shapes of problem b (columns src of X, conjugated where conjq, times
  // scale[q] when given) expanded to all nodes on the device and sent to the
  // caller's buffer by the context's deliverer (the copy overlaps later work).
  void expand_and_send(int b, HostProblem* hp, const double2* X, long long ldx, const std::vector<int>& src,
                       const std::vector<int>& conjq, const std::vector<double>& scale) {
    const int nm = static_cast<int>(src.size());
    const size_t bytes = sizeof(double2) * 3 * (size_t)S_.nn * nm;
    DevBuf dsrc = upload(pool(), src, stream()), dconj = upload(pool(), conjq, stream());
    DevBuf dscale;
    if (!scale.empty()) dscale = upload(pool(), scale, stream());
    Deliverer& dl = ctx_->deliverer();
    size_t cap = 0;
    void* stage = dl.stage(bytes, cap);
    try {
      const double alpha = 2.0 * M_PI / S_.N;
      const long long tot = (long long)S_.nn * nm;
      hipLaunchKernelGGL(k_expand_modes, dim3((tot + 255) / 256), dim3(256), 0, stream(), S_.nn,
                         M_.d_free_of_full.as<int>(), M_.d_node_class.as<int>(), std::cos(alpha), std::sin(alpha),
                         S_.c1, S_.c2, S_.cax, X + b * ldx, nvec_, dsrc.as<int>(), dconj.as<int>(), nm, hp->pre,
                         hp->pim, dscale.p ? dscale.as<double>() : (const double*)nullptr,
                         static_cast<double2*>(stage));
      TM_HIP(hipGetLastError());
      dl.copy(stream(), stage, cap, hp->out->mode_shapes, bytes);
    } catch (...) {
      TM_HIP(hipStreamSynchronize(stream()));
      dl.unstage(stage, cap);
      throw;
    }
  }

  // Results of a Coriolis problem (src/modal_solver.cpp:258-307, src/cyclic_solver.cpp:1058-1096):
  // drop |omega| < 1e-6, keep up to num_modes forward (omega > 0) and backward
  // (omega < 0) pairs by |omega|, min_frequency filter, truncate to num_modes;
  // f = |omega| / 2 pi, whirl = sign(omega), no duplication.  The mode is the
  // a half of z, taken as omega b: the Ritz vectors are accurate in the L2
  // norm, which weighs b with K and a only with M.
  void deliver_qep(int b, HostProblem* hp, const double2* X, long long ldx) {
    tm_result* out = hp->out;
    const int M = hp->in.num_modes;
    std::vector<double> om(hp->n_extract);
    std::vector<int> ord;
    for (int u = 0; u < hp->n_extract; ++u) {
      const double mu = hp->theta_sorted_lambda[u];
      om[u] = mu != 0.0 ? 1.0 / mu : 0.0;
      if (std::fabs(om[u]) >= 1e-6) ord.push_back(u);
    }
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return std::fabs(om[x]) < std::fabs(om[y]); });
    std::vector<int> sel;
    int nf = 0, nb = 0;
    for (int u : ord) {
      const double w = om[u];
      if (w > 0 && nf < M) { sel.push_back(u); ++nf; }
      else if (w < 0 && nb < M) { sel.push_back(u); ++nb; }
      if (nf >= M && nb >= M) break;
    }
    std::vector<int> keep;
    for (int u : sel) {
      const double f = std::fabs(om[u]) / (2.0 * M_PI);
      if (hp->in.min_frequency > 0.0 && !(f > hp->in.min_frequency)) continue;
      keep.push_back(u);
    }
    if ((int)keep.size() > M) keep.resize(M);
    if (keep.empty()) {
      out->status = TM_ERR_NUMERIC;
      out->n_modes = 0;
      return;
    }
    const int nm = static_cast<int>(keep.size());
    std::vector<int> conjq(nm, 0);
    for (int q = 0; q < nm; ++q) {
      const int u = keep[q];
      const double w = om[u];
      if (out->frequencies) out->frequencies[q] = std::fabs(w) / (2.0 * M_PI);
      if (out->whirl) out->whirl[q] = w > 0 ? 1 : -1;
      if (out->eigenvalues) out->eigenvalues[q] = w;
      if (out->residuals) out->residuals[q] = hp->residual[u];
      if (out->modal_mass) out->modal_mass[q] = static_cast<double>(S_.N) * w * w * hp->xmx[u];
    }
    out->n_modes = nm;
    out->n_unique = nm;
    out->converged = hp->converged ? 1 : 0;
    out->restarts = hp->restarts;
    out->status = TM_OK;
    if (out->mode_shapes) {  // omega b
      std::vector<double> sc(nm);
      for (int q = 0; q < nm; ++q) sc[q] = om[keep[q]];
      expand_and_send(b, hp, X + n_, ldx, keep, conjq, sc);
    }
  }

  // -------------------------------------------------------------------------
  // Block shift-invert Krylov with explicit FP64 Rayleigh-Ritz.
  //
  // The basis V (M-orthonormal, up to mmax columns per problem) grows by
  // blocks of pb vectors: W = (K - sigma M)^-1 M Y for the block Y of V last
  // added (one multi-RHS triangular sweep pair per block); W is
  // M-orthogonalised against V twice (classical Gram-Schmidt) and within
  // itself (Cholesky QR).  K V and M V are kept beside V, and the projected
  // matrices H = V^H K V, G = V^H M V are extended by one column block per
  // step.  Near a full basis (and after every restart) a Rayleigh-Ritz step
  // with the FP64 K(k), M(k) checks the nev lowest Ritz pairs: accepted when
  // ||K x - lambda M x|| / ||K x|| <= max(kBlockTol, 4 x the FP64 floor of
  // that residual), or when it is below 1e-8 and stopped improving (the
  // rounding of x = V s sets a level a few times the floor).  A full basis
  // restarts from its nkeep lowest Ritz vectors and expands from the lowest
  // unconverged ones.
  // -------------------------------------------------------------------------
  static constexpr double kBlockTol = 5e-9;
  static constexpr double kBlockStall = 1e-8;
  static constexpr double kBlockStallMax = 1e-6;
  static constexpr double kQepStallMax = 8e-9;
  static constexpr int kBlockMaxRestarts = 40;
  // TMHIP_TRACE=1 prints the convergence history of the block method to stderr
  const bool trace_ = std::getenv("TMHIP_TRACE") != nullptr;
  bool block_ = false;
  // Coriolis pencil (L1, L2) on 2n-long vectors z = [a; b] instead of (K, M) on n
  bool qep_ = false;
  // TM_PREC_MIXED: complex-FP32 factor; every shift-invert solve refined in FP64
  bool mixed_ = false;
  bool tune_ = true;  // the batched factorizations may run the Schur-tile autotuner
  long long nvec_ = 0;  // basis vector length: n, or 2n for the pencil
  long long cs_ = 0;    // column stride of the vectors handed to the SpMVs
  std::vector<DevBuf> dblk_;
  std::vector<const double2*> dptr_host_;
  DevBuf dptr_;
  int pb_ = 0, mmax_ = 0, nkeep_ = 0, nts_ = 0;
  long long ldv_ = 0, ldw_ = 0, ldk_ = 0;
  DevBuf Vb_, MVb_, KVb_, Wb_, MWb_, KWb_, Xb_, KXb_, MXb_, tsp_, tsc_, tsp2_, tsc2_;
  std::vector<DevBuf> kblk_;
  std::vector<const double2*> kptr_host_;
  DevBuf kptr_;
  std::vector<SplitKM> sk_host_;  // per slot: the split K/M product's operands
  DevBuf sk_;
  std::vector<int> iter_;  // slot is iterating
  std::vector<int> nv_, y0_;
  std::vector<std::vector<std::complex<double>>> Hm_, Gm_;  // mmax x mmax per problem
  std::vector<std::vector<double>> prev_res_;

  void block_setup() {
    pb_ = block_width(nevmax_, qep_);
    mmax_ = basis_length(pb_, qep_, nevmax_);
    if (mmax_ > 2 * kTsMaxRows || mmax_ % pb_) throw ArgError("basis length");
    nkeep_ = mmax_ / 2;
    static_assert(kTsMaxRows >= 128, "tsmm coefficient rows");
    ldv_ = (long long)mmax_ * nvec_;
    ldw_ = (long long)pb_ * nvec_;
    ldk_ = (long long)nkeep_ * nvec_;
    nts_ = static_cast<int>((nvec_ + kTsChunk - 1) / kTsChunk);
    for (DevBuf* b : {&Vb_, &MVb_, &KVb_}) b->alloc(pool(), sizeof(double2) * ldv_ * B_);
    for (DevBuf* b : {&Wb_, &MWb_, &KWb_}) b->alloc(pool(), sizeof(double2) * ldw_ * B_);
    for (DevBuf* b : {&Xb_, &KXb_, &MXb_}) b->alloc(pool(), sizeof(double2) * ldk_ * B_);
    for (DevBuf* b : {&tsp_, &tsp2_}) b->alloc(pool(), sizeof(double2) * (size_t)B_ * mmax_ * 16 * nts_);
    for (DevBuf* b : {&tsc_, &tsc2_}) b->alloc(pool(), sizeof(double2) * (size_t)B_ * mmax_ * 16);
    Hm_.assign(B_, std::vector<std::complex<double>>((size_t)mmax_ * mmax_));
    Gm_.assign(B_, std::vector<std::complex<double>>((size_t)mmax_ * mmax_));
    prev_res_.assign(B_, std::vector<double>());
    kblk_.resize(B_);
    kptr_host_.assign(B_, nullptr);
    dblk_.resize(B_);
    dptr_host_.assign(B_, nullptr);
  }

  // FP64 K(k) blocks of the problems in the listed slots (residuals, Rayleigh-Ritz)
  void build_stiffness_blocks(const std::vector<int>& slots) {
    if (sk_host_.size() != (size_t)B_) sk_host_.assign(B_, SplitKM{});
    for (int b : slots) {
      HostProblem* hp = probs_[b];
      const Operator& sop = *static_cast<const Operator*>(hp->in.op);
      sk_host_[b] = SplitKM{sop.real_pack.as<double2>(), sop.real_slot.as<unsigned char>(), hp->s};
      if (!kblk_[b].p) kblk_[b].alloc(pool(), sizeof(double2) * 9 * op_.nblk);
      hipLaunchKernelGGL(k_build_phase_blocks, dim3((9 * (long long)op_.nblk + 255) / 256), dim3(256), 0, stream(),
                         (long long)op_.nblk, M_.d_slot_row.as<int>(), M_.d_colidx.as<int>(), M_.d_tslot.as<int>(),
                         M_.d_pinned.as<unsigned char>(), sop.Kc.as<double>(), sop.Kp.as<double>(),
                         sop.dKc.as<double>(), sop.dKp.as<double>(), hp->s, hp->pre, hp->pim,
                         kblk_[b].as<double2>(), 0, make_double2(1.0, 0.0));
      kptr_host_[b] = kblk_[b].as<double2>();
      if (qep_) {  // -D(k) = -i 2 omega G(k), G skew: Gc + P Gp - conj(P) Gp^T
        if (!dblk_[b].p) dblk_[b].alloc(pool(), sizeof(double2) * 9 * op_.nblk);
        hipLaunchKernelGGL(k_build_phase_blocks, dim3((9 * (long long)op_.nblk + 255) / 256), dim3(256), 0,
                           stream(), (long long)op_.nblk, M_.d_slot_row.as<int>(), M_.d_colidx.as<int>(),
                           M_.d_tslot.as<int>(),
                           M_.d_pinned.as<unsigned char>(), sop.Gc.as<double>(), sop.Gp.as<double>(),
                           (const double*)nullptr, (const double*)nullptr, 0.0, hp->pre, hp->pim,
                           dblk_[b].as<double2>(), 1, make_double2(0.0, -2.0 * hp->omega));
        dptr_host_[b] = dblk_[b].as<double2>();
      }
    }
    TM_HIP(hipGetLastError());
    kptr_ = upload(pool(), kptr_host_, stream());
    sk_ = upload(pool(), sk_host_, stream());
    if (qep_) dptr_ = upload(pool(), dptr_host_, stream());
  }

  DevBuf ints(const std::vector<int>& v) { return upload(pool(), v, stream()); }
  // v into the device buffer d, which stays in place from call to call (the
  // factorization's and the block sweeps' captured launches read it)
  template <class T>
  void put(DevBuf& d, const std::vector<T>& v) {
    if (d.bytes < sizeof(T) * v.size()) throw std::logic_error("put: buffer too small");
    TM_HIP(hipMemcpyAsync(d.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, stream()));
    TM_HIP(hipStreamSynchronize(stream()));
  }

  // C_b = A_b^H B_b (ka_b x kb, kb <= 16) over the first len rows (default: all)
  // of columns nvec_ apart, returned on the host as [b][a][c] (rows of 16)
  std::vector<double2> tsgemm(const double2* A, long long lda, const DevBuf& ka, int kamax, const double2* Bm,
                              long long ldb, int kb, const DevBuf& act, long long len = -1) {
    return tsgemm2(A, lda, ka, kamax, Bm, nullptr, ldb, kb, act, len).first;
  }
  // C_b = A_b^H B_b and, when B2 is given, C2_b = A_b^H B2_b from one pass over A
  std::pair<std::vector<double2>, std::vector<double2>> tsgemm2(const double2* A, long long lda, const DevBuf& ka,
                                                                int kamax, const double2* Bm, const double2* B2,
                                                                long long ldb, int kb, const DevBuf& act,
                                                                long long len = -1) {
    if (len < 0) len = nvec_;
    KTimer t(ctx_, "gram", 0, 8.0 * len * kamax * kb * B_ * (B2 ? 2 : 1));
    if (kb < 1 || kb > 16 || kamax < 1 || kamax > mmax_ || len > nvec_) throw std::logic_error("tsgemm shape");
    // 16-column blocks of A over the block's four waves: one wave per block up to four
    // blocks (row groups share the rest), two per wave beyond
    const int nab = (kamax + 15) / 16;
    const int wa = nab >= 3 ? 4 : nab, per = nab > 4 ? 2 : 1;
    const dim3 g(nts_, B_, (nab + wa * per - 1) / (wa * per));
    auto launch = [&](auto kern) {
      hipLaunchKernelGGL(kern, g, dim3(256), 0, stream(), A, lda, Bm, B2, ldb, len, nvec_, ka.as<int>(), kb,
                         act.as<int>(), mmax_, tsp_.as<double2>(), B2 ? tsp2_.as<double2>() : (double2*)nullptr,
                         nts_);
    };
    auto by_shape = [&](auto two) {
      constexpr bool T = decltype(two)::value;
      if (kb <= 8) {
        if (per == 2) launch(k_tsgemm_mfma<T, 1, 2, 4>);
        else if (wa == 4) launch(k_tsgemm_mfma<T, 1, 1, 4>);
        else if (wa == 2) launch(k_tsgemm_mfma<T, 1, 1, 2>);
        else launch(k_tsgemm_mfma<T, 1, 1, 1>);
      } else {
        if (per == 2) launch(k_tsgemm_mfma<T, 2, 2, 4>);
        else if (wa == 4) launch(k_tsgemm_mfma<T, 2, 1, 4>);
        else if (wa == 2) launch(k_tsgemm_mfma<T, 2, 1, 2>);
        else launch(k_tsgemm_mfma<T, 2, 1, 1>);
      }
    };
    if (B2) by_shape(std::true_type{});
    else by_shape(std::false_type{});
    hipLaunchKernelGGL(k_tsgemm_reduce, dim3((kamax * 16 + 255) / 256, B_), dim3(256), 0, stream(),
                       tsp_.as<double2>(), ka.as<int>(), kb, act.as<int>(), mmax_, nts_, tsc_.as<double2>());
    if (B2)
      hipLaunchKernelGGL(k_tsgemm_reduce, dim3((kamax * 16 + 255) / 256, B_), dim3(256), 0, stream(),
                         tsp2_.as<double2>(), ka.as<int>(), kb, act.as<int>(), mmax_, nts_, tsc2_.as<double2>());
    std::pair<std::vector<double2>, std::vector<double2>> h;
    h.first.resize((size_t)B_ * mmax_ * 16);
    TM_HIP(hipMemcpyAsync(h.first.data(), tsc_.p, sizeof(double2) * h.first.size(), hipMemcpyDeviceToHost, stream()));
    if (B2) {
      h.second.resize((size_t)B_ * mmax_ * 16);
      TM_HIP(hipMemcpyAsync(h.second.data(), tsc2_.p, sizeof(double2) * h.second.size(), hipMemcpyDeviceToHost,
                            stream()));
    }
    TM_HIP(hipStreamSynchronize(stream()));
    return h;
  }

  // Y_b[:, yoff_b + c] = beta Y + alpha sum_a A_b[:, a] C_b[a][c], c < kc (C on the host, [b][a][16])
  // (first len rows, default all; columns nvec_ apart)
  void tsmm(const double2* A, long long lda, const DevBuf& ka, const std::vector<double2>& C, int kc, double alpha,
            double beta, double2* Y, long long ldy, const DevBuf* yoff, const DevBuf& act, long long len = -1) {
    if (len < 0) len = nvec_;
    DevBuf dc = upload(pool(), C, stream());
    KTimer t(ctx_, "vec", 0, 0);
    for (int r0 = 0; r0 < mmax_; r0 += kTsMaxRows)
      hipLaunchKernelGGL(k_tsmm, dim3((len + 255) / 256, B_), dim3(256), 0, stream(), A, lda, len, nvec_,
                         dc.as<double2>(), ka.as<int>(), kc, act.as<int>(), mmax_, alpha, r0 ? 1.0 : beta, Y, ldy,
                         yoff ? yoff->as<int>() : (const int*)nullptr, r0);
  }

  // Y_b[:, c] (+)= A_b X_b[:, c], c < ncol_b (<= cmax), n rows; problems ld apart,
  // columns cs_ apart, for X and Y alike
  // (cmul ncol_b columns, cstride apart: default cs_; cmul = 2 with cstride
  // n_ treats the two halves of each 2n-long column as two columns)
  void spmv_cols(const DevBuf& blocks, const double2* X, long long ld, const DevBuf& ncol, int cmax, double2* Y,
                 bool accumulate = false, long long cstride = -1, int cmul = 1) {
    spmv_pair(blocks, nullptr, X, ld, ncol, cmax, Y, nullptr, accumulate, false, cstride, cmul);
  }
  // both Y1 (+)= A1 X and Y2 (+)= A2 X (blocks2 optional) from one pass over X
  void spmv_pair(const DevBuf& blocks1, const DevBuf* blocks2, const double2* X, long long ld, const DevBuf& ncol,
                 int cmax, double2* Y1, double2* Y2, bool acc1, bool acc2, long long cstride = -1, int cmul = 1) {
    KTimer t(ctx_, "spmv_m", 0, 0);
    if (cstride < 0) cstride = cs_ > 0 ? cs_ : n_;
    const long long ntask = (long long)((S_.nf + 3) / 4) * B_;
    if (ntask > std::numeric_limits<int>::max()) throw ArgError("spmv: too many (row quad, problem) tasks for one launch");
    const double2* const* b2 = blocks2 ? blocks2->as<const double2*>() : nullptr;
    for (int c0 = 0; c0 < cmax; c0 += 16) {  // cmax: the most columns (cmul ncol_b) of any problem
      if (cmax - c0 > 8)
        hipLaunchKernelGGL(k_spmv_pair<2>, dim3((unsigned)ntask), dim3(256), 0, stream(), S_.nf, B_,
                           M_.d_rowptr.as<int>(), M_.d_colidx.as<int>(), blocks1.as<const double2*>(), b2, X, ld,
                           cstride, ncol.as<int>(), cmul, c0, Y1, Y2, ld, acc1 ? 1 : 0, acc2 ? 1 : 0);
      else
        hipLaunchKernelGGL(k_spmv_pair<1>, dim3((unsigned)ntask), dim3(256), 0, stream(), S_.nf, B_,
                           M_.d_rowptr.as<int>(), M_.d_colidx.as<int>(), blocks1.as<const double2*>(), b2, X, ld,
                           cstride, ncol.as<int>(), cmul, c0, Y1, Y2, ld, acc1 ? 1 : 0, acc2 ? 1 : 0);
    }
  }

  // Y1 = K(k) X and Y2 = M(k) X of the Hermitian pencil in one pass, real slots
  // from the operators' shared arrays (k_spmv_km)
  void spmv_km(const double2* X, long long ld, const DevBuf& ncol, int cmax, double2* Y1, double2* Y2) {
    KTimer t(ctx_, "spmv_m", 0, 0);
    const long long cstride = cs_ > 0 ? cs_ : n_;
    const long long ntask = (long long)((S_.nf + 3) / 4) * B_;
    if (ntask > std::numeric_limits<int>::max()) throw ArgError("spmv: too many (row quad, problem) tasks for one launch");
    for (int c0 = 0; c0 < cmax; c0 += 16) {
      if (cmax - c0 > 8)
        hipLaunchKernelGGL(k_spmv_km<2>, dim3((unsigned)ntask), dim3(256), 0, stream(), S_.nf, B_,
                           M_.d_rowptr.as<int>(), M_.d_colidx.as<int>(), sk_.as<SplitKM>(),
                           kptr_.as<const double2*>(), mk_ptrs_.as<const double2*>(), X, ld, cstride,
                           ncol.as<int>(), c0, Y1, Y2, ld);
      else
        hipLaunchKernelGGL(k_spmv_km<1>, dim3((unsigned)ntask), dim3(256), 0, stream(), S_.nf, B_,
                           M_.d_rowptr.as<int>(), M_.d_colidx.as<int>(), sk_.as<SplitKM>(),
                           kptr_.as<const double2*>(), mk_ptrs_.as<const double2*>(), X, ld, cstride,
                           ncol.as<int>(), c0, Y1, Y2, ld);
    }
  }

  // The Rayleigh-Ritz pencil (A, B) with B positive definite: (K, M) with
  // eigenvalues lambda = omega^2, or for the Coriolis problem on z = [a; b]
  // (a = omega x, b = x) the definite pencil (Bc, L2), Bc = [[0, M], [M, -D]],
  // L2 = diag(M, K), with eigenvalues mu = 1 / omega.  Bc = L2 Op for the
  // shift-invert operator Op = L1^-1 L2 of the Lancaster pencil
  // L1 = [[D, K], [K, 0]] (L1 z = omega L2 z), so the Ritz values of this pencil
  // are those of Op in the L2 inner product: bounded by its spectrum, with the
  // wanted largest |mu| at its two ends (a Rayleigh-Ritz step on the indefinite
  // L1 itself puts spurious values into the gap around omega = 0).
  //
  // B z: M x, or L2 z = [M a; K b]
  void apply_B(const double2* X, long long ld, const DevBuf& ncol, int cmax, double2* Y) {
    spmv_cols(mk_ptrs_, X, ld, ncol, cmax, Y);
    if (qep_) spmv_cols(kptr_, X + n_, ld, ncol, cmax, Y + n_);
  }
  // both products, A z into AX and B z into BX.  For the Coriolis pencil one
  // pass of M over both halves gives M a and M b; then Bc z = [M b; M a - D b]
  // and L2 z = [M a; K b] need only K b and D b besides.
  void apply_AB(const double2* X, long long ld, const DevBuf& ncol, int cmax, double2* AX, double2* BX) {
    if (!qep_) {
      spmv_km(X, ld, ncol, cmax, AX, BX);
      return;
    }
    spmv_cols(mk_ptrs_, X, ld, ncol, 2 * cmax, BX, false, n_, 2);  // [M a; M b]
    {
      KTimer t(ctx_, "vec", 0, 0);
      hipLaunchKernelGGL(k_qep_swap_halves, dim3((n_ + 255) / 256, B_, cmax), dim3(256), 0, stream(), BX, AX, ld,
                         n_, nvec_, ncol.as<int>());
    }
    spmv_pair(kptr_, &dptr_, X + n_, ld, ncol, cmax, BX + n_, AX + n_, false, true);  // K b; + (-D) b
  }

  // A z: K x, or Bc z = [M b; M a - D b] (dptr_ holds -D)
  void apply_A(const double2* X, long long ld, const DevBuf& ncol, int cmax, double2* Y) {
    if (!qep_) {
      spmv_cols(kptr_, X, ld, ncol, cmax, Y);
      return;
    }
    spmv_cols(mk_ptrs_, X + n_, ld, ncol, cmax, Y);
    spmv_cols(mk_ptrs_, X, ld, ncol, cmax, Y + n_);
    spmv_cols(dptr_, X + n_, ld, ncol, cmax, Y + n_, true);
  }

  // M-orthogonalise W against V[:, 0:nv_b] (twice), orthonormalise it (Cholesky
  // QR), append it to V, MV, KV at column nv_b and extend H, G.  Returns the
  // problems whose block was rank deficient (nothing appended for them).
  std::vector<int> append_block(const std::vector<int>& act) {
    DevBuf dact = ints(act);
    std::vector<int> nvc(B_, 0), pbv(B_, 0);
    int nvmax = 0;
    for (int b = 0; b < B_; ++b)
      if (act[b]) {
        nvc[b] = nv_[b];
        pbv[b] = pb_;
        nvmax = std::max(nvmax, nv_[b]);
      }
    DevBuf dnv = ints(nvc), dpb = ints(pbv);
    if (nvmax > 0)
      for (int pass = 0; pass < 2; ++pass) {
        std::vector<double2> C = tsgemm(MVb_.as<double2>(), ldv_, dnv, nvmax, Wb_.as<double2>(), ldw_, pb_, dact);
        tsmm(Vb_.as<double2>(), ldv_, dnv, C, pb_, -1.0, 1.0, Wb_.as<double2>(), ldw_, nullptr, dact);
      }
    if (!qep_)
      apply_AB(Wb_.as<double2>(), ldw_, dpb, pb_, KWb_.as<double2>(), MWb_.as<double2>());
    else
      apply_B(Wb_.as<double2>(), ldw_, dpb, pb_, MWb_.as<double2>());
    std::vector<double2> G = tsgemm(Wb_.as<double2>(), ldw_, dpb, pb_, MWb_.as<double2>(), ldw_, pb_, dact);
    std::vector<double2> Rinv((size_t)B_ * mmax_ * 16, make_double2(0.0, 0.0));
    std::vector<int> bad, ok(act);
    for (int b = 0; b < B_; ++b) {
      if (!act[b]) continue;
      if (!chol_inverse(&G[(size_t)b * mmax_ * 16], &Rinv[(size_t)b * mmax_ * 16])) {
        bad.push_back(b);
        ok[b] = 0;
      }
    }
    DevBuf dok = ints(ok), doff = ints(nvc);
    tsmm(Wb_.as<double2>(), ldw_, dpb, Rinv, pb_, 1.0, 0.0, Vb_.as<double2>(), ldv_, &doff, dok);
    if (!qep_) {
      tsmm(MWb_.as<double2>(), ldw_, dpb, Rinv, pb_, 1.0, 0.0, MVb_.as<double2>(), ldv_, &doff, dok);
      tsmm(KWb_.as<double2>(), ldw_, dpb, Rinv, pb_, 1.0, 0.0, KVb_.as<double2>(), ldv_, &doff, dok);
    } else {
      // the Coriolis pencil's products of the new block by fresh SpMVs: scaling
      // the remainder's products by Rinv leaves errors of eps cond(R) (large once
      // Ritz vectors converge and the remainder is nearly singular), and the
      // shift-invert pencil does not filter them from its Ritz vectors
      tsmm(Wb_.as<double2>(), ldw_, dpb, Rinv, pb_, 1.0, 0.0, MWb_.as<double2>(), ldw_, nullptr, dok);
      std::vector<int> pok(B_, 0);
      for (int b = 0; b < B_; ++b) pok[b] = ok[b] ? pb_ : 0;
      DevBuf dpok = ints(pok);
      apply_AB(MWb_.as<double2>(), ldw_, dpok, pb_, KWb_.as<double2>(), Wb_.as<double2>());
      KTimer t(ctx_, "vec", 0, 0);
      hipLaunchKernelGGL(k_copy_cols, dim3((nvec_ + 255) / 256, B_), dim3(256), 0, stream(), Wb_.as<double2>(), ldw_,
                         (const int*)nullptr, MVb_.as<double2>(), ldv_, doff.as<int>(), pb_, nvec_, dok.as<int>());
      hipLaunchKernelGGL(k_copy_cols, dim3((nvec_ + 255) / 256, B_), dim3(256), 0, stream(), KWb_.as<double2>(), ldw_,
                         (const int*)nullptr, KVb_.as<double2>(), ldv_, doff.as<int>(), pb_, nvec_, dok.as<int>());
      std::fill(Rinv.begin(), Rinv.end(), make_double2(0.0, 0.0));  // the products below are of V_new itself
      for (int b = 0; b < B_; ++b)
        for (int c = 0; c < pb_; ++c) Rinv[(size_t)b * mmax_ * 16 + c * 16 + c] = make_double2(1.0, 0.0);
    }
    // new column block of H and G: V[:, 0:nv+pb]^H (K W, M W) Rinv
    std::vector<int> nvn(B_, 0);
    int nvnmax = 0;
    for (int b = 0; b < B_; ++b)
      if (ok[b]) {
        nvn[b] = nv_[b] + pb_;
        nvnmax = std::max(nvnmax, nvn[b]);
      }
    if (nvnmax > 0) {
      DevBuf dnvn = ints(nvn);
      // V^H (K W) and V^H (M W) from one pass over V
      auto hkm = tsgemm2(Vb_.as<double2>(), ldv_, dnvn, nvnmax, KWb_.as<double2>(),
                         (qep_ ? Wb_ : MWb_).as<double2>(), ldw_, pb_, dok);
      const std::vector<double2>& hk = hkm.first;
      const std::vector<double2>& hm = hkm.second;
      using cd = std::complex<double>;
      for (int b = 0; b < B_; ++b) {
        if (!ok[b]) continue;
        const int n0 = nv_[b], n1 = nvn[b];
        const double2* R = &Rinv[(size_t)b * mmax_ * 16];
        for (int a = 0; a < n1; ++a)
          for (int c = 0; c < pb_; ++c) {
            cd sk(0, 0), sm(0, 0);
            for (int q = 0; q <= c; ++q) {
              const cd r(R[q * 16 + c].x, R[q * 16 + c].y);
              const double2 xk = hk[((size_t)b * mmax_ + a) * 16 + q], xm = hm[((size_t)b * mmax_ + a) * 16 + q];
              sk += cd(xk.x, xk.y) * r;
              sm += cd(xm.x, xm.y) * r;
            }
            Hm_[b][(size_t)a * mmax_ + n0 + c] = sk;
            Gm_[b][(size_t)a * mmax_ + n0 + c] = sm;
          }
        for (int a = n0; a < n1; ++a)  // Hermitian completion (the new diagonal block is averaged)
          for (int c = 0; c < n1; ++c) {
            if (c >= n0 && c < a) continue;
            const cd hk2 = std::conj(Hm_[b][(size_t)c * mmax_ + a]), gm2 = std::conj(Gm_[b][(size_t)c * mmax_ + a]);
            if (c >= n0) {
              const cd h1 = 0.5 * (Hm_[b][(size_t)a * mmax_ + c] + hk2);
              const cd g1 = 0.5 * (Gm_[b][(size_t)a * mmax_ + c] + gm2);
              Hm_[b][(size_t)a * mmax_ + c] = h1;
              Hm_[b][(size_t)c * mmax_ + a] = std::conj(h1);
              Gm_[b][(size_t)a * mmax_ + c] = g1;
              Gm_[b][(size_t)c * mmax_ + a] = std::conj(g1);
            } else {
              Hm_[b][(size_t)a * mmax_ + c] = hk2;
              Gm_[b][(size_t)a * mmax_ + c] = gm2;
            }
          }
        nv_[b] = n1;
      }
    }
    return bad;
  }

  // G (pb x pb, rows of 16) = R^H R; Rinv = R^-1 (upper).  False if G is not
  // numerically positive definite.
  bool chol_inverse(const double2* G, double2* Rinv) const {
    using cd = std::complex<double>;
    const int p = pb_;
    std::vector<cd> R((size_t)p * p, cd(0, 0));
    double gmax = 0;
    for (int i = 0; i < p; ++i) gmax = std::max(gmax, std::fabs(G[i * 16 + i].x));
    for (int j = 0; j < p; ++j) {
      double d = G[j * 16 + j].x;
      for (int k = 0; k < j; ++k) d -= std::norm(R[(size_t)k * p + j]);
      if (!(d > 1e-20 * gmax) || !(gmax > 0)) return false;
      const double rjj = std::sqrt(d);
      R[(size_t)j * p + j] = rjj;
      for (int i = j + 1; i < p; ++i) {
        cd s(G[j * 16 + i].x, G[j * 16 + i].y);  // (W^H M W)[j][i]
        for (int k = 0; k < j; ++k) s -= std::conj(R[(size_t)k * p + j]) * R[(size_t)k * p + i];
        R[(size_t)j * p + i] = s / rjj;
      }
    }
    for (int c = 0; c < p; ++c)  // R X = I, column by column
      for (int i = c; i >= 0; --i) {
        cd s = i == c ? cd(1, 0) : cd(0, 0);
        for (int k = i + 1; k <= c; ++k) s -= R[(size_t)i * p + k] * cd(Rinv[k * 16 + c].x, Rinv[k * 16 + c].y);
        const cd x = s / R[(size_t)i * p + i];
        Rinv[i * 16 + c] = make_double2(x.real(), x.imag());
      }
    return true;
  }

  // W[:, 0:pb] of the selected slots <- seeded normal vectors (zero on pinned
  // DOFs); the seed depends on the column and the slot's salt only, so a
  // problem's start and restart vectors do not depend on its batch
  void random_block(const std::vector<int>& sel, const std::vector<int>& salt) {
    std::map<int, std::vector<int>> by_salt;
    for (int b = 0; b < B_; ++b)
      if (sel[b]) by_salt[salt[b]].push_back(b);
    for (const auto& g : by_salt) {
      if (g.first == 0) {  // every problem's first start block: made once per slot setup, then copied
        const size_t bytes = sizeof(double2) * (size_t)pb_ * nvec_;
        if (!rand0_.p) {
          rand0_.alloc(pool(), bytes);
          for (int c = 0; c < pb_; ++c) {
            std::vector<double> v = random_vector(7919u * (uint32_t)(c + 1));
            if (qep_) {  // the b half from its own seed
              std::vector<double> v2 = random_vector(7919u * (uint32_t)(c + 1) + 31u);
              v.insert(v.end(), v2.begin(), v2.end());
            }
            TM_HIP(hipMemcpyAsync(rand0_.as<double2>() + (long long)c * nvec_, v.data(), sizeof(double2) * nvec_,
                                  hipMemcpyHostToDevice, stream()));
            TM_HIP(hipStreamSynchronize(stream()));
          }
        }
        for (int b : g.second)  // slot b's block: pb columns nvec apart, contiguous
          TM_HIP(hipMemcpyAsync(Wb_.as<double2>() + b * ldw_, rand0_.p, bytes, hipMemcpyDeviceToDevice, stream()));
        continue;
      }
      for (int c = 0; c < pb_; ++c) {
        std::vector<double> v = random_vector(7919u * (uint32_t)(c + 1) + 104729u * (uint32_t)g.first);
        if (qep_) {  // the b half from its own seed
          std::vector<double> v2 = random_vector(7919u * (uint32_t)(c + 1) + 104729u * (uint32_t)g.first + 31u);
          v.insert(v.end(), v2.begin(), v2.end());
        }
        const int first = g.second[0];
        TM_HIP(hipMemcpyAsync(Wb_.as<double2>() + first * ldw_ + (long long)c * nvec_, v.data(),
                              sizeof(double2) * nvec_, hipMemcpyHostToDevice, stream()));
        for (size_t i = 1; i < g.second.size(); ++i)
          TM_HIP(hipMemcpyAsync(Wb_.as<double2>() + g.second[i] * ldw_ + (long long)c * nvec_,
                                Wb_.as<double2>() + first * ldw_ + (long long)c * nvec_, sizeof(double2) * nvec_,
                                hipMemcpyDeviceToDevice, stream()));
        TM_HIP(hipStreamSynchronize(stream()));
      }
    }
  }
  DevBuf rand0_;  // the salt-0 start block (pb columns, nvec apart), made on first use per slot setup

  // ---- slots: continuous batching ----------------------------------------
 public:
  // a serving loop: factorizations use tuned tables but never tune (no device-wide pause)
  void no_tuning() { tune_ = false; }
  // Allocates the workspaces of `capacity` slots for problems of at most nevmax
  // wanted pairs (one block size).
  // elem: the factor's element type (FactorElem::C32 is TM_PREC_MIXED; R64 takes real
  // harmonics only)
  void slot_setup(int capacity, int nevmax, bool qep = false, FactorElem elem = FactorElem::C64) {
    HostTimer h(ctx_, "setup");
    B_ = capacity;
    probs_.assign(B_, nullptr);
    hold_.assign(B_, 0);
    n_ = 3LL * S_.nf;
    qep_ = qep;
    mixed_ = elem == FactorElem::C32;
    nvec_ = qep ? 2 * n_ : n_;
    cs_ = nvec_;
    nevmax_ = nevmax;
    ncvmax_ = 1;
    nchunk_ = static_cast<int>((nvec_ + kDotChunk - 1) / kDotChunk);
    const int pcols = std::min(16, std::max(block_width(nevmax_, qep_), nevmax_));
    bm_.alloc(pool(), P_, B_, pcols, elem);
    act_.alloc(pool(), sizeof(int) * B_);
    gmask_.alloc(pool(), sizeof(int) * B_);
    gact_.alloc(pool(), sizeof(int) * B_);
    gops_.alloc(pool(), sizeof(OpArrays) * B_);
    h_params_.assign(B_, ProbParam{1.0, 0.0, 0.0, 1.0});
    h_ops_.assign(B_, op_arrays(op_));
    mk_.assign(B_, nullptr);
    h_ncv_.assign(B_, 0);
    iter_.assign(B_, 0);
    nv_.assign(B_, 0);
    y0_.assign(B_, 0);
    block_setup();
    rand0_ = DevBuf();
  }

  // a slot is free once it holds no problem and no problem in flight reads the shared
  // fronts its last factorization left there
  bool slot_free(int b) const { return !probs_[b] && hold_[b] == 0; }
  // classes (shared-front leaders) of the slots in `mask`; -1 without a class table
  int classes(const std::vector<int>& mask) const {
    if (bm_.nlead == 0 || (int)bm_.lead_h.size() != B_) return -1;
    std::vector<char> seen(B_, 0);
    int n = 0;
    for (int b = 0; b < B_; ++b)
      if (mask[b] && !seen[bm_.lead_h[b]]) {
        seen[bm_.lead_h[b]] = 1;
        ++n;
      }
    return n;
  }
  int free_slots() const {
    int n = 0;
    for (int b = 0; b < B_; ++b) n += slot_free(b) ? 1 : 0;
    return n;
  }
  bool busy() const { return free_slots() < B_; }
  bool stepping() const { return std::any_of(iter_.begin(), iter_.end(), [](int v) { return v != 0; }); }
  int capacity() const { return B_; }

  // Puts problems into free slots, factors them (only them), and starts their
  // block Krylov bases; they then step with the problems already in flight.
  void admit(std::vector<HostProblem*> adm) {
    if (!admit_prepare(std::move(adm))) return;
    admit_factor(ctx_);
    admit_complete();
  }

  // The same in two halves for a serving loop: admit_async puts the problems into
  // free slots and launches their factorization on `side` (a context of its own:
  // stream, graph cache), after the slots' uploads and block builds on this
  // context's stream; the slots already in flight keep stepping meanwhile (the
  // factorization writes only the new slots' workspaces).  admit_ready tells when it
  // has finished, admit_finish then starts the new slots' bases.  One admission at
  // a time; the new slots neither step nor retire before admit_finish.
  void admit_async(std::vector<HostProblem*> adm, Context* side) {
    if (admitting()) throw std::logic_error("admit_async: an admission is pending");
    if (!admit_prepare(std::move(adm))) return;
    if (!side_ev_[0]) {
      TM_HIP(hipEventCreateWithFlags(&side_ev_[0], hipEventDisableTiming));
      TM_HIP(hipEventCreateWithFlags(&side_ev_[1], hipEventDisableTiming));
    }
    TM_HIP(hipEventRecord(side_ev_[0], stream()));  // neither stream is capturing here
    TM_HIP(hipStreamWaitEvent(side->stream, side_ev_[0], 0));
    admit_factor(side);
    TM_HIP(hipEventRecord(side_ev_[1], side->stream));
    pend_side_ = side;
  }
  bool admitting() const { return pend_side_ != nullptr; }
  bool admit_ready() const {
    if (!pend_side_) return false;
    const hipError_t q = hipEventQuery(side_ev_[1]);
    if (q == hipErrorNotReady) return false;
    TM_HIP(q);
    return true;
  }
  void admit_finish() {
    if (!pend_side_) return;
    TM_HIP(hipEventSynchronize(side_ev_[1]));
    pend_side_ = nullptr;
    admit_complete();
  }

 private:
  // slots, parameters, M(k) and K(k) blocks of the admitted problems (this context's stream)
  bool admit_prepare(std::vector<HostProblem*> adm) {
    adm_slots_.clear();
    adm_mask_.assign(B_, 0);
    const auto now = std::chrono::steady_clock::now();
    adm_t0_ = now;
    for (int b = 0; b < B_ && !adm.empty(); ++b)
      if (slot_free(b)) {
        HostProblem* hp = adm.front();
        adm.erase(adm.begin());
        probs_[b] = hp;
        hp->t_admit = now;
        // the pencil's shift-invert operator needs K(k) itself: shift 0
        h_params_[b] = {hp->pre, hp->pim, hp->s, qep_ ? 0.0 : hp->in.shift};
        Operator& sop = *static_cast<Operator*>(hp->in.op);
        if (sop.mesh != &M_) throw ArgError("admit: problem of another mesh");
        if (bm_.elem == FactorElem::R64 && !real_harmonic(S_.N, hp->k))
          throw std::logic_error("admit: a complex harmonic in a real-factor slot group");
        mk_[b] = harmonic_mass(ctx_, sop, hp->k);
        h_ops_[b] = op_arrays(sop);
        nv_[b] = 0;
        y0_[b] = 0;
        prev_res_[b].clear();
        adm_slots_.push_back(b);
        adm_mask_[b] = 1;
      }
    if (!adm.empty()) throw ArgError("admit: more problems than free slots");
    if (adm_slots_.empty()) return false;
    admit_lap(0);
    TM_HIP(hipMemcpyAsync(bm_.params.p, h_params_.data(), sizeof(ProbParam) * B_, hipMemcpyHostToDevice, stream()));
    std::vector<int> zero(B_, 0);
    TM_HIP(hipMemcpyAsync(bm_.status.p, zero.data(), sizeof(int) * B_, hipMemcpyHostToDevice, stream()));
    mk_ptrs_ = upload(pool(), mk_, stream());
    build_stiffness_blocks(adm_slots_);
    put(gmask_, adm_mask_);
    put(gops_, h_ops_);
    admit_lap(1);
    return true;
  }
  // the admitted slots' factorization on c's stream
  void admit_factor(Context* c) {
    HostTimer h(ctx_, "factor");
    bm_.set_classes(adm_slots_, h_params_, h_ops_, B_, c->stream);
    for (int b : adm_slots_) hold_[bm_.lead_h[b]] += 1;
    BatchBufs bb = bm_.bufs(P_, B_);
    factor_batch(c, M_, P_, op_arrays(op_), bb, gmask_.as<int>(), static_cast<int>(adm_slots_.size()),
                 gops_.as<OpArrays>(), true, tune_);
    admit_lap(2);
  }
  // status of the factorizations, start blocks of the admitted slots
  void admit_complete() {
    std::vector<int>& mask = adm_mask_;
    {
      std::vector<int> status(B_);
      TM_HIP(hipMemcpyAsync(status.data(), bm_.status.p, sizeof(int) * B_, hipMemcpyDeviceToHost, stream()));
      TM_HIP(hipStreamSynchronize(stream()));
      admit_lap(3);
      for (int b : adm_slots_) {
        probs_[b]->factor_flags = status[b];
        if (status[b]) {
          probs_[b]->failed = true;
          probs_[b]->message = factor_message(status[b]);
          mask[b] = 0;
        }
      }
    }
    HostTimer h(ctx_, "iterate");
    random_block(mask, std::vector<int>(B_, 0));
    for (int b : adm_slots_) {  // rpm continuation: the adjacent point's Ritz vectors lead the start block
      const std::pair<DevBuf, int>* src = mask[b] ? warm_source(probs_[b]) : nullptr;
      if (!src) continue;
      TM_HIP(hipMemcpyAsync(Wb_.as<double2>() + b * ldw_, src->first.p, sizeof(double2) * (size_t)src->second * n_,
                            hipMemcpyDeviceToDevice, stream()));
      probs_[b]->warm_started = true;
    }
    warm_prune(adm_slots_);
    if (qep_) smooth_block(mask);
    append_or_retry(mask);
    for (int b : adm_slots_)
      if (mask[b]) {
        iter_[b] = 1;
        y0_[b] = nv_[b] - pb_;
      }
    admit_lap(4);
    adm_slots_.clear();
  }
  void admit_lap(int part) {
    const auto t2 = std::chrono::steady_clock::now();
    admit_s_[part] += std::chrono::duration<double>(t2 - adm_t0_).count();
    adm_t0_ = t2;
  }
  std::vector<int> adm_slots_, adm_mask_;  // the admission in progress
  std::chrono::steady_clock::time_point adm_t0_;
  Context* pend_side_ = nullptr;           // the context of a pending asynchronous factorization
  hipEvent_t side_ev_[2] = {nullptr, nullptr};

 public:
  // host seconds of the admissions so far: M(k) lookups, uploads and block builds,
  // factorization launches (captures included), waiting for the factorizations, start blocks
  double admit_s_[5] = {0, 0, 0, 0, 0};

  // One block step of every iterating slot: expansion, orthogonalisation,
  // Rayleigh-Ritz and convergence (finished slots stop iterating).
  void block_step() {
    HostTimer h(ctx_, "iterate");
    bool any = false;
    for (int b = 0; b < B_; ++b) any = any || iter_[b];
    if (!any) return;
    BatchBufs bb = bm_.bufs(P_, B_);
    put(gact_, iter_);
    const DevBuf& dact = gact_;
    DevBuf dy0 = ints(y0_);
    const int nit = (int)std::count_if(iter_.begin(), iter_.end(), [](int v) { return v != 0; });
    if (!qep_) {
      // W = (K - sigma M)^-1 M Y, Y = V[:, y0 : y0 + pb]
      {
        KTimer t(ctx_, "vec", 0, 0);
        hipLaunchKernelGGL(k_copy_cols, dim3((n_ + 255) / 256, B_), dim3(256), 0, stream(), MVb_.as<double2>(), ldv_,
                           dy0.as<int>(), Wb_.as<double2>(), ldw_, (const int*)nullptr, pb_, n_, dact.as<int>());
      }
      trisolve_batch(ctx_, M_, P_, bb, Wb_.as<double2>(), ldw_, dact.as<int>(), pb_, -1, nit, true, classes(iter_));
      if (mixed_) {
        std::vector<int> nc(B_, 0);
        for (int b = 0; b < B_; ++b) nc[b] = iter_[b] ? pb_ : 0;
        mixed_refine(Wb_.as<double2>(), MVb_.as<double2>(), ldv_, &dy0, nc);
      }
    } else {
      // W = L1^-1 L2 Y: for Y = [a; b], W = [b; K^-1 (M a - D b)] (one solve with K(k))
      std::vector<int> pbv(B_, 0);
      for (int b = 0; b < B_; ++b) pbv[b] = iter_[b] ? pb_ : 0;
      DevBuf dpb = ints(pbv);
      {
        KTimer t(ctx_, "vec", 0, 0);
        // W_a <- Y_b ;  W_b <- (M Y_a) = the a half of L2 Y
        hipLaunchKernelGGL(k_axpy_half, dim3((n_ + 255) / 256, B_), dim3(256), 0, stream(), Vb_.as<double2>(), ldv_,
                           dy0.as<int>(), n_, Wb_.as<double2>(), ldw_, (const int*)nullptr, 0LL, pb_, n_, nvec_, 1.0,
                           0.0, dact.as<int>());
        hipLaunchKernelGGL(k_axpy_half, dim3((n_ + 255) / 256, B_), dim3(256), 0, stream(), MVb_.as<double2>(), ldv_,
                           dy0.as<int>(), 0LL, Wb_.as<double2>(), ldw_, (const int*)nullptr, n_, pb_, n_, nvec_, 1.0,
                           0.0, dact.as<int>());
      }
      // W_b += (-D) Y_b
      spmv_cols(dptr_, Wb_.as<double2>(), ldw_, dpb, pb_, Wb_.as<double2>() + n_, true);
      trisolve_batch(ctx_, M_, P_, bb, Wb_.as<double2>() + n_, ldw_, dact.as<int>(), pb_, nvec_, nit, true,
                     classes(iter_));
    }
    for (int b = 0; b < B_; ++b)
      if (iter_[b]) probs_[b]->steps += 1;
    append_or_retry(iter_);
    for (int b = 0; b < B_; ++b)
      if (iter_[b]) y0_[b] = nv_[b] - pb_;
    rayleigh_ritz(iter_);
    ctx_->flush_timing();
  }

  // Extracts the results of every finished slot and frees it; returns the
  // problems that finished.
  std::vector<HostProblem*> retire() {
    std::vector<int> sel(B_, 0);
    std::vector<HostProblem*> out;
    for (int b = 0; b < B_; ++b)
      if (probs_[b] && !iter_[b] && !(pend_side_ && adm_mask_[b])) {
        sel[b] = 1;
        out.push_back(probs_[b]);
      }
    if (out.empty()) return out;
    {
      HostTimer h(ctx_, "extract");
      extract(sel);
      warm_keep(sel);
    }
    const auto now = std::chrono::steady_clock::now();
    for (int b = 0; b < B_; ++b)
      if (sel[b]) {
        HostProblem* hp = probs_[b];
        if (hp->out) hp->out->seconds = std::chrono::duration<double>(now - hp->t_admit).count();
        probs_[b] = nullptr;
        hold_[bm_.lead_h[b]] -= 1;
      }
    return out;
  }

  // Streams the queue through the slots: refills once `refill` slots (or all
  // the rest of the queue) can be admitted together.  refill = capacity drains
  // the slots between admissions (best when the whole queue is known: the
  // factorizations stay fully batched); smaller values let problems that
  // converge early hand their slots on.
  void run_slots(std::vector<HostProblem*> queue, int capacity, int refill, bool qep = false,
                 FactorElem elem = FactorElem::C64) {
    int nevmax = 1;
    for (auto* hp : queue) nevmax = std::max(nevmax, hp->nev);
    slot_setup(capacity, nevmax, qep, elem);
    warm_rpms_.clear();
    warm_store_.clear();
    warm_admitted_.clear();
    if (!qep)
      for (auto* hp : queue)
        if (hp->in.warm_start) warm_rpms_[warm_series(hp)].push_back(hp->in.rpm);
    for (auto& kv : warm_rpms_) {
      std::sort(kv.second.begin(), kv.second.end());
      kv.second.erase(std::unique(kv.second.begin(), kv.second.end()), kv.second.end());
    }
    refill = std::max(1, std::min(refill, capacity));
    size_t next = 0;
    while (true) {
      const int nfree = free_slots();
      const size_t left = queue.size() - next;
      if (left > 0 && nfree > 0 && (nfree >= refill || (size_t)nfree >= left)) {
        const size_t take = std::min(left, (size_t)nfree);
        admit(std::vector<HostProblem*>(queue.begin() + next, queue.begin() + next + take));
        next += take;
      }
      if (!busy()) break;
      block_step();
      retire();
    }
  }

 private:
  std::vector<ProbParam> h_params_;
  std::vector<OpArrays> h_ops_;  // per slot: the problem's operator (material) on the shared mesh
  // per slot, in place for the life of the slots: the factorization's mask and
  // operators and the block sweeps' active mask (read by captured launches)
  DevBuf gmask_, gact_, gops_;

  // rpm continuation (tm_problem.warm_start): the converged Ritz vectors of every
  // finished problem of the run, by (operator, harmonic, wanted pairs) and rpm, so that a
  // problem admitted later at an adjacent rpm point starts from them.  Only problems of
  // this slot group and of earlier waves are used: the choice depends on the run alone.
  using WarmSeries = std::tuple<const void*, int, int>;
  std::map<WarmSeries, std::vector<double>> warm_rpms_;                 // sorted rpm points
  std::map<std::pair<WarmSeries, double>, std::pair<DevBuf, int>> warm_store_;  // vectors, columns
  std::set<std::pair<WarmSeries, double>> warm_admitted_;
  static WarmSeries warm_series(const HostProblem* hp) {
    return WarmSeries{hp->in.op, hp->k, hp->nev};
  }
  // the stored Ritz vectors of the adjacent rpm point (lower one first), or null
  const std::pair<DevBuf, int>* warm_source(const HostProblem* hp) const {
    if (!hp->in.warm_start || qep_) return nullptr;
    const WarmSeries key = warm_series(hp);
    auto it = warm_rpms_.find(key);
    if (it == warm_rpms_.end()) return nullptr;
    const auto& r = it->second;
    const long idx = std::lower_bound(r.begin(), r.end(), hp->in.rpm) - r.begin();
    for (long j : {idx - 1, idx + 1}) {
      if (j < 0 || j >= (long)r.size()) continue;
      auto f = warm_store_.find({key, r[j]});
      if (f != warm_store_.end()) return &f->second;
    }
    return nullptr;
  }
  // records the admission of the problems in slots and frees the stored vectors that no
  // problem still to come can use (both adjacent rpm points admitted)
  void warm_prune(const std::vector<int>& slots) {
    for (int b : slots)
      if (probs_[b]->in.warm_start) warm_admitted_.insert({warm_series(probs_[b]), probs_[b]->in.rpm});
    for (auto it = warm_store_.begin(); it != warm_store_.end();)
      it = warm_needed(it->first.first, it->first.second) ? std::next(it) : warm_store_.erase(it);
  }
  bool warm_needed(const WarmSeries& key, double rpm) const {
    const auto& r = warm_rpms_.at(key);
    const long idx = std::lower_bound(r.begin(), r.end(), rpm) - r.begin();
    for (long j : {idx - 1, idx + 1})
      if (j >= 0 && j < (long)r.size() && !warm_admitted_.count({key, r[j]})) return true;
    return false;
  }
  // keeps the Ritz vectors of the finished, converged problems of sel (columns 0.. of X)
  void warm_keep(const std::vector<int>& sel) {
    for (int b = 0; b < B_; ++b) {
      HostProblem* hp = sel[b] ? probs_[b] : nullptr;
      if (!hp || !hp->in.warm_start || qep_ || hp->failed || !hp->converged || !hp->done) continue;
      const int nc = std::min(hp->n_extract, pb_);
      if (nc <= 0 || !warm_needed(warm_series(hp), hp->in.rpm)) continue;
      DevBuf v(pool(), sizeof(double2) * (size_t)nc * n_);
      TM_HIP(hipMemcpyAsync(v.p, Xb_.as<double2>() + b * ldk_, sizeof(double2) * (size_t)nc * n_,
                            hipMemcpyDeviceToDevice, stream()));
      warm_store_[{warm_series(hp), hp->in.rpm}] = std::make_pair(std::move(v), nc);
    }
  }

  // Appends W; a problem whose block lost rank (an invariant subspace was
  // reached) continues from fresh random directions, up to three times.
  // Coriolis pencil: W <- Op^2 W for the selected slots.  A random block has
  // O(1) content in the stiff modes, which the L1 residuals amplify by
  // |omega_max / omega|; two applications of Op damp it by (omega / omega_max)^2
  // before it enters the basis.
  void smooth_block(const std::vector<int>& sel) {
    std::vector<int> pbv(B_, 0);
    for (int b = 0; b < B_; ++b) pbv[b] = sel[b] ? pb_ : 0;
    DevBuf dsel = ints(sel), dpb = ints(pbv);
    BatchBufs bb = bm_.bufs(P_, B_);
    const dim3 g((n_ + 255) / 256, B_);
    for (int rep = 0; rep < 2; ++rep) {
      spmv_cols(mk_ptrs_, Wb_.as<double2>(), ldw_, dpb, pb_, MWb_.as<double2>());  // M a
      {
        KTimer t(ctx_, "vec", 0, 0);
        hipLaunchKernelGGL(k_axpy_half, g, dim3(256), 0, stream(), Wb_.as<double2>(), ldw_, (const int*)nullptr, n_,
                           Wb_.as<double2>(), ldw_, (const int*)nullptr, 0LL, pb_, n_, nvec_, 1.0, 0.0,
                           dsel.as<int>());  // a <- b
        hipLaunchKernelGGL(k_axpy_half, g, dim3(256), 0, stream(), MWb_.as<double2>(), ldw_, (const int*)nullptr, 0LL,
                           Wb_.as<double2>(), ldw_, (const int*)nullptr, n_, pb_, n_, nvec_, 1.0, 0.0,
                           dsel.as<int>());  // b <- M a
      }
      spmv_cols(dptr_, Wb_.as<double2>(), ldw_, dpb, pb_, Wb_.as<double2>() + n_, true);  // += (-D) b
      trisolve_batch(ctx_, M_, P_, bb, Wb_.as<double2>() + n_, ldw_, dsel.as<int>(), pb_, nvec_,
                     (int)std::count_if(sel.begin(), sel.end(), [](int v) { return v != 0; }), false, classes(sel));
    }
  }

  void append_or_retry(std::vector<int>& act) {
    std::vector<int> bad = append_block(act);
    for (int attempt = 1; attempt <= 3 && !bad.empty(); ++attempt) {
      std::vector<int> sub(B_, 0), salt(B_, 0);
      for (int b : bad) {
        sub[b] = 1;
        salt[b] = 1000 * probs_[b]->steps + attempt;
      }
      random_block(sub, salt);
      if (qep_) smooth_block(sub);
      bad = append_block(sub);
    }
    for (int b : bad) fail_block(b, act);
  }

  void fail_block(int b, std::vector<int>& act) {
    probs_[b]->failed = true;
    probs_[b]->message = "block Krylov basis lost rank";
    act[b] = 0;
    if (b < (int)iter_.size()) iter_[b] = 0;
  }

  // X[:, c0:c1] = V S[:, c0:c1] for the problems in sel; K X and M X by fresh
  // products (no drift of K V, M V across restarts)
  // For the Coriolis problem with mu given (mu = 1 / omega per column), the a
  // half is set to omega b: z = [omega x; x] with x = b, the vector delivered.
  void ritz_vectors(const std::vector<int>& sel, const std::vector<std::vector<double>>& Sv, int c0, int c1,
                    const std::vector<std::vector<double>>* mu = nullptr) {
    std::vector<int> nvc(B_, 0);
    for (int b = 0; b < B_; ++b)
      if (sel[b]) nvc[b] = nv_[b];
    DevBuf dsel = ints(sel), dnv = ints(nvc);
    for (int q0 = c0; q0 < c1; q0 += 16) {
      const int kc = std::min(16, c1 - q0);
      std::vector<double2> C((size_t)B_ * mmax_ * 16, make_double2(0.0, 0.0));
      for (int b = 0; b < B_; ++b)
        if (sel[b])
          for (int a = 0; a < nvc[b]; ++a)
            for (int c = 0; c < kc && q0 + c < nvc[b]; ++c) {
              const size_t s = (size_t)a * nvc[b] + q0 + c;
              C[((size_t)b * mmax_ + a) * 16 + c] = make_double2(Sv[b][2 * s], Sv[b][2 * s + 1]);
            }
      std::vector<int> off(B_, q0);
      DevBuf doff = ints(off);
      tsmm(Vb_.as<double2>(), ldv_, dnv, C, kc, 1.0, 0.0, Xb_.as<double2>(), ldk_, &doff, dsel);
      if (qep_ && mu) {
        for (int b = 0; b < B_; ++b)
          if (sel[b])
            for (int a = 0; a < nvc[b]; ++a)
              for (int c = 0; c < kc && q0 + c < nvc[b]; ++c) {
                const double m = (*mu)[b][q0 + c], w = m != 0.0 ? 1.0 / m : 0.0;
                double2& x = C[((size_t)b * mmax_ + a) * 16 + c];
                x = make_double2(w * x.x, w * x.y);
              }
        tsmm(Vb_.as<double2>() + n_, ldv_, dnv, C, kc, 1.0, 0.0, Xb_.as<double2>(), ldk_, &doff, dsel, n_);
      }
    }
    std::vector<int> ncol(B_, 0);
    for (int b = 0; b < B_; ++b)
      if (sel[b]) ncol[b] = c1 - c0;
    DevBuf dn = ints(ncol);
    const long long o = (long long)c0 * nvec_;
    apply_AB(Xb_.as<double2>() + o, ldk_, dn, c1 - c0, KXb_.as<double2>() + o, MXb_.as<double2>() + o);
  }

  // Rayleigh-Ritz on span(V) for the problems near a full basis or restarted;
  // convergence, restart and the next expansion block.
  void rayleigh_ritz(std::vector<int>& act) {
    std::vector<int> rr(B_, 0);
    bool anyrr = false;
    for (int b = 0; b < B_; ++b) {
      if (!act[b]) continue;
      HostProblem* hp = probs_[b];
      // a warm-started problem also checks after its fifth and later blocks
      const bool early = hp->warm_started && nv_[b] >= 5 * pb_;
      const bool due = nv_[b] >= mmax_ - pb_ || hp->restarts > 0 || nv_[b] + pb_ > mmax_ || early;
      if (due && nv_[b] >= std::min(hp->nev + pb_, mmax_)) {
        rr[b] = 1;
        anyrr = true;
      }
    }
    if (!anyrr) return;
    std::vector<std::vector<double>> lam(B_), Sv(B_);
    std::vector<int> ok(B_, 1);
    {
      HostTimer h(ctx_, "rr_eig");
      parallel_for(B_, [&](int b) {
        if (!rr[b]) return;
        // (V^H A V, V^H B V): (K, M), or the Coriolis definite pencil (Bc, L2)
        // whose eigenvalues are mu = 1 / omega
        const int m = nv_[b];
        std::vector<double> H(2 * (size_t)m * m), G(2 * (size_t)m * m);
        for (int i = 0; i < m; ++i)
          for (int j = 0; j < m; ++j) {
            const std::complex<double> x = Hm_[b][(size_t)i * mmax_ + j], y = Gm_[b][(size_t)i * mmax_ + j];
            H[2 * ((size_t)i * m + j)] = x.real();
            H[2 * ((size_t)i * m + j) + 1] = x.imag();
            G[2 * ((size_t)i * m + j)] = y.real();
            G[2 * ((size_t)i * m + j) + 1] = y.imag();
          }
        ok[b] = gen_eig_chol(m, H, G, lam[b], Sv[b]) ? 1 : 0;
        if (ok[b] && qep_) {  // the pencil's wanted pairs: largest |mu| = smallest |omega| first
          std::vector<int> idx(m);
          std::iota(idx.begin(), idx.end(), 0);
          std::stable_sort(idx.begin(), idx.end(),
                           [&](int x, int y) { return std::fabs(lam[b][x]) > std::fabs(lam[b][y]); });
          std::vector<double> l2(m), s2(Sv[b].size());
          for (int q = 0; q < m; ++q) {
            l2[q] = lam[b][idx[q]];
            for (int a = 0; a < m; ++a) {
              s2[2 * ((size_t)a * m + q)] = Sv[b][2 * ((size_t)a * m + idx[q])];
              s2[2 * ((size_t)a * m + q) + 1] = Sv[b][2 * ((size_t)a * m + idx[q]) + 1];
            }
          }
          lam[b].swap(l2);
          Sv[b].swap(s2);
        }
      });
    }
    for (int b = 0; b < B_; ++b)
      if (rr[b] && !ok[b]) {
        fail_block(b, act);
        rr[b] = 0;
      }
    // Ritz vectors of the nev lowest, their residuals
    std::vector<int> ncheck(B_, 0);
    int ncmax = 0;
    for (int b = 0; b < B_; ++b)
      if (rr[b]) {
        ncheck[b] = std::min(probs_[b]->nev, nv_[b]);
        ncmax = std::max(ncmax, ncheck[b]);
      }
    if (ncmax == 0) return;
    ritz_vectors(rr, Sv, 0, ncmax, &lam);
    std::vector<std::vector<double>> theta(B_);
    for (int b = 0; b < B_; ++b)
      if (rr[b]) theta[b].assign(lam[b].begin(), lam[b].begin() + ncheck[b]);
    std::vector<double> lamv((size_t)B_ * nkeep_, 0.0);
    for (int b = 0; b < B_; ++b)
      for (int c = 0; c < ncheck[b]; ++c) lamv[(size_t)b * nkeep_ + c] = qep_ ? 1.0 / theta[b][c] : theta[b][c];
    DevBuf dl = upload(pool(), lamv, stream()), dnc = ints(ncheck);
    DevBuf part(pool(), sizeof(double) * 3 * (size_t)B_ * nkeep_ * nchunk_);
    if (qep_)  // ||K x + omega D x - omega^2 M x|| / ||K x|| = ||(L2 z)_b - omega (Bc z)_b|| / ||(L2 z)_b||
      hipLaunchKernelGGL(k_residual_partial, dim3(nchunk_, B_, ncmax), dim3(256), 0, stream(),
                         Xb_.as<double2>() + n_, MXb_.as<double2>() + n_, KXb_.as<double2>() + n_, ldk_, n_, nvec_,
                         dl.as<double>(), nkeep_, dnc.as<int>(), part.as<double>(), nchunk_);
    else
      hipLaunchKernelGGL(k_residual_partial, dim3(nchunk_, B_, ncmax), dim3(256), 0, stream(), Xb_.as<double2>(),
                         KXb_.as<double2>(), MXb_.as<double2>(), ldk_, nvec_, nvec_, dl.as<double>(), nkeep_,
                         dnc.as<int>(), part.as<double>(), nchunk_);
    std::vector<double> hpart(3 * (size_t)B_ * nkeep_ * nchunk_);
    TM_HIP(hipMemcpyAsync(hpart.data(), part.p, sizeof(double) * hpart.size(), hipMemcpyDeviceToHost, stream()));
    TM_HIP(hipStreamSynchronize(stream()));
    std::vector<std::vector<double>> res(B_), den(B_);
    std::vector<int> need_floor(B_, 0);
    int nfmax = 0;
    for (int b = 0; b < B_; ++b) {
      if (!rr[b]) continue;
      double worst = 0;
      for (int c = 0; c < ncheck[b]; ++c) {
        double num = 0, dd = 0;
        for (int q = 0; q < nchunk_; ++q) {
          const double* pq = &hpart[(((size_t)b * nkeep_ + c) * nchunk_ + q) * 3];
          num += pq[0];
          dd += pq[1];
        }
        res[b].push_back(dd > 0 ? std::sqrt(num / dd) : std::sqrt(num));
        den[b].push_back(dd);
        worst = std::max(worst, res[b].back());
      }
      // the floor only matters once the residuals are near the tolerance
      if (worst > kBlockTol && worst < 1e-6) {
        need_floor[b] = ncheck[b];
        nfmax = std::max(nfmax, ncheck[b]);
      }
    }
    std::vector<std::vector<double>> floor(B_);
    if (nfmax > 0) {
      DevBuf fpart(pool(), sizeof(double) * (size_t)B_ * nkeep_ * nchunk_);
      for (int c0 = 0; c0 < nfmax; c0 += pb_) {
        std::vector<int> nc2(B_, 0);
        for (int b = 0; b < B_; ++b) nc2[b] = std::max(0, std::min(pb_, need_floor[b] - c0));
        DevBuf dnc2 = ints(nc2);
        const int cm = std::min(pb_, nfmax - c0);
        // |K||x|, |M||x| (and |D||x|) into the (consumed) W, MW (KW) buffers, pb
        // columns at a time; x is the b half for the Coriolis problem
        const double2* xs = Xb_.as<double2>() + (qep_ ? n_ : 0LL) + (long long)c0 * nvec_;
        const dim3 ga((unsigned)(((S_.nf + 3) / 4) * (long long)B_), (cm + 7) / 8);  // (row quad, problem) x columns
        hipLaunchKernelGGL(k_spmv_abs, ga, dim3(256), 0, stream(), S_.nf, M_.d_rowptr.as<int>(),
                           M_.d_colidx.as<int>(), kptr_.as<const double2*>(), mk_ptrs_.as<const double2*>(),
                           qep_ ? dptr_.as<const double2*>() : (const double2* const*)nullptr, xs, ldk_, nvec_,
                           dnc2.as<int>(), Wb_.as<double2>(), MWb_.as<double2>(),
                           qep_ ? KWb_.as<double2>() : (double2*)nullptr, ldw_,
                           qep_ || !sk_.p ? (const SplitKM*)nullptr : sk_.as<SplitKM>());
        hipLaunchKernelGGL(k_floor_partial, dim3(nchunk_, B_, cm), dim3(256), 0, stream(), Wb_.as<double2>(),
                           qep_ ? KWb_.as<double2>() : (const double2*)nullptr, MWb_.as<double2>(), ldw_, n_, nvec_,
                           dl.as<double>() + c0, nkeep_, dnc2.as<int>(), fpart.as<double>() + (size_t)c0 * nchunk_,
                           nchunk_);
      }
      std::vector<double> fl((size_t)B_ * nkeep_ * nchunk_);
      TM_HIP(hipMemcpyAsync(fl.data(), fpart.p, sizeof(double) * fl.size(), hipMemcpyDeviceToHost, stream()));
      TM_HIP(hipStreamSynchronize(stream()));
      for (int b = 0; b < B_; ++b)
        for (int c = 0; c < need_floor[b]; ++c) {
          double fs = 0;
          for (int q = 0; q < nchunk_; ++q) fs += fl[((size_t)b * nkeep_ + c) * nchunk_ + q];
          floor[b].push_back(den[b][c] > 0 ? 2.220446049250313e-16 * std::sqrt(fs / den[b][c]) : 0.0);
        }
    }
    // decisions
    std::vector<int> restart(B_, 0);
    for (int b = 0; b < B_; ++b) {
      if (!rr[b]) continue;
      HostProblem* hp = probs_[b];
      int first_bad = -1;
      for (int c = 0; c < ncheck[b]; ++c) {
        const double r = res[b][c];
        const double fl = c < (int)floor[b].size() ? floor[b][c] : 0.0;
        bool good = r <= std::max(kBlockTol, 4.0 * fl);
        // stagnated (no 2x gain since the previous check) at the rounding level of
        // the basis: accept; the extraction refines what is left above 1e-8 with
        // fresh triangular solves.  The Coriolis problem has no such refinement:
        // it accepts a stall only below kQepStallMax.
        if (!good && r <= (qep_ ? kQepStallMax : kBlockStallMax) && c < (int)prev_res_[b].size() &&
            r > 0.5 * prev_res_[b][c])
          good = true;
        if (!good && first_bad < 0) first_bad = c;
        if (trace_)
          std::fprintf(stderr, "[tmhip] k=%d rpm=%.0f warm=%d nv=%d rst=%d c=%d lam=%.10e res=%.2e floor=%.2e\n",
                       hp->k, hp->in.rpm, hp->warm_started ? 1 : 0, nv_[b], hp->restarts, c, theta[b][c], r, fl);
      }
      prev_res_[b] = res[b];
      const bool converged = first_bad < 0 && ncheck[b] >= hp->nev;
      if (converged || hp->restarts >= std::min(hp->max_restarts, kBlockMaxRestarts)) {
        hp->converged = converged;
        hp->done = true;
        hp->n_extract = ncheck[b];
        hp->theta_sorted_lambda = theta[b];
        act[b] = 0;
        continue;
      }
      if (nv_[b] + pb_ > mmax_) {
        restart[b] = 1;
        hp->restarts += 1;
        const int nx = std::min(nv_[b], nkeep_);
        y0_[b] = std::max(0, std::min(first_bad < 0 ? 0 : first_bad, nx - pb_));
      }
    }
    bool anyr = false;
    int ncm = 0;
    for (int b = 0; b < B_; ++b)
      if (restart[b]) {
        anyr = true;
        ncm = std::max(ncm, ncheck[b]);
      }
    if (!anyr) return;
    // restart: V <- the nkeep lowest Ritz vectors (columns ncheck.. formed now)
    ritz_vectors(restart, Sv, ncm, nkeep_);
    {
      DevBuf dr = ints(restart);
      KTimer t(ctx_, "vec", 0, 0);
      hipLaunchKernelGGL(k_copy_cols, dim3((nvec_ + 255) / 256, B_), dim3(256), 0, stream(), Xb_.as<double2>(), ldk_,
                         (const int*)nullptr, Vb_.as<double2>(), ldv_, (const int*)nullptr, nkeep_, nvec_, dr.as<int>());
      hipLaunchKernelGGL(k_copy_cols, dim3((nvec_ + 255) / 256, B_), dim3(256), 0, stream(), KXb_.as<double2>(), ldk_,
                         (const int*)nullptr, KVb_.as<double2>(), ldv_, (const int*)nullptr, nkeep_, nvec_, dr.as<int>());
      hipLaunchKernelGGL(k_copy_cols, dim3((nvec_ + 255) / 256, B_), dim3(256), 0, stream(), MXb_.as<double2>(), ldk_,
                         (const int*)nullptr, MVb_.as<double2>(), ldv_, (const int*)nullptr, nkeep_, nvec_, dr.as<int>());
    }
    // projected matrices of the kept block, recomputed: X^H K X and X^H M X
    std::vector<int> nk(B_, 0);
    for (int b = 0; b < B_; ++b)
      if (restart[b]) nk[b] = std::min(nv_[b], nkeep_);
    DevBuf dr = ints(restart), dnk = ints(nk);
    for (int c0 = 0; c0 < nkeep_; c0 += 16) {
      const int kb = std::min(16, nkeep_ - c0);
      auto hkm = tsgemm2(Vb_.as<double2>(), ldv_, dnk, nkeep_, KVb_.as<double2>() + (long long)c0 * nvec_,
                         MVb_.as<double2>() + (long long)c0 * nvec_, ldv_, kb, dr);
      const std::vector<double2>& hk = hkm.first;
      const std::vector<double2>& hm = hkm.second;
      for (int b = 0; b < B_; ++b)
        if (restart[b])
          for (int a = 0; a < nk[b]; ++a)
            for (int c = 0; c < kb && c0 + c < nk[b]; ++c) {
              const double2 x = hk[((size_t)b * mmax_ + a) * 16 + c], y = hm[((size_t)b * mmax_ + a) * 16 + c];
              Hm_[b][(size_t)a * mmax_ + c0 + c] = std::complex<double>(x.x, x.y);
              Gm_[b][(size_t)a * mmax_ + c0 + c] = std::complex<double>(y.x, y.y);
            }
    }
    for (int b = 0; b < B_; ++b)
      if (restart[b]) nv_[b] = nk[b];
  }

  void finish_failed(const std::vector<int>& sel) {
    for (int b = 0; b < B_; ++b)
      if (sel[b] && probs_[b] && probs_[b]->out) {
        probs_[b]->out->status = TM_ERR_NUMERIC;
        probs_[b]->out->n_modes = 0;
      }
  }
};

// The block method serves problems with at least this many free DOFs (its basis
// holds up to 64 vectors); smaller ones, and the reference convergence policy,
// run the single-vector Lanczos.
static constexpr int kBlockMinFree = 256;

static bool is_qep(const Symbolic& S, const tm_problem& p) {
  HostProblem hp;
  hp.in = p;
  fill_policy(S, hp);
  return hp.qep;
}

// the Coriolis pencil always runs the block method (2n >= 128 vectors' room)
static constexpr int kQepMinFree = 64;

static bool use_block(const Symbolic& S, const tm_problem& p) {
  HostProblem hp;
  hp.in = p;
  fill_policy(S, hp);
  if (hp.qep) {
    if (S.n_free_dofs < kQepMinFree) throw ArgError("the Coriolis QEP needs at least 64 free DOFs");
    // a basis longer than the default one needs vectors (2 n long) of twice its length
    if (hp.nev > kBlockMaxNev || (hp.nev > 24 && S.n_free_dofs < basis_length(16, true, hp.nev)))
      throw ArgError("the Coriolis QEP takes at most 110 modes per harmonic, and 31 or more modes need at least "
                     "as many free DOFs as basis vectors");
    return true;
  }
  if (p.method != TM_METHOD_BLOCK || p.conv_policy == TM_CONV_REFERENCE || S.n_free_dofs < kBlockMinFree)
    return false;
  // the basis must stay well inside the space; longer requests run the Lanczos method
  return hp.nev <= kBlockMaxNev && S.n_free_dofs >= 2 * basis_length(block_width(hp.nev, false), false, hp.nev);
}

static int64_t problem_footprint(const Operator& op, const tm_problem& p) {
  const Symbolic& S = op.mesh->S;
  const Plan& P = op.mesh->plan;
  HostProblem hp;
  hp.in = p;
  fill_policy(S, hp);
  const int64_t n = 3LL * S.nf;
  const bool block = use_block(S, hp.in);
  const int pb = block_width(hp.nev, hp.qep);
  const int pcols = std::min(16, std::max(block ? pb : 1, hp.nev));
  // factor entry: complex FP32 or real FP64 8 bytes, complex FP64 16
  const bool real = block && !hp.qep && p.precision != TM_PREC_MIXED && real_factor_enabled() &&
                    real_harmonic(S.N, hp.k);
  const int64_t es = p.precision == TM_PREC_MIXED || real ? 8 : 16;
  int64_t bytes = es * (P.panel_entries + P.upool_entries + P.inv_scratch_entries) + 16 * pcols * P.tvec_entries;
  if (block) {
    const int mmax = basis_length(pb, hp.qep, hp.nev);
    const int64_t nv = hp.qep ? 2 * n : n;           // pencil vectors are 2n long
    bytes += 3 * 16 * (int64_t)mmax * nv;            // V, MV, KV
    bytes += 3 * 16 * (int64_t)pb * nv;              // W, MW, KW
    bytes += 4 * 16 * (int64_t)(mmax / 2) * nv;      // Ritz X, KX, MX, refinement
    if (hp.qep) bytes += 16 * 9 * (int64_t)op.nblk;  // D(k)
  } else {
    bytes += 2 * 16 * (int64_t)(hp.ncv + 1) * n;     // V, MV
    bytes += 2 * 16 * n;                             // W, MW
    bytes += 4 * 16 * (int64_t)(hp.nev + 1) * n;     // restart / Ritz / residual vectors
  }
  bytes += 2 * 16 * 9 * (int64_t)op.nblk;            // K(k), M(k)
  return bytes;
}

int64_t problem_footprint(const Operator& op, int num_modes, int harmonic, int method, int precision) {
  tm_problem p{};
  p.harmonic = harmonic;
  p.num_modes = num_modes;
  p.method = method;
  p.precision = precision;
  return problem_footprint(op, p);
}

// Releases what a context holds besides its pool's live buffers: waits for its
// stream, drains its deliverer, frees the cached pool blocks, events and stream.
static void release_context(Context& c) {
  (void)hipSetDevice(c.device);
  (void)hipStreamSynchronize(c.stream);
  c.groups.clear();
  c.deliver.reset();  // drains its copies
  c.clear_graphs();
  c.pool.trim();
  for (auto& p : c.pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  c.pending.clear();
  for (auto ev : c.event_pool) (void)hipEventDestroy(ev);
  c.event_pool.clear();
  c.release_side();
  (void)hipStreamDestroy(c.stream);
  c.stream = nullptr;
}

// a context of its own (stream, pool) on the parent's device, for one slot group
struct SubContext {
  Context c;
  explicit SubContext(int device) {
    c.device = device;
    TM_HIP(hipSetDevice(device));
    TM_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  }
  ~SubContext() { release_context(c); }
};

// The context of concurrent slot group g of ctx, created on first use.
static Context* group_context(Context* ctx, size_t g) {
  while (ctx->groups.size() <= g) {
    std::unique_ptr<Context> c(new Context);
    c->device = ctx->device;
    TM_HIP(hipSetDevice(ctx->device));
    TM_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    ctx->groups.emplace_back(c.release(), [](Context* p) {
      release_context(*p);
      delete p;
    });
  }
  Context* c = ctx->groups[g].get();
  c->profiling = ctx->profiling;
  return c;
}

// Folds a group context's timings into its parent's.
static void merge_stats(Context* ctx, Context* g) {
  g->flush_timing();
  for (const auto& kv : g->stats) {
    KernelStat& d = ctx->stats[kv.first];
    d.launches += kv.second.launches;
    d.total_ms += kv.second.total_ms;
    d.bytes += kv.second.bytes;
    d.flops += kv.second.flops;
  }
  g->stats.clear();
}

// Runs of one mesh's block-method problems (one block size each) as slot groups
// side by side, each on its own host thread and context, so that one group's
// host work (Gram reductions, Rayleigh-Ritz, extraction) and small launches
// overlap the other's kernels; the slot capacities share the memory budget in
// proportion to the runs' sizes.  One run (or TMHIP_SERIAL_GROUPS=1): one after
// another on the caller's context.
static void run_block_groups(Context* ctx, Operator& op, std::vector<std::vector<HostProblem*>>& runs,
                             bool qep, const std::vector<FactorElem>& elems) {
  // device bytes per problem of each run: the largest footprint of its problems
  std::vector<int64_t> per(runs.size(), 1);
  for (size_t g = 0; g < runs.size(); ++g)
    for (const HostProblem* hp : runs[g]) per[g] = std::max<int64_t>(per[g], problem_footprint(op, hp->in));
  // a warm-started run of at least 3 x 32 problems takes at most a third of them per wave,
  // so that its later waves find adjacent rpm points solved.  Smaller runs stay in one
  // wave: a sweep launch over fewer than about 32 problems takes nearly as long as over
  // more (the upper tree levels are latency-bound), so three waves would cost more sweeps
  // than continuation saves (measured on the 40 real-harmonic problems of the bladed
  // sector's 20-point sweep)
  auto wave_cap = [&](const std::vector<HostProblem*>& q, int64_t cap) {
    if (!qep && q.size() >= 96 &&
        std::any_of(q.begin(), q.end(), [](const HostProblem* h) { return h->in.warm_start != 0; }))
      cap = std::min<int64_t>(cap, std::max<int64_t>(1, ((int64_t)q.size() + 2) / 3));
    return (int)cap;
  };
  const char* ser = std::getenv("TMHIP_SERIAL_GROUPS");
  const bool serial = runs.size() == 1 || (ser && std::atoi(ser) != 0);
  size_t free_b = 0, total_b = 0;
  if (serial) {
    struct Fork {  // one group at a time: the factorization may fork onto the side stream
      Context* c;
      bool was;
      explicit Fork(Context* x) : c(x), was(x->fork) { c->fork = true; }
      ~Fork() { c->fork = was; }
    } fork(ctx);
    for (auto& g : ctx->groups) g->pool.trim();  // cached there, of no use here
    for (size_t g = 0; g < runs.size(); ++g) {
      auto& q = runs[g];
      TM_HIP(hipMemGetInfo(&free_b, &total_b));
      // blocks cached by the context's pool are reusable by this call
      const int64_t budget = (int64_t)(ctx->batch_fraction * (double)(free_b + ctx->pool.cached()));
      const int cap = wave_cap(q, std::max<int64_t>(1, std::min<int64_t>({64, (int64_t)q.size(), budget / per[g]})));
      BatchLanczos bl(ctx, op, {}, true);
      bl.run_slots(q, cap, cap, qep, elems[g]);
    }
    return;
  }
  ctx->pool.trim();
  // the memory budget is shared in proportion to the runs' total footprints
  double total = 0;
  int64_t cached = 0;
  for (size_t g = 0; g < runs.size(); ++g) total += (double)runs[g].size() * (double)per[g];
  std::vector<Context*> gc(runs.size());
  for (size_t g = 0; g < runs.size(); ++g) {
    gc[g] = group_context(ctx, g);
    cached += (int64_t)gc[g]->pool.cached();
  }
  TM_HIP(hipMemGetInfo(&free_b, &total_b));
  const double budget = ctx->batch_fraction * (double)((int64_t)free_b + cached);
  std::vector<std::exception_ptr> err(runs.size());
  std::vector<std::thread> th;
  for (size_t g = 0; g < runs.size(); ++g)
    th.emplace_back([&, g] {
      Context* c = gc[g];
      try {
        TM_HIP(hipSetDevice(c->device));
        const int64_t share = (int64_t)(budget * (double)runs[g].size() * (double)per[g] / total);
        const int cap =
            wave_cap(runs[g], std::max<int64_t>(1, std::min<int64_t>({64, (int64_t)runs[g].size(), share / per[g]})));
        {
          BatchLanczos bl(c, op, {}, true);
          bl.run_slots(runs[g], cap, cap, qep, elems[g]);
        }
        if (c->deliver) c->deliver->drain();
      } catch (...) {
        err[g] = std::current_exception();
        if (c->deliver) c->deliver->drain_quiet();
      }
    });
  for (auto& t : th) t.join();
  for (Context* c : gc) merge_stats(ctx, c);
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
}

static std::vector<int> solve_problems_impl(Context* ctx, const tm_problem* problems, int n, tm_result* results);

// Failed problems are solved again, each alone, as the reference retries its failed
// harmonics one at a time (src/cyclic_solver.cpp:1334-1342).  Where the factorization of
// a failed problem met a zero, non-finite or near-singular pivot (the reference falls
// back from LDL^H to LU on a zero pivot, src/hermitian_lanczos.cpp:27-55), the retry moves
// the shift down by 1 % of max(|sigma|, 1), up to three times: the block method's Ritz values come from
// K(k) and M(k) and do not depend on the shift, and the Lanczos method maps its Ritz
// values back with the shift it used (tm_result.shift_used).  Other failures get one
// retry.  The Coriolis pencil factors K_eff(k) itself, with no shift to move: a flagged
// factor there stays failed.
static void retry_failed(Context* ctx, const tm_problem* problems, int n, tm_result* results,
                         std::vector<int> flags) {
  std::vector<tm_problem> cur(problems, problems + n);
  for (int attempt = 1; attempt <= 3; ++attempt)
    for (int i = 0; i < n; ++i) {
      if (results[i].status == TM_OK) continue;
      const Operator& op = *static_cast<const Operator*>(cur[i].op);
      if (flags[i] && is_qep(op.mesh->S, cur[i])) continue;
      if (!flags[i] && attempt > 1) continue;
      if (flags[i]) cur[i].shift -= 0.01 * std::max(std::fabs(cur[i].shift), 1.0);
      flags[i] = solve_problems_impl(ctx, &cur[i], 1, &results[i])[0];
      results[i].factor_retries = attempt;
      results[i].pivot_flags = flags[i];
      if (results[i].status == TM_OK)
        results[i].shift_used = is_qep(op.mesh->S, cur[i]) ? 0.0 : cur[i].shift;
    }
}

// Mode shapes reach the caller's buffers through the context's deliverer; every
// copy has landed when this returns, also when it throws.
void solve_problems(Context* ctx, const tm_problem* problems, int n, tm_result* results) {
  try {
    std::vector<int> flags = solve_problems_impl(ctx, problems, n, results);
    if (results) {
      for (int i = 0; i < n; ++i) {
        const Operator& op = *static_cast<const Operator*>(problems[i].op);
        results[i].factor_retries = 0;
        results[i].pivot_flags = flags[i];
        results[i].shift_used = is_qep(op.mesh->S, problems[i]) ? 0.0 : problems[i].shift;
      }
      if (ctx->deliver) ctx->deliver->drain();
      retry_failed(ctx, problems, n, results, flags);
    }
  } catch (...) {
    if (ctx->deliver) ctx->deliver->drain_quiet();
    throw;
  }
  if (ctx->deliver) ctx->deliver->drain();
}

static std::vector<int> solve_problems_impl(Context* ctx, const tm_problem* problems, int n, tm_result* results) {
  std::vector<HostProblem> hps(n);
  // batches: problems of one operator and one method, ordered by nev so a
  // batch shares its block size and basis length
  // key: (mesh, 1) for the block method, whose slots may mix materials of one mesh;
  // (operator, 0) for Lanczos batches
  std::map<std::pair<const void*, int>, std::vector<HostProblem*>> groups;
  std::map<std::pair<const void*, int>, Operator*> rep;
  std::vector<std::pair<const void*, int>> order;
  for (int i = 0; i < n; ++i) {
    hps[i].in = problems[i];
    hps[i].out = results ? &results[i] : nullptr;
    if (!problems[i].op) throw ArgError("problem without operator");
    if (problems[i].num_modes <= 0) throw ArgError("num_modes must be positive");
    if (problems[i].method != TM_METHOD_LANCZOS && problems[i].method != TM_METHOD_BLOCK)
      throw ArgError("unknown method");
    Operator* op = static_cast<Operator*>(problems[i].op);
    const int max_k = op->mesh->S.N / 2;
    if (problems[i].harmonic < 0 || problems[i].harmonic > max_k)
      throw ArgError("harmonic index out of range 0..N/2");
    const bool blk = use_block(op->mesh->S, problems[i]);
    const bool qep = blk && is_qep(op->mesh->S, problems[i]);
    const bool mixed = problems[i].precision == TM_PREC_MIXED;
    if (mixed && (!blk || qep))
      throw ArgError("TM_PREC_MIXED runs the block method on the (K, M) pencil only: not the Lanczos method, the "
                     "reference convergence policy, the Coriolis pencil or systems below 256 free DOFs");
    // key: 0 Lanczos (per operator); per mesh: 1 block method, 2 Coriolis pencil, 3 block method with
    // the complex-FP32 factor
    const auto key = std::make_pair(blk ? (const void*)op->mesh : (const void*)op, blk ? (qep ? 2 : (mixed ? 3 : 1)) : 0);
    if (!groups.count(key)) {
      order.push_back(key);
      rep[key] = op;
    }
    groups[key].push_back(&hps[i]);
  }
  for (const auto& key : order) {
    Operator* op = rep[key];
    auto& list = groups[key];
    int64_t per = 0;
    for (auto* hp : list) {
      fill_policy(op->mesh->S, *hp);
      per = std::max<int64_t>(per, problem_footprint(*op, hp->in));
    }
    std::stable_sort(list.begin(), list.end(), [](HostProblem* a, HostProblem* b) { return a->nev < b->nev; });
    if (key.second >= 1) {
      // block method: one run per block size (the Coriolis group is one run)
      const bool qep = key.second == 2;
      // a real harmonic on the (K, M) pencil in FP64 takes the real factor
      // (TMHIP_REAL_FACTOR=0: the complex one)
      const auto real = [&](const HostProblem* h) {
        return key.second == 1 && real_factor_enabled() && real_harmonic(op->mesh->S.N, h->k);
      };
      std::stable_partition(list.begin(), list.end(), [&](const HostProblem* h) { return !real(h); });
      std::vector<std::vector<HostProblem*>> runs;
      for (size_t i = 0; i < list.size();) {
        size_t j = i;
        const int pb_i = block_width(list[i]->nev, qep);
        const int mm_i = basis_length(pb_i, qep, list[i]->nev);
        // one block size, basis length and factor type per run: a problem's arithmetic
        // does not depend on what shares its run
        while (j < list.size() && block_width(list[j]->nev, qep) == pb_i &&
               basis_length(pb_i, qep, list[j]->nev) == mm_i && real(list[j]) == real(list[i]))
          ++j;
        runs.emplace_back(list.begin() + i, list.begin() + j);
        i = j;
      }
      // rpm continuation (tm_problem.warm_start): a run's problems in the order of a third of
      // its rpm points (every third one), then the next third, then the rest, harmonic by
      // harmonic within each point, so that most problems find an adjacent rpm point solved in
      // an earlier wave of their slot group
      const auto warm_run = [&](const std::vector<HostProblem*>& r) {
        return !qep && std::any_of(r.begin(), r.end(), [](const HostProblem* h) { return h->in.warm_start != 0; });
      };
      for (auto& r : runs) {
        if (!warm_run(r)) continue;
        std::vector<double> pts;
        for (auto* h : r) pts.push_back(h->in.rpm);
        std::sort(pts.begin(), pts.end());
        pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
        auto idx = [&](const HostProblem* h) { return std::lower_bound(pts.begin(), pts.end(), h->in.rpm) - pts.begin(); };
        std::stable_sort(r.begin(), r.end(), [&](const HostProblem* x, const HostProblem* y) {
          const long ix = idx(x), iy = idx(y);
          if (ix % 3 != iy % 3) return ix % 3 < iy % 3;
          if (ix != iy) return ix < iy;
          return x->k < y->k;
        });
      }
      // while fewer than three groups: the largest run (16 problems or more) in two
      // interleaved halves, so that groups overlap one another's host work (a warm run by
      // harmonics, alternately, so that each harmonic's rpm points stay in one group)
      while (runs.size() < 3) {
        size_t big = 0;
        for (size_t g = 1; g < runs.size(); ++g)
          if (runs[g].size() > runs[big].size()) big = g;
        if (runs[big].size() < 16) break;
        std::vector<HostProblem*> a, b;
        if (warm_run(runs[big])) {
          std::vector<int> ks;
          for (auto* h : runs[big]) ks.push_back(h->k);
          std::sort(ks.begin(), ks.end());
          ks.erase(std::unique(ks.begin(), ks.end()), ks.end());
          if (ks.size() < 2) break;
          for (auto* h : runs[big])
            ((std::lower_bound(ks.begin(), ks.end(), h->k) - ks.begin()) % 2 ? b : a).push_back(h);
        } else {
          for (size_t i = 0; i < runs[big].size(); ++i) (i % 2 ? b : a).push_back(runs[big][i]);
        }
        runs[big].swap(a);
        runs.insert(runs.begin() + big + 1, std::move(b));
      }
      std::vector<FactorElem> elems;
      for (const auto& r : runs)
        elems.push_back(key.second == 3 ? FactorElem::C32 : (real(r[0]) ? FactorElem::R64 : FactorElem::C64));
      run_block_groups(ctx, *op, runs, qep, elems);
      continue;
    }
    size_t free_b = 0, total_b = 0;
    TM_HIP(hipMemGetInfo(&free_b, &total_b));
    // blocks cached by the context's pool are reusable by this call
    int64_t budget = (int64_t)(ctx->batch_fraction * (double)(free_b + ctx->pool.cached()));
    size_t max_batch = (size_t)std::max<int64_t>(1, std::min<int64_t>(64, budget / std::max<int64_t>(per, 1)));
    for (size_t i = 0; i < list.size();) {  // Lanczos: lockstep batches
      size_t j = i;
      while (j < list.size() && j - i < max_batch) ++j;
      std::vector<HostProblem*> chunk(list.begin() + i, list.begin() + j);
      BatchLanczos bl(ctx, *op, chunk, false);
      bl.run();
      i = j;
    }
  }
  std::vector<int> flags(n);
  for (int i = 0; i < n; ++i) flags[i] = hps[i].factor_flags;
  return flags;
}

}  // namespace tme

// ---------------------------------------------------------------------------
// Continuous-batching server.  A dispatcher thread takes submitted problems and
// hands them to slot groups, one per (mesh, block size): a group may mix
// materials and rpm of one mesh.  Each group runs on its own host thread and
// HIP stream with its own device pool, so the groups' kernels overlap on the
// device; its slot capacity comes from the problems' device footprint within
// a server-wide HBM budget (a group holds at most a quarter of it).  A group
// admits queued problems as slots free up (factorizations batched, masked to
// the new slots), steps all its slots together, retires finished problems into
// the caller's result buffers, and returns its memory when it has been dry for
// kIdleKeep or another group waits for memory.
// Problems too small for the block method (and the Lanczos method's requests) run
// as Lanczos batches on a worker thread of their own, with its own stream and pool,
// so they never hold up the dispatcher's admissions.
// ---------------------------------------------------------------------------
namespace tme {

struct Server {
  struct Req {
    HostProblem hp;
    bool done = false;
  };
  struct Group {
    Server* srv = nullptr;
    const Mesh* mesh = nullptr;
    Operator* rep = nullptr;  // any operator of the mesh (plan, pattern)
    int pbclass = 8;
    int nevclass = 0;  // a wanted-pair count of its basis class (any one: they share the basis length)
    bool qep = false;
    std::unique_ptr<SubContext> sub;
    std::unique_ptr<SubContext> side;  // stream of the admissions' factorizations
    std::unique_ptr<BatchLanczos> bl;
    std::deque<HostProblem*> queue;  // guarded by srv->mu
    double reserved = 0;             // guarded by srv->mu
    std::thread th;
    bool quit = false;               // guarded by srv->mu
    // counters (TMHIP_SERVER_STATS=1 prints them when the server stops)
    long admits = 0, admitted = 0, steps = 0, slot_steps = 0, setups = 0;
    double t_admit = 0, t_step = 0, t_setup = 0;
    double admit_s[5] = {0, 0, 0, 0, 0};  // the admissions' parts (BatchLanczos::admit_s_)
  };

  // how long an idle group keeps its slots before it returns their memory
  static constexpr std::chrono::milliseconds kIdleKeep{1000};
  static constexpr int kMinSlots = 16;  // a group's smallest capacity (below max_slots)
  Context* ctx;
  int max_slots;
  double frac, budget = 0, reserved = 0;
  int mem_waiters = 0;  // groups waiting for device memory (guarded by mu)
  std::thread th;
  std::mutex mu;
  std::condition_variable cv_in, cv_done, cv_group;
  bool stop = false;
  std::string fatal;
  std::deque<uint64_t> incoming;
  std::map<uint64_t, std::unique_ptr<Req>> reqs;
  std::map<HostProblem*, uint64_t> ticket_of;
  uint64_t next_ticket = 1;
  std::map<std::pair<const Mesh*, int>, std::unique_ptr<Group>> groups;  // (mesh, block size + 1000 qep + basis)
  // Lanczos-method requests: batches per operator for the worker thread (guarded by mu)
  std::deque<std::pair<Operator*, std::vector<HostProblem*>>> lanczos_q;
  bool lanczos_quit = false;  // guarded by mu
  std::thread lanczos_th;

  Server(Context* c, int slots, double f) : ctx(c), max_slots(std::max(1, std::min(64, slots))), frac(f) {
    if (!(frac > 0.0 && frac <= 0.95)) frac = 0.85;
    size_t free_b = 0, total_b = 0;
    TM_HIP(hipMemGetInfo(&free_b, &total_b));
    budget = frac * (double)(free_b + ctx->pool.cached());
    th = std::thread([this] { dispatch(); });
  }
  ~Server() {
    {
      std::lock_guard<std::mutex> g(mu);
      stop = true;
    }
    cv_in.notify_all();
    if (th.joinable()) th.join();
  }

  uint64_t submit(const tm_problem& p, tm_result* out) {
    if (!p.op) throw ArgError("problem without operator");
    if (p.num_modes <= 0) throw ArgError("num_modes must be positive");
    if (p.method != TM_METHOD_LANCZOS && p.method != TM_METHOD_BLOCK) throw ArgError("unknown method");
    Operator* op = static_cast<Operator*>(p.op);
    if (p.harmonic < 0 || p.harmonic > op->mesh->S.N / 2) throw ArgError("harmonic index out of range 0..N/2");
    if (p.precision != TM_PREC_FP64) throw ArgError("precision: the server runs TM_PREC_FP64 problems");
    auto r = std::make_unique<Req>();
    r->hp.in = p;
    r->hp.out = out;
    if (out) {
      out->n_modes = 0;
      out->n_unique = 0;
      out->converged = 0;
      out->restarts = 0;
      out->status = TM_ERR_RUNTIME;
      out->seconds = 0;
      out->factor_retries = 0;  // the server drops a failed problem without a retry
      out->pivot_flags = 0;
      out->shift_used = is_qep(op->mesh->S, p) ? 0.0 : p.shift;
    }
    std::lock_guard<std::mutex> g(mu);
    if (!fatal.empty()) throw HipError("server stopped: " + fatal);
    const uint64_t t = next_ticket++;
    ticket_of[&r->hp] = t;
    reqs[t] = std::move(r);
    incoming.push_back(t);
    cv_in.notify_all();
    return t;
  }

  bool wait(uint64_t t, double timeout_s) {
    std::unique_lock<std::mutex> lk(mu);
    if (!reqs.count(t)) throw ArgError("unknown ticket");
    auto pred = [&] { return reqs[t]->done || !fatal.empty(); };
    if (timeout_s < 0) cv_done.wait(lk, pred);
    else if (!cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) return false;
    if (!reqs[t]->done) throw HipError("server stopped: " + fatal);
    ticket_of.erase(&reqs[t]->hp);
    reqs.erase(t);
    return true;
  }

  void finish(const std::vector<HostProblem*>& fin) {
    std::lock_guard<std::mutex> g(mu);
    for (auto* hp : fin) {
      auto it = ticket_of.find(hp);
      if (it != ticket_of.end()) reqs[it->second]->done = true;
    }
    cv_done.notify_all();
  }

  void fail(const std::string& msg) {
    std::lock_guard<std::mutex> g(mu);
    if (fatal.empty()) fatal = msg;
    for (auto& kv : reqs)
      if (!kv.second->done && kv.second->hp.out) kv.second->hp.out->status = TM_ERR_HIP;
    cv_done.notify_all();
    cv_group.notify_all();
  }

  // one group's thread: admit, step, retire until the server stops
  void run_group(Group* g) {
    try {
      TM_HIP(hipSetDevice(ctx->device));
      while (true) {
        std::vector<HostProblem*> adm;
        bool need_setup = false;
        int64_t per = 0;
        {
          std::unique_lock<std::mutex> lk(mu);
          const bool busy = g->bl && g->bl->busy();
          if (!busy && g->queue.empty()) {
            if (g->bl && !g->quit && mem_waiters == 0) {
              // ran dry: keep the slots (their workspaces, captured launches) for a while,
              // as the next request of a stream usually comes soon, unless another group
              // waits for memory
              cv_group.wait_for(lk, kIdleKeep, [&] {
                return g->quit || !g->queue.empty() || !fatal.empty() || mem_waiters > 0;
              });
              if (!fatal.empty()) return;
              if (!g->queue.empty()) continue;
            }
            if (g->bl) {  // still dry: give the memory back
              lk.unlock();
              for (int q = 0; q < 5; ++q) g->admit_s[q] += g->bl->admit_s_[q];
              g->bl.reset();
              g->sub->c.pool.trim();
              lk.lock();
              reserved -= g->reserved;
              g->reserved = 0;
              cv_group.notify_all();
            }
            cv_group.wait(lk, [&] { return g->quit || !g->queue.empty() || !fatal.empty(); });
            if (!fatal.empty()) return;
            if (g->quit && g->queue.empty()) return;
            continue;
          }
          if (!g->bl) {
            for (const HostProblem* h : g->queue)
              per = std::max<int64_t>(per, problem_footprint(*g->rep, h->in));
            per = std::max<int64_t>(per, 1) * 21 / 20;
            const double avail = std::min(budget - reserved, 0.25 * budget);
            // sized to the demand: twice the queue (at least kMinSlots), up to max_slots and
            // the memory share.  Workspaces of a full 64 slots run to tens of GB for the
            // bladed sector; allocating them whenever a group of a thin stream starts (and
            // returning them when it runs dry) stalled the stream for seconds
            int want = kMinSlots;
            while (want < 2 * (int)g->queue.size()) want *= 2;
            const int cap = (int)std::min<double>({(double)max_slots, (double)want, std::floor(avail / (double)per)});
            if (cap < 1) {  // no memory now: wait for another group to return some
              ++mem_waiters;
              cv_group.notify_all();  // idle groups release theirs
              cv_group.wait_for(lk, std::chrono::milliseconds(20));
              --mem_waiters;
              continue;
            }
            g->reserved = (double)cap * per;
            reserved += g->reserved;
            need_setup = true;
            per = cap;  // capacity
          }
        }
        if (need_setup) {
          const auto ts = std::chrono::steady_clock::now();
          g->bl = std::make_unique<BatchLanczos>(&g->sub->c, *g->rep, std::vector<HostProblem*>{}, true);
          g->bl->no_tuning();
          g->bl->slot_setup((int)per, g->pbclass == 8 ? kNarrowNev : std::max(24, g->nevclass), g->qep);
          g->setups += 1;
          g->t_setup += std::chrono::duration<double>(std::chrono::steady_clock::now() - ts).count();
        }
        const auto ta = std::chrono::steady_clock::now();
        // a finished admission's slots join the stepping ones; with nothing else to
        // step, wait for it
        if (g->bl->admitting() && (g->bl->admit_ready() || !g->bl->stepping())) g->bl->admit_finish();
        if (!g->bl->admitting()) {
          std::lock_guard<std::mutex> lk(mu);
          const int nfree = g->bl->free_slots();
          const int refill = std::max(1, g->bl->capacity() / 8);
          if (!g->queue.empty() && nfree > 0 &&
              (nfree >= refill || !g->bl->busy() || g->queue.size() <= (size_t)nfree))
            while (!g->queue.empty() && (int)adm.size() < nfree) {
              adm.push_back(g->queue.front());
              g->queue.pop_front();
            }
        }
        if (!adm.empty()) {
          // the new slots' factorization runs beside the steps of those in flight
          g->bl->admit_async(adm, &g->side->c);
          g->admits += 1;
          g->admitted += (long)adm.size();
        }
        const auto tb = std::chrono::steady_clock::now();
        g->t_admit += std::chrono::duration<double>(tb - ta).count();
        if (g->bl->stepping()) {
          g->steps += 1;
          g->slot_steps += g->bl->capacity() - g->bl->free_slots();
          g->bl->block_step();
          g->t_step += std::chrono::duration<double>(std::chrono::steady_clock::now() - tb).count();
        }
        // finished problems, and admitted ones whose factorization failed
        std::vector<HostProblem*> fin = g->bl->retire();
        if (!fin.empty())  // done once their mode shapes have landed
          g->sub->c.deliverer().then(g->sub->c.stream, [this, fin](const std::string& e) {
            if (e.empty()) finish(fin);
            else fail("result copy: " + e);
          });
      }
    } catch (const std::exception& e) {
      fail(e.what());
    }
  }

  // Lanczos batches, one operator at a time (requests queued meanwhile for the same
  // operator join the next batch), on a context of the worker's own
  void run_lanczos() {
    try {
      SubContext sc(ctx->device);
      while (true) {
        std::map<Operator*, std::vector<HostProblem*>> batch;
        {
          std::unique_lock<std::mutex> lk(mu);
          cv_group.wait(lk, [&] { return lanczos_quit || !lanczos_q.empty() || !fatal.empty(); });
          if (!fatal.empty()) return;
          if (lanczos_q.empty()) return;  // quit with nothing left
          while (!lanczos_q.empty()) {
            auto& f = lanczos_q.front();
            auto& v = batch[f.first];
            v.insert(v.end(), f.second.begin(), f.second.end());
            lanczos_q.pop_front();
          }
        }
        for (auto& kv : batch) {
          for (size_t i = 0; i < kv.second.size(); i += 64) {  // at most 64 problems per batch
            std::vector<HostProblem*> chunk(kv.second.begin() + i,
                                            kv.second.begin() + std::min(kv.second.size(), i + 64));
            BatchLanczos bl(&sc.c, *kv.first, chunk, false);
            bl.run();
            sc.c.deliverer().drain();
            finish(chunk);
          }
        }
      }
    } catch (const std::exception& e) {
      fail(e.what());
    }
  }

  void dispatch() {
    try {
      TM_HIP(hipSetDevice(ctx->device));
      while (true) {
        std::vector<uint64_t> fresh;
        {
          std::unique_lock<std::mutex> lk(mu);
          cv_in.wait(lk, [&] { return stop || !incoming.empty(); });
          if (stop && incoming.empty()) break;
          fresh.assign(incoming.begin(), incoming.end());
          incoming.clear();
        }
        std::map<Operator*, std::vector<HostProblem*>> small;
        for (uint64_t t : fresh) {
          HostProblem* hp;
          {
            std::lock_guard<std::mutex> g(mu);
            hp = &reqs[t]->hp;
          }
          Operator* op = static_cast<Operator*>(hp->in.op);
          fill_policy(op->mesh->S, *hp);
          if (!use_block(op->mesh->S, hp->in)) {
            small[op].push_back(hp);
            continue;
          }
          const int pbc = block_width(hp->nev, hp->qep);
          const int mmc = basis_length(pbc, hp->qep, hp->nev);
          std::lock_guard<std::mutex> g(mu);
          auto& gp = groups[{op->mesh, pbc + (hp->qep ? 1000 : 0) + 10000 * mmc}];
          if (!gp) {
            gp = std::make_unique<Group>();
            gp->srv = this;
            gp->mesh = op->mesh;
            gp->rep = op;
            gp->pbclass = pbc;
            gp->nevclass = hp->nev;
            gp->qep = hp->qep;
            gp->sub = std::make_unique<SubContext>(ctx->device);
            gp->side = std::make_unique<SubContext>(ctx->device);
            Group* raw = gp.get();
            gp->th = std::thread([this, raw] { run_group(raw); });
          }
          gp->queue.push_back(hp);
          cv_group.notify_all();
        }
        if (!small.empty()) {
          std::lock_guard<std::mutex> g(mu);
          if (!lanczos_th.joinable()) lanczos_th = std::thread([this] { run_lanczos(); });
          for (auto& kv : small) lanczos_q.emplace_back(kv.first, std::move(kv.second));
          cv_group.notify_all();
        }
      }
    } catch (const std::exception& e) {
      fail(e.what());
    }
    // drain: every group and the Lanczos worker finish their queues, then stop
    {
      std::lock_guard<std::mutex> g(mu);
      for (auto& kv : groups) kv.second->quit = true;
      lanczos_quit = true;
      cv_group.notify_all();
    }
    for (auto& kv : groups)
      if (kv.second->th.joinable()) kv.second->th.join();
    if (lanczos_th.joinable()) lanczos_th.join();
    if (std::getenv("TMHIP_SERVER_STATS"))
      for (auto& kv : groups) {
        const Group& g = *kv.second;
        std::fprintf(stderr,
                     "[tmhip] group nf=%d pb=%d qep=%d: %ld admissions of %ld problems (%.1f s: %.2f / %.2f / "
                     "%.2f / %.2f / %.2f s), %ld steps (%.1f s, %.1f slots busy on average), %ld slot setups "
                     "(%.1f s)\n",
                     g.mesh->S.nf, g.pbclass, g.qep ? 1 : 0, g.admits, g.admitted, g.t_admit, g.admit_s[0],
                     g.admit_s[1], g.admit_s[2], g.admit_s[3], g.admit_s[4], g.steps, g.t_step,
                     g.steps ? (double)g.slot_steps / g.steps : 0.0, g.setups, g.t_setup);
      }
    groups.clear();
  }
};

Server* server_create(Context* ctx, int max_slots, double hbm_fraction) {
  return new Server(ctx, max_slots, hbm_fraction);
}
void server_destroy(Server* s) { delete s; }
uint64_t server_submit(Server* s, const tm_problem& p, tm_result* out) { return s->submit(p, out); }
bool server_wait(Server* s, uint64_t ticket, double timeout_s) { return s->wait(ticket, timeout_s); }

}  // namespace tme
