// extern "C" entry points of include/tmhip.h.  Every call catches all C++
// exceptions and turns them into a status code plus tm_last_error(); the
// error classes keep the reference's distinction between invalid arguments
// (std::invalid_argument, src/material.cpp:15-22) and topology/runtime
// failures (std::runtime_error, src/mesh.cpp:239-243,
// src/cyclic_solver.cpp:303-307).
#include <cstring>

#include "kernels.hpp"

using namespace tme;

namespace {

thread_local std::string g_error;

// Device scratch of tm_operator_inertia_noalloc: one problem's factor workspace for
// meshes up to this size (the 24-sector fixtures and the 4.9k-DOF flat disk fit; the
// 45k-DOF bladed sector does not).  Every device the library runs on holds its own copy
// until the process ends.  Host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device (Context::scratch), never through
// the array's name.
constexpr size_t kScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

template <class F>
int guarded(Context* ctx, F&& f) {
  try {
    f();
    return TM_OK;
  } catch (const ArgError& e) {
    (ctx ? ctx->last_error : g_error) = e.what();
    return TM_ERR_ARG;
  } catch (const std::invalid_argument& e) {
    (ctx ? ctx->last_error : g_error) = e.what();
    return TM_ERR_ARG;
  } catch (const HipError& e) {
    (ctx ? ctx->last_error : g_error) = e.what();
    const std::string m = e.what();
    return m.find("hipMalloc") != std::string::npos ? TM_ERR_MEMORY : TM_ERR_HIP;
  } catch (const NumericError& e) {
    (ctx ? ctx->last_error : g_error) = e.what();
    return TM_ERR_NUMERIC;
  } catch (const std::bad_alloc& e) {
    (ctx ? ctx->last_error : g_error) = "host out of memory";
    return TM_ERR_MEMORY;
  } catch (const std::exception& e) {
    (ctx ? ctx->last_error : g_error) = e.what();
    return TM_ERR_RUNTIME;
  }
}

void upload_mesh(Mesh& M, hipStream_t s) {
  const Symbolic& S = M.S;
  M.d_coords = upload(S.coords, s);
  M.d_elements = upload(S.elements, s);
  M.d_rowptr = upload(S.rowptr, s);
  M.d_colidx = upload(S.colidx, s);
  M.d_tslot = upload(S.tslot, s);
  {
    std::vector<int> row(S.colidx.size());
    for (int I = 0; I + 1 < (int)S.rowptr.size(); ++I)
      for (int q = S.rowptr[I]; q < S.rowptr[I + 1]; ++q) row[q] = I;
    M.d_slot_row = upload(row, s);
  }
  {
    std::vector<int> diag(S.nf, -1);
    for (int I = 0; I < S.nf; ++I)
      for (int q = S.rowptr[I]; q < S.rowptr[I + 1]; ++q)
        if (S.colidx[q] == I) diag[I] = q;
    for (int I = 0; I < S.nf; ++I)
      if (diag[I] < 0) throw std::runtime_error("pattern without a diagonal block");
    M.d_diag_slot = upload(diag, s);
  }
  M.d_kc_ptr = upload(S.kc_ptr, s);
  M.d_kc_src = upload(S.kc_src, s);
  M.d_kp_ptr = upload(S.kp_ptr, s);
  M.d_kp_src = upload(S.kp_src, s);
  M.d_f_ptr = upload(S.f_ptr, s);
  M.d_f_src = upload(S.f_src, s);
  M.d_pinned = upload(S.pinned, s);
  M.d_free_of_full = upload(S.free_of_full, s);
  M.d_node_class = upload(S.node_class, s);
  M.d_struct = upload(S.struct_nodes, s);
  M.d_relmap = upload(S.relmap, s);
  M.d_asm_slot = upload(S.asm_slot, s);
  M.d_asm_pos = upload(S.asm_pos, s);
}

}  // namespace

extern "C" {

int tm_context_create(int device, tm_context** out) {
  if (!out) return TM_ERR_ARG;
  *out = nullptr;
  return guarded(nullptr, [&] {
    int ndev = 0;
    TM_HIP(hipGetDeviceCount(&ndev));
    if (ndev <= 0) throw HipError("no HIP device visible");
    if (device < 0) TM_HIP(hipGetDevice(&device));
    if (device >= ndev) throw ArgError("device index out of range");
    TM_HIP(hipSetDevice(device));
    // this device's copy of the scratch array (with the device current)
    void* sp = nullptr;
    TM_HIP(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
    auto* ctx = new tm_context();
    ctx->device = device;
    ctx->scratch = sp;
    ctx->scratch_bytes = kScratchBytes;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete ctx;
      throw HipError(std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    *out = ctx;
  });
}

void tm_context_destroy(tm_context* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->groups.clear();  // the slot groups' contexts
  ctx->deliver.reset();  // drains its copies
  ctx->clear_graphs();
  for (auto& p : ctx->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  ctx->release_side();
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* tm_last_error(const tm_context* ctx) {
  return ctx ? ctx->last_error.c_str() : g_error.c_str();
}

int tm_set_profiling(tm_context* ctx, int enabled) {
  if (!ctx) return TM_ERR_ARG;
  return guarded(ctx, [&] {
    ctx->flush_timing();
    ctx->profiling = enabled != 0;
  });
}

int tm_kernel_stats(tm_context* ctx, tm_kernel_stat* stats, int max_stats) {
  if (!ctx) return TM_ERR_ARG;
  int count = 0;
  int rc = guarded(ctx, [&] {
    ctx->flush_timing();
    for (auto& [name, s] : ctx->stats) {
      if (stats && count < max_stats) {
        tm_kernel_stat& o = stats[count];
        std::memset(&o, 0, sizeof(o));
        std::strncpy(o.name, name.c_str(), sizeof(o.name) - 1);
        o.launches = s.launches;
        o.total_ms = s.total_ms;
        o.alg_bytes = s.bytes;
        o.alg_flops = s.flops;
      }
      ++count;
    }
  });
  return rc == TM_OK ? count : rc;
}

int tm_reset_kernel_stats(tm_context* ctx) {
  if (!ctx) return TM_ERR_ARG;
  return guarded(ctx, [&] {
    ctx->flush_timing();
    ctx->stats.clear();
  });
}

int tm_mesh_create(tm_context* ctx, const double* nodes, int32_t n_nodes, const int32_t* elements,
                   int32_t n_elems, const int32_t* node_class, const int32_t* right_partner,
                   const uint8_t* constrained, int32_t num_sectors, int32_t rotation_axis,
                   tm_mesh** out) {
  if (!ctx || !out) return TM_ERR_ARG;
  *out = nullptr;
  return guarded(ctx, [&] {
    if (!nodes || !elements || !node_class || !right_partner || !constrained)
      throw ArgError("null mesh array");
    if (n_nodes <= 0 || n_elems <= 0) throw ArgError("empty mesh");
    if (num_sectors <= 0) throw ArgError("num_sectors must be positive");
    if (rotation_axis < 0 || rotation_axis > 2)
      throw ArgError("rotation_axis must be 0 (X), 1 (Y), or 2 (Z)");
    for (int64_t i = 0; i < 10LL * n_elems; ++i)
      if (elements[i] < 0 || elements[i] >= n_nodes) throw ArgError("element references unknown node");
    TM_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<tm_mesh> m(new tm_mesh());
    m->ctx = ctx;
    Symbolic& S = m->S;
    S.nn = n_nodes;
    S.ne = n_elems;
    S.N = num_sectors;
    S.axis = rotation_axis;
    S.coords.assign(nodes, nodes + 3LL * n_nodes);
    S.elements.assign(elements, elements + 10LL * n_elems);
    S.node_class.assign(node_class, node_class + n_nodes);
    S.partner.assign(right_partner, right_partner + n_nodes);
    S.cmask.assign(constrained, constrained + 3LL * n_nodes);
    build_symbolic(S);
    upload_mesh(*m, ctx->stream);
    build_plan(S, m->plan, ctx->stream);
    TM_HIP(hipStreamSynchronize(ctx->stream));
    *out = m.release();
  });
}

void tm_mesh_destroy(tm_mesh* mesh) {
  if (!mesh) return;
  (void)hipSetDevice(mesh->ctx->device);
  (void)hipStreamSynchronize(mesh->ctx->stream);
  delete mesh;
}

int tm_mesh_get_info(const tm_mesh* mesh, tm_mesh_info* info) {
  if (!mesh || !info) return TM_ERR_ARG;
  const Symbolic& S = mesh->S;
  info->n_nodes = S.nn;
  info->n_elems = S.ne;
  info->n_reduced_nodes = S.n_reduced_nodes;
  info->n_free_nodes = S.nf;
  info->n_free = S.n_free_dofs;
  info->n_dof_system = 3 * S.nf;
  info->nnz_blocks = static_cast<int64_t>(S.colidx.size());
  info->nnz_L = S.nnz_L;
  info->factor_flops = S.flops;
  info->n_supernodes = static_cast<int32_t>(S.sn.size());
  info->tree_levels = static_cast<int32_t>(S.levels.size());
  info->max_front = S.max_front;
  info->front_bytes_fp64 =
      16 * (mesh->plan.panel_entries + mesh->plan.upool_entries + mesh->plan.tvec_entries);
  return TM_OK;
}

int tm_operator_create(tm_context* ctx, tm_mesh* mesh, double E, double nu, double rho,
                       int32_t prestress, tm_operator** out) {
  if (!ctx || !mesh || !out) return TM_ERR_ARG;
  *out = nullptr;
  return guarded(ctx, [&] {
    if (E <= 0.0) throw ArgError("Young's modulus E must be positive");
    if (nu <= 0.0 || nu >= 0.5) throw ArgError("Poisson's ratio nu must be in (0, 0.5)");
    if (rho <= 0.0) throw ArgError("Density rho must be positive");
    TM_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<tm_operator> op(new tm_operator());
    op->mesh = mesh;
    op->E = E;
    op->nu = nu;
    op->rho = rho;
    op->prestress = prestress != 0;
    create_operator(ctx, *mesh, *op);
    *out = op.release();
  });
}

void tm_operator_destroy(tm_operator* op) {
  if (!op) return;
  (void)hipSetDevice(op->mesh->ctx->device);
  (void)hipStreamSynchronize(op->mesh->ctx->stream);
  delete op;
}

int tm_solve_batch(tm_context* ctx, const tm_problem* problems, int32_t n, tm_result* results) {
  if (!ctx || (n > 0 && !problems)) return TM_ERR_ARG;
  if (n <= 0) return TM_OK;
  return guarded(ctx, [&] {
    if (ctx->server) throw std::runtime_error("the context is owned by a server; submit to it instead");
    for (int i = 0; i < n; ++i)
      if (problems[i].precision != TM_PREC_FP64 && problems[i].precision != TM_PREC_MIXED)
        throw ArgError("precision must be TM_PREC_FP64 or TM_PREC_MIXED");
    TM_HIP(hipSetDevice(ctx->device));
    for (int i = 0; i < n; ++i)
      if (results) {
        results[i].n_modes = 0;
        results[i].n_unique = 0;
        results[i].converged = 0;
        results[i].restarts = 0;
        results[i].status = TM_ERR_RUNTIME;
        results[i].seconds = 0;
        results[i].factor_retries = 0;
        results[i].pivot_flags = 0;
        results[i].shift_used = problems[i].shift;
      }
    solve_problems(ctx, problems, n, results);
  });
}

int tm_server_create(tm_context* ctx, int32_t max_slots, double hbm_fraction, tm_server** out) {
  if (!ctx || !out) return TM_ERR_ARG;
  *out = nullptr;
  return guarded(ctx, [&] {
    if (ctx->server) throw std::runtime_error("the context already has a server");
    TM_HIP(hipSetDevice(ctx->device));
    ctx->flush_timing();
    auto* s = new tm_server{ctx, server_create(ctx, max_slots, hbm_fraction)};
    ctx->server = s->s;
    *out = s;
  });
}

int tm_server_submit(tm_server* s, const tm_problem* p, tm_result* r, uint64_t* ticket) {
  if (!s || !p || !ticket) return TM_ERR_ARG;
  return guarded(s->ctx, [&] { *ticket = server_submit(s->s, *p, r); });
}

int tm_server_wait(tm_server* s, uint64_t ticket, double timeout_s, int32_t* done) {
  if (!s || !done) return TM_ERR_ARG;
  *done = 0;
  return guarded(s->ctx, [&] { *done = server_wait(s->s, ticket, timeout_s) ? 1 : 0; });
}

void tm_server_destroy(tm_server* s) {
  if (!s) return;
  server_destroy(s->s);
  s->ctx->server = nullptr;
  delete s;
}

int64_t tm_problem_footprint(const tm_operator* op, int32_t num_modes, int32_t precision,
                             int32_t harmonic, int32_t method) {
  if (!op || (precision != TM_PREC_FP64 && precision != TM_PREC_MIXED)) return -1;
  try {
    return problem_footprint(*op, num_modes, harmonic, method, precision);
  } catch (...) {
    return -1;
  }
}

int tm_set_memory_reserve(tm_context* ctx, double reserve_fraction) {
  if (!ctx) return TM_ERR_ARG;
  return guarded(ctx, [&] {
    if (!(reserve_fraction >= 0.0 && reserve_fraction <= 1.0))
      throw ArgError("memory_reserve_fraction must lie in [0, 1]");
    ctx->batch_fraction = 1.0 - reserve_fraction;
  });
}

int64_t tm_device_free_bytes(tm_context* ctx) {
  if (!ctx) return -1;
  size_t f = 0, t = 0;
  if (hipSetDevice(ctx->device) != hipSuccess || hipMemGetInfo(&f, &t) != hipSuccess) return -1;
  return static_cast<int64_t>(f);
}

int tm_operator_static_u(tm_operator* op, double* u_full) {
  if (!op || !u_full) return TM_ERR_ARG;
  std::copy(op->u_static.begin(), op->u_static.end(), u_full);
  return TM_OK;
}

int tm_operator_add_springs(tm_operator* op, const double* spring_full) {
  if (!op || !spring_full) return TM_ERR_ARG;
  Context* ctx = op->mesh->ctx;
  return guarded(ctx, [&] {
    TM_HIP(hipSetDevice(ctx->device));
    add_springs(ctx, *op, spring_full);
  });
}

int tm_operator_inertia(tm_operator* op, double rpm, int32_t harmonic, double shift, int64_t* n_negative,
                        double* min_pivot_ratio, int32_t* pivot_status) {
  if (!op || !n_negative) return TM_ERR_ARG;
  Context* ctx = op->mesh->ctx;
  return guarded(ctx, [&] {
    if (harmonic < 0 || harmonic > op->mesh->S.N / 2) throw ArgError("harmonic index out of range 0..N/2");
    if (!std::isfinite(shift) || !std::isfinite(rpm)) throw ArgError("rpm and shift must be finite");
    TM_HIP(hipSetDevice(ctx->device));
    double r = 0;
    int st = 0;
    operator_inertia(ctx, *op, rpm, harmonic, shift, n_negative, &r, &st);
    if (min_pivot_ratio) *min_pivot_ratio = r;
    if (pivot_status) *pivot_status = st;
  });
}

int tm_operator_inertia_noalloc(tm_operator* op, double rpm, int32_t harmonic, double shift,
                                int64_t* n_negative, double* min_pivot_ratio, int32_t* pivot_status) {
  if (!op || !n_negative) return TM_ERR_ARG;
  Context* ctx = op->mesh->ctx;
  return guarded(ctx, [&] {
    if (harmonic < 0 || harmonic > op->mesh->S.N / 2) throw ArgError("harmonic index out of range 0..N/2");
    if (!std::isfinite(shift) || !std::isfinite(rpm)) throw ArgError("rpm and shift must be finite");
    if (!ctx->scratch) throw HipError("device scratch not resolved");
    TM_HIP(hipSetDevice(ctx->device));
    double r = 0;
    int st = 0;
    operator_inertia_in(ctx, *op, rpm, harmonic, shift, ctx->scratch, ctx->scratch_bytes, n_negative, &r, &st);
    if (min_pivot_ratio) *min_pivot_ratio = r;
    if (pivot_status) *pivot_status = st;
  });
}

int tm_operator_export(tm_operator* op, double rpm, int32_t harmonic, int32_t which, int64_t* nnz,
                       int32_t* rows, int32_t* cols, double* vals) {
  if (!op || !nnz) return TM_ERR_ARG;
  Context* ctx = op->mesh->ctx;
  return guarded(ctx, [&] {
    const Symbolic& S = op->mesh->S;
    const size_t nb = op->nblk;
    auto dl = [&](const DevBuf& b) {
      std::vector<double> h(9 * nb);
      TM_HIP(hipMemcpy(h.data(), b.p, sizeof(double) * 9 * nb, hipMemcpyDeviceToHost));
      return h;
    };
    std::vector<double> Xc, Xp, dXc, dXp;
    if (which == 0) {
      Xc = dl(op->Kc); Xp = dl(op->Kp); dXc = dl(op->dKc); dXp = dl(op->dKp);
    } else if (which == 1) {
      Xc = dl(op->Mc); Xp = dl(op->Mp);
    } else if (which == 2) {
      Xc = dl(op->Gc); Xp = dl(op->Gp);
    } else {
      throw ArgError("which must be 0 (K), 1 (M) or 2 (G)");
    }
    const double w = rpm * 2.0 * M_PI / 60.0, s = w * w;
    double pre, pim;
    harmonic_phase(S.N, harmonic, pre, pim);
    std::vector<int> sys_to_ref(3 * S.nf, -1);
    for (size_t q = 0; q < S.ref_free_to_sys.size(); ++q) sys_to_ref[S.ref_free_to_sys[q]] = (int)q;
    int64_t cnt = 0;
    for (int I = 0; I < S.nf; ++I)
      for (int sl = S.rowptr[I]; sl < S.rowptr[I + 1]; ++sl) {
        const int J = S.colidx[sl], tt = S.tslot[sl];
        for (int ci = 0; ci < 3; ++ci)
          for (int cj = 0; cj < 3; ++cj) {
            const int r = sys_to_ref[3 * I + ci], c = sys_to_ref[3 * J + cj];
            if (r < 0 || c < 0) continue;
            const size_t e1 = (size_t)sl * 9 + 3 * ci + cj, e2 = (size_t)tt * 9 + 3 * cj + ci;
            double cv = Xc[e1], pv = Xp[e1], qv = Xp[e2];
            if (!dXc.empty()) {
              cv += s * dXc[e1];
              pv += s * dXp[e1];
              qv += s * dXp[e2];
            }
            if (rows) {
              rows[cnt] = r;
              cols[cnt] = c;
              if (which == 2) {  // G is skew: G(k) = G_c + P G_p - conj(P) G_p^T
                vals[2 * cnt] = cv + pre * (pv - qv);
                vals[2 * cnt + 1] = pim * (pv + qv);
              } else {           // K, M symmetric: X_c + P X_p + conj(P) X_p^T
                vals[2 * cnt] = cv + pre * (pv + qv);
                vals[2 * cnt + 1] = pim * (pv - qv);
              }
            }
            ++cnt;
          }
      }
    *nnz = cnt;
  });
}

}  // extern "C"
