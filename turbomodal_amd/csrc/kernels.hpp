// Launchers and device-side plan structures shared by the engine sources.
#pragma once

#include <cmath>

#include "engine.hpp"

namespace tme {

struct Mesh;
// ---- FEM (fem.hip) ----------------------------------------------------------
void launch_element_base(Context* ctx, Mesh& M, double E, double nu, double rho, DevBuf& ke,
                         DevBuf& m_abs, DevBuf& m_raw, DevBuf& fe);
void launch_element_sigma(Context* ctx, Mesh& M, double E, double nu, const DevBuf& u_full,
                          DevBuf& s_ab);
// kind: 0 = K (element stiffness), 1 = M, 2 = Ks1 - Kw1, 3 = G
void launch_assemble(Context* ctx, Mesh& M, int kind, bool phase, const DevBuf& ke,
                     const DevBuf& m_abs, const DevBuf& m_raw, const DevBuf* s_ab, DevBuf& out);
void launch_assemble_load(Context* ctx, Mesh& M, const DevBuf& fe, DevBuf& f);
void launch_expand_real(Context* ctx, Mesh& M, const double2* u, DevBuf& out);

// ---- per-problem parameters --------------------------------------------------
struct ProbParam {
  double pre, pim;   // P = exp(i k alpha)
  double s;          // omega^2 (rad^2/s^2): K_eff = K + s (Ks1 - Kw1)
  double sigma;      // shift
};

// Operator arrays (9 doubles per pattern slot).
struct OpArrays {
  const double *Kc, *Kp, *dKc, *dKp, *Mc, *Mp;
};

// ---- plans (built once per mesh from the symbolic analysis) ------------------
struct AsmTask {    // front assembly: pivot node columns [n0, n1) of snode sn
  int sn, n0, n1, a0, a1;
  int clo[2], chi[2];  // child struct node ranges that land in [n0, n1)
};
struct RowTask {    // rows [r0, r1) (DOF) of snode sn
  int sn, r0, r1;
};
struct TileTask {   // 64x64 lower tile of the trailing front
  int sn, i0, j0;
  int tk, kd;        // Schur update source columns [tk, tk + kd) (factor tiles)
  int jend;          // update columns < jend only (< 0: no limit)
};

// Inversion of the unit lower L11 of every front, 64-column block J (from the last).
struct InvStep {
  int J;
  int diag0, ndiag;    // snode ids (range in inv_diag)
  int tile0, ntile;    // (snode, row block I > J) tasks (range in inv_tiles; TileTask.i0 = I, j0 = J)
};

// Dense partial factorization plan step at one (level, kb).
// One sub-panel step of the blocked LDL^H: diagonal block and panel solve of
// columns [kb, kb + 32), then the Schur updates of its tile tasks (each names
// its source columns and column limit: the rest of the current super-panel, or
// the whole trailing matrix once a front's super-panel is complete).
struct FactorStep {
  int kb;
  int diag0, ndiag;    // range in diag task list (snode ids)
  int trsm0, ntrsm;    // range in row tasks
  int tile0, ntile;    // range in tile tasks
  int pull;            // some tile is a front's first trailing pass (writes its update matrix)
};

struct LevelPlan {
  int asm0, nasm;          // front assembly tasks
  std::vector<FactorStep> steps;
  std::vector<InvStep> inv;
  int sn0, nsn;            // snode ids of this level (range in level_sn)
  int finv0, nfinv;        // forward: rows of L11^-1 y (64-row chunks of the pivot block)
  int fup0, nfup;          // forward: rows [0, m) (pivot write-back + update vector)
  int bcol0, nbcol;        // backward: pivot columns (64 per task)
};

struct Plan {
  uint64_t serial = 0;     // distinct for every plan built in the process (graph keys)
  // the factorization's task ranges per level for the shared fronts and for the others
  // (both index the same task tables as `levels`); shared fronts' panels come first
  std::vector<LevelPlan> slevels, clevels;
  bool has_shared = false;
  int64_t shared_panel_entries = 0;
  double flops_shared = 0, flops_other = 0;  // real flops of one complex factorization
  int nb = 32;             // diagonal block width
  int tile = 64;           // Schur tile
  int kw = 128;            // super-panel width: depth of the trailing Schur update
  std::vector<LevelPlan> levels;
  std::vector<AsmTask> asm_tasks;
  std::vector<int> diag_tasks;
  std::vector<RowTask> row_tasks;
  std::vector<TileTask> tile_tasks;
  std::vector<int> level_sn;
  std::vector<RowTask> finv_tasks, fup_tasks, bcol_tasks;
  std::vector<int> inv_diag;
  std::vector<TileTask> inv_tiles;
  int64_t inv_scratch_entries = 0;  // complex entries per problem for the inversion
  // device copies
  DevBuf d_asm, d_diag, d_row, d_tile, d_level_sn, d_finv, d_fup, d_bcol, d_inv_diag, d_inv_tiles;
  // per-snode device data
  DevBuf d_sn;             // DevSn[]
  DevBuf d_invmap;         // per front and child: front node position -> child update node, or -1
  int64_t panel_entries = 0, upool_entries = 0, tvec_entries = 0;
  double trisolve_bytes_c16 = 0;  // algorithmic bytes of one L/D/L^H pair (complex fp64)
  double trisolve_shared_bytes_c16 = 0;  // the shared fronts' part of it (factor entries only)
};

struct DevSn {
  int col0, ncol, nstr, str0, rel0, child0, child1;
  int inv0;  // offset into invmap: per child, front node -> child update node (-1: none)
  int shared;  // the front and its subtree assemble no phase-coupled slot: A(k) there is real and
               // the same for every harmonic at one (operator, rpm, shift)
  int pad_;
  long long poff, uoff;
  long long toff;   // per column: [r update-vector entries][w pivot temp entries]
  long long soff;   // inversion scratch (w x 64) within its level
};

void build_plan(const Symbolic& S, Plan& P, hipStream_t stream);

// Device-resident mesh: topology, symbolic analysis and plan.
struct Mesh {
  Context* ctx = nullptr;
  Symbolic S;
  Plan plan;
  DevBuf d_coords, d_elements, d_rowptr, d_colidx, d_tslot, d_slot_row;  // slot_row: block row of each slot
  DevBuf d_kc_ptr, d_kc_src, d_kp_ptr, d_kp_src, d_f_ptr, d_f_src;
  DevBuf d_pinned, d_free_of_full, d_node_class;
  DevBuf d_struct, d_relmap, d_asm_slot, d_asm_pos;
  DevBuf d_diag_slot;  // per free node: slot of its diagonal block (I, I)
};

// Assembled operator (mesh + material): FP64 phase-split node blocks, 9 doubles
// per pattern slot.  K_eff(omega) = K + omega^2 dK with dK = Ks1 - Kw1.
struct Operator {
  Mesh* mesh = nullptr;
  double E = 0, nu = 0, rho = 0;
  bool prestress = false;
  size_t nblk = 0;
  DevBuf Kc, Kp, dKc, dKp, Mc, Mp, Gc, Gp;
  DevBuf real_slot;  // per slot: K(k), M(k) blocks real and unpinned for every harmonic (k_real_slots)
  DevBuf real_pack;  // per slot: Kc, dKc, Mc packed for the split K/M product (k_pack_real_slots)
  std::vector<double> u_static;       // 3*nn at unit omega
  // harmonic -> complex M(k) blocks, built on the operator's own stream (which runs
  // nothing else, so waiting for a build waits for that build only) before they are
  // handed to any user
  std::map<int, DevBuf> Mk_cache;
  hipStream_t mk_stream = nullptr;
  Operator() = default;
  Operator(const Operator&) = delete;
  Operator& operator=(const Operator&) = delete;
  ~Operator() {
    if (mk_stream) (void)hipStreamDestroy(mk_stream);
  }
};

// ---- factorization (factor.hip) ------------------------------------------------
// element type of a batch's factor: complex FP64, complex FP32 (TM_PREC_MIXED) or
// real FP64 (real harmonics, k = 0 and N/2, FP64)
enum class FactorElem : int { C64 = 0, C32 = 1, R64 = 2 };

// e^{i 2 pi k / N}, exactly 1 at k = 0 and -1 at k = N/2 (N even): the harmonics whose
// K(k), M(k) are real (the reference builds them from P = +-1, src/cyclic_solver.cpp:673-723)
inline bool real_harmonic(int N, int k) { return k == 0 || (N % 2 == 0 && 2 * k == N); }
inline void harmonic_phase(int N, int k, double& re, double& im) {
  if (real_harmonic(N, k)) {
    re = k == 0 ? 1.0 : -1.0;
    im = 0.0;
  } else {
    const double alpha = 2.0 * M_PI / N;
    re = std::cos(k * alpha);
    im = std::sin(k * alpha);
  }
}
inline size_t factor_elem_bytes(FactorElem e) { return e == FactorElem::C64 ? 16 : 8; }

struct BatchBufs {
  // the factor's panels, update matrices and inversion scratch, of element type elem
  void* L;           // B * panel_entries
  void* U;           // B * upool_entries
  void* S;           // B * inv_scratch_entries
  FactorElem elem = FactorElem::C64;
  double2* T;        // B * p * tvec_entries (p right-hand sides; FP64 in either case)
  int* status;       // B: bit 0 zero / non-finite pivot, bit 1 near-singular pivot
                     // (|d_j| <= kTinyPivot |a_jj|, a_jj the diagonal of A(k) before elimination)
  int* pivinfo;      // 2 B (may be null): negative pivots, bits of the smallest |d_j| / |a_jj| (float)
  const ProbParam* params;  // B
  int B;
  // Shared fronts (DevSn::shared) factored once per class of problems with the same
  // operator, rpm and shift: lead[b] is the first problem of b's class (itself for a
  // class's first problem), nlead the number of classes among the factored problems,
  // share_ws 4 B ints of workspace.  The shared fronts' panels live in the leader's
  // storage only: every member's triangular sweeps read them there, so the caller keeps
  // a leader's storage in place until no member of its class is solved any more.
  // lead null: every problem factors every front into its own storage.
  const int* lead = nullptr;
  int* share_ws = nullptr;
  int nlead = 0;
};

// A factorization is flagged near-singular (status bit 1) when a pivot d_j has
// |d_j| <= kTinyPivot |a_jj| (a_jj the diagonal of A(k) before elimination) and
// |d_j| <= kTinyShiftPivot |sigma M_jj|.  On the reference's fixtures a shift on an
// eigenvalue leaves |d_j| / |a_jj| at 1e-14 .. 1e-10 (thin plates at the top of that
// range), a shift 1e-6 away from one at 1e-9 .. 1e-7, and sigma = 1 below the spectrum
// at 1e-6 .. 1e-4; sigma = 1 above the rigid-body modes of a free structure leaves
// about 1e-12, but there the pivot is of the size of sigma M_jj, not far below it.
constexpr double kTinyPivot = 1e-10;
constexpr double kTinyShiftPivot = 1e-3;

void factor_batch(Context* ctx, const Mesh& M, const Plan& P, const OpArrays& ops,
                  const BatchBufs& bb, const int* active = nullptr, int nactive = -1,
                  const OpArrays* ops_b = nullptr, bool graph = false, bool tune = false);
// in-place solve A X = B for each problem's block X_b = X + b*ldx: p columns of
// length n = 3 nf, column c at X_b + c*n.  bb.T must hold p * tvec_entries per problem.
// nactive: how many problems `active` selects (-1: all B), for the byte count only;
// nclass: how many classes (bb.lead) they form, whose shared fronts count once each
// (-1: once per problem).
// graph: replay the level launches from ctx's graph cache (run_captured); for callers
// whose argument buffers stay in place from call to call.
void trisolve_batch(Context* ctx, const Mesh& M, const Plan& P, const BatchBufs& bb, double2* X,
                    long long ldx, const int* active, int p = 1, long long colstride = -1, int nactive = -1,
                    bool graph = false, int nclass = -1);

struct HostProblem;
void create_operator(Context* ctx, Mesh& M, Operator& op);
void add_springs(Context* ctx, Operator& op, const double* spring_full);
void operator_inertia(Context* ctx, Operator& op, double rpm, int k, double shift, int64_t* neg, double* ratio,
                      int* status);
// the same, its one-problem factor workspace carved out of work (work_bytes of device
// memory the caller owns); no allocation
void operator_inertia_in(Context* ctx, Operator& op, double rpm, int k, double shift, void* work, size_t work_bytes,
                         int64_t* neg, double* ratio, int* status);
OpArrays op_arrays(const Operator& op);
const double2* harmonic_mass(Context* ctx, Operator& op, int k);
// solves every problem; results written through each problem's tm_result
void solve_problems(Context* ctx, const tm_problem* problems, int n, tm_result* results);

// continuous-batching server (driver.hip)
struct Server;
Server* server_create(Context* ctx, int max_slots, double hbm_fraction);
void server_destroy(Server* s);
uint64_t server_submit(Server* s, const tm_problem& p, tm_result* out);
bool server_wait(Server* s, uint64_t ticket, double timeout_s);
int64_t problem_footprint(const Operator& op, int num_modes, int harmonic, int method, int precision);

}  // namespace tme

struct tm_context : tme::Context {
  tme::Server* server = nullptr;  // while set, the server's thread owns the context
};
struct tm_server {
  tm_context* ctx;
  tme::Server* s;
};
struct tm_mesh : tme::Mesh {};
struct tm_operator : tme::Operator {};
