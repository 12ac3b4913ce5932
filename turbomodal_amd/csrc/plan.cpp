// Execution plan of the supernodal multifrontal LDL^H: per tree level, the
// task tables of every kernel (front assembly, diagonal-block factorization,
// panel solve, Schur tiles, triangular-solve row/column chunks) and the
// memory layout of one problem (L panels, a level-scheduled pool for the
// update matrices, update vectors of the triangular solves).  Built once per
// mesh; every numeric factorization and solve of every problem that shares
// the mesh replays it.
#include <algorithm>
#include <atomic>
#include <limits>
#include <stdexcept>

#include "kernels.hpp"

namespace tme {

namespace {

// First-fit interval allocator (deterministic) used to lay out update matrices.
class Pool {
 public:
  int64_t alloc(int64_t n) {
    if (n == 0) return 0;
    for (size_t i = 0; i < free_.size(); ++i)
      if (free_[i].second >= n) {
        int64_t off = free_[i].first;
        free_[i].first += n;
        free_[i].second -= n;
        if (free_[i].second == 0) free_.erase(free_.begin() + i);
        return off;
      }
    int64_t off = top_;
    top_ += n;
    return off;
  }
  void release(int64_t off, int64_t n) {
    if (n == 0) return;
    free_.push_back({off, n});
    std::sort(free_.begin(), free_.end());
    std::vector<std::pair<int64_t, int64_t>> merged;
    for (auto& f : free_) {
      if (!merged.empty() && merged.back().first + merged.back().second == f.first)
        merged.back().second += f.second;
      else
        merged.push_back(f);
    }
    free_ = merged;
    // shrink the top when its last block is free
    while (!free_.empty() && free_.back().first + free_.back().second == top_) {
      top_ = free_.back().first;
      free_.pop_back();
    }
  }
  int64_t peak() const { return peak_; }
  void note() { peak_ = std::max(peak_, top_); }

 private:
  std::vector<std::pair<int64_t, int64_t>> free_;
  int64_t top_ = 0, peak_ = 0;
};

}  // namespace

void build_plan(const Symbolic& S, Plan& P, hipStream_t stream) {
  static std::atomic<uint64_t> plans{0};
  P.serial = ++plans;
  const int nsn = static_cast<int>(S.sn.size());
  const std::vector<char> shared = shared_fronts(S);
  std::vector<DevSn> dsn(nsn);
  int64_t poff = 0, toff = 0;
  P.flops_shared = P.flops_other = 0;
  for (int pass = 1; pass >= 0; --pass) {  // shared fronts' panels first
    for (int id = 0; id < nsn; ++id) {
      if (shared[id] != pass) continue;
      const Snode& s = S.sn[id];
      dsn[id].poff = poff;
      poff += int64_t(s.m()) * s.w();
      double fl = 0;
      for (int64_t j = 0; j < s.w(); ++j) fl += 4.0 * double(s.m() - j) * double(s.m() - j);
      (pass ? P.flops_shared : P.flops_other) += fl;
    }
    if (pass) P.shared_panel_entries = poff;
  }
  P.has_shared = P.shared_panel_entries > 0;
  for (int id = 0; id < nsn; ++id) {
    const Snode& s = S.sn[id];
    DevSn& d = dsn[id];
    d.col0 = s.col0;
    d.ncol = s.ncol;
    d.nstr = s.nstr;
    d.str0 = s.str0;
    d.rel0 = s.rel0;
    d.child0 = s.child[0];
    d.child1 = s.child[1];
    d.inv0 = 0;
    d.shared = shared[id];
    d.pad_ = 0;
    d.toff = toff;
    toff += s.r() + s.w();
    d.soff = 0;
  }
  // inverse relmaps: for each child of a front (child0's map, then child1's),
  // the child's update node at each front node position, or -1
  std::vector<int> invmap;
  for (int id = 0; id < nsn; ++id) {
    const Snode& s = S.sn[id];
    dsn[id].inv0 = static_cast<int>(invmap.size());
    const int nodes = s.ncol + s.nstr;
    for (int k = 0; k < 2; ++k) {
      const int c = s.child[k];
      if (c < 0) continue;
      const size_t base = invmap.size();
      invmap.resize(base + nodes, -1);
      const Snode& cs = S.sn[c];
      for (int q = 0; q < cs.nstr; ++q) invmap[base + S.relmap[cs.rel0 + q]] = q;
    }
    if (invmap.size() > (size_t)std::numeric_limits<int>::max()) throw std::length_error("invmap");
  }
  if (invmap.empty()) invmap.push_back(-1);
  // inversion scratch: w x 64 per front (Z = L X_JJ of the current block column), laid out per level
  int64_t scratch = 0;
  for (const auto& lvl : S.levels) {
    int64_t o = 0;
    for (int id : lvl) {
      dsn[id].soff = o;
      o += int64_t(S.sn[id].w()) * 64;
    }
    scratch = std::max(scratch, o);
  }
  P.inv_scratch_entries = std::max<int64_t>(scratch, 1);
  Pool pool;
  for (const auto& lvl : S.levels) {
    for (int id : lvl) {
      const Snode& s = S.sn[id];
      dsn[id].uoff = pool.alloc(int64_t(s.r()) * s.r());
    }
    pool.note();
    for (int id : lvl)
      for (int c : S.sn[id].child)
        if (c >= 0) pool.release(dsn[c].uoff, int64_t(S.sn[c].r()) * S.sn[c].r());
  }
  P.panel_entries = poff;
  P.tvec_entries = std::max<int64_t>(toff, 1);
  P.upool_entries = std::max<int64_t>(pool.peak(), 1);

  P.levels.clear();
  P.slevels.clear();
  P.clevels.clear();
  P.asm_tasks.clear();
  P.diag_tasks.clear();
  P.row_tasks.clear();
  P.tile_tasks.clear();
  P.level_sn.clear();
  P.finv_tasks.clear();
  P.fup_tasks.clear();
  P.bcol_tasks.clear();
  P.inv_diag.clear();
  P.inv_tiles.clear();
  const int NB = P.nb, TL = P.tile;
  constexpr int kAsmChunk = 16;   // node columns per assembly task
  constexpr int kRowChunk = 256;  // rows per panel-solve / forward-update task
  constexpr int kInvChunk = 64;   // rows per L11^-1 task
  constexpr int kColChunk = 64;   // pivot columns per backward task
  // the factorization's tasks of the fronts `fr` of one level, appended to the task tables
  auto factor_tasks = [&](const std::vector<int>& fr, LevelPlan& L) {
    L.asm0 = static_cast<int>(P.asm_tasks.size());
    int maxw = 0;
    for (int id : fr) {
      const Snode& s = S.sn[id];
      maxw = std::max(maxw, s.w());
      // pivot node columns only: the update matrix is written by the first
      // trailing Schur pass from the children's update matrices
      for (int n0 = 0; n0 < s.ncol; n0 += kAsmChunk) {
        AsmTask t{};
        t.sn = id;
        t.n0 = n0;
        t.n1 = std::min(s.ncol, n0 + kAsmChunk);
        // asm entries are ordered by column node (colpos = pos & 0xffff)
        int a = s.asm0, b = s.asm1;
        auto colpos = [&](int q) { return S.asm_pos[q] & 0xffff; };
        int lo = a, hi = a;
        while (lo < b && colpos(lo) < t.n0) ++lo;
        hi = lo;
        while (hi < b && colpos(hi) < t.n1) ++hi;
        t.a0 = lo;
        t.a1 = hi;
        for (int k = 0; k < 2; ++k) {
          int c = s.child[k];
          t.clo[k] = t.chi[k] = 0;
          if (c < 0) continue;
          const Snode& cs = S.sn[c];
          const int* rm = S.relmap.data() + cs.rel0;
          int l = static_cast<int>(std::lower_bound(rm, rm + cs.nstr, t.n0) - rm);
          int h = static_cast<int>(std::lower_bound(rm, rm + cs.nstr, t.n1) - rm);
          t.clo[k] = l;
          t.chi[k] = h;
        }
        P.asm_tasks.push_back(t);
      }
    }
    L.nasm = static_cast<int>(P.asm_tasks.size()) - L.asm0;
    // super-panels of KW columns made of NB-wide sub-panels: each sub-panel
    // updates the rest of its super-panel (narrow update, depth NB); the last
    // one updates the trailing matrix with the whole super-panel (depth <= KW)
    const int KW = P.kw;
    for (int ks = 0; ks < maxw; ks += KW) {
      for (int kb = ks; kb < std::min(maxw, ks + KW); kb += NB) {
        const bool last = kb + NB >= ks + KW;
        FactorStep st{};
        st.kb = kb;
        st.diag0 = static_cast<int>(P.diag_tasks.size());
        st.trsm0 = static_cast<int>(P.row_tasks.size());
        st.tile0 = static_cast<int>(P.tile_tasks.size());
        for (int id : fr) {
          const Snode& s = S.sn[id];
          const int w = s.w(), m = s.m();
          if (w <= kb) continue;
          const int jb = std::min(NB, w - kb);
          P.diag_tasks.push_back(id);
          const int t0 = kb + jb;
          for (int r0 = t0; r0 < m; r0 += kRowChunk)
            P.row_tasks.push_back({id, r0, std::min(m, r0 + kRowChunk)});
          const int se = std::min(w, ks + KW);           // end of this front's super-panel
          if (last || kb + NB >= w) {
            // super-panel complete: trailing matrix, depth se - ks, in groups of
            // 4 x 4 tiles so neighbouring tasks share row and column panels
            const int G = 4 * TL;
            if (ks == 0 && se < m) st.pull = 1;
            for (int J = se; J < m; J += G)
              for (int I = J; I < m; I += G)
                for (int j0 = J; j0 < std::min(m, J + G); j0 += TL)
                  for (int i0 = std::max(I, j0); i0 < std::min(m, I + G); i0 += TL)
                    P.tile_tasks.push_back({id, i0, j0, ks, se - ks, -1});
          } else {
            // the rest of the super-panel, columns [t0, se), depth jb
            for (int j0 = t0; j0 < se; j0 += TL)
              for (int i0 = j0; i0 < m; i0 += TL) P.tile_tasks.push_back({id, i0, j0, kb, jb, se});
          }
        }
        st.ndiag = static_cast<int>(P.diag_tasks.size()) - st.diag0;
        st.ntrsm = static_cast<int>(P.row_tasks.size()) - st.trsm0;
        st.ntile = static_cast<int>(P.tile_tasks.size()) - st.tile0;
        L.steps.push_back(st);
      }
    }
    const int maxblk = (maxw + 63) / 64;
    for (int J = maxblk - 1; J >= 0; --J) {
      InvStep iv{};
      iv.J = J;
      iv.diag0 = static_cast<int>(P.inv_diag.size());
      iv.tile0 = static_cast<int>(P.inv_tiles.size());
      for (int id : fr) {
        const int nblk = (S.sn[id].w() + 63) / 64;
        if (nblk <= J) continue;
        P.inv_diag.push_back(id);
        for (int I = J + 1; I < nblk; ++I) P.inv_tiles.push_back({id, I, J, 0, 0, -1});
      }
      iv.ndiag = static_cast<int>(P.inv_diag.size()) - iv.diag0;
      iv.ntile = static_cast<int>(P.inv_tiles.size()) - iv.tile0;
      L.inv.push_back(iv);
    }
  };
  double tri_bytes = 0, tri_shared = 0;
  for (const auto& lvl : S.levels) {
    LevelPlan L;
    L.sn0 = static_cast<int>(P.level_sn.size());
    L.nsn = static_cast<int>(lvl.size());
    P.level_sn.insert(P.level_sn.end(), lvl.begin(), lvl.end());
    L.finv0 = static_cast<int>(P.finv_tasks.size());
    L.fup0 = static_cast<int>(P.fup_tasks.size());
    L.bcol0 = static_cast<int>(P.bcol_tasks.size());
    for (int id : lvl) {
      const Snode& s = S.sn[id];
      for (int r0 = 0; r0 < s.w(); r0 += kInvChunk)
        P.finv_tasks.push_back({id, r0, std::min(s.w(), r0 + kInvChunk)});
      for (int r0 = 0; r0 < s.m(); r0 += kRowChunk)
        P.fup_tasks.push_back({id, r0, std::min(s.m(), r0 + kRowChunk)});
      for (int c0 = 0; c0 < s.w(); c0 += kColChunk)
        P.bcol_tasks.push_back({id, c0, std::min(s.w(), c0 + kColChunk)});
      // per sweep pair: L21 read twice (forward, backward), L11^-1 triangle read twice
      const double w = s.w(), r = s.r();
      const double fb = 2.0 * (r * w + w * (w - 1) / 2) * 20.0;
      tri_bytes += fb;
      if (shared[id]) tri_shared += fb;
    }
    L.nfinv = static_cast<int>(P.finv_tasks.size()) - L.finv0;
    L.nfup = static_cast<int>(P.fup_tasks.size()) - L.fup0;
    L.nbcol = static_cast<int>(P.bcol_tasks.size()) - L.bcol0;
    factor_tasks(lvl, L);
    std::vector<int> sh, ot;
    for (int id : lvl) (shared[id] ? sh : ot).push_back(id);
    LevelPlan Ls, Lo;
    factor_tasks(sh, Ls);
    factor_tasks(ot, Lo);
    P.levels.push_back(L);
    P.slevels.push_back(Ls);
    P.clevels.push_back(Lo);
  }
  P.trisolve_bytes_c16 = tri_bytes + 5.0 * 3 * S.nf * 16.0;
  P.trisolve_shared_bytes_c16 = tri_shared;
  P.d_sn = upload(dsn, stream);
  P.d_invmap = upload(invmap, stream);
  P.d_asm = upload(P.asm_tasks, stream);
  P.d_diag = upload(P.diag_tasks, stream);
  P.d_row = upload(P.row_tasks, stream);
  P.d_tile = upload(P.tile_tasks, stream);
  P.d_level_sn = upload(P.level_sn, stream);
  P.d_finv = upload(P.finv_tasks, stream);
  P.d_fup = upload(P.fup_tasks, stream);
  P.d_bcol = upload(P.bcol_tasks, stream);
  P.d_inv_diag = upload(P.inv_diag, stream);
  P.d_inv_tiles = upload(P.inv_tiles, stream);
}

}  // namespace tme
