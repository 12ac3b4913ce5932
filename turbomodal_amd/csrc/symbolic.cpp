// Symbolic analysis of a cyclic sector mesh: all integer work, done once per
// mesh on the host.
//
// * reduced/free node map: interior and left nodes are the reduced nodes
//   (src/cyclic_solver.cpp:214-257); every right node stands for its matched
//   left partner; nodes whose three DOFs are constrained are removed, nodes
//   with one or two constrained DOFs keep them as pinned rows
//   (build_constrained_dofs, src/cyclic_solver.cpp:107-193);
// * the K(k) pattern as 3x3 node blocks: the union of the phase-free part
//   (same right-membership of row and column) and the phase part (row not
//   right, column right) plus its transpose (src/cyclic_solver.cpp:606-631);
// * per pattern slot, the ordered list of element node-pair blocks that sum
//   into it (elements ascending, then a, then b: the reference's serial
//   triplet order, src/assembler.cpp:43-62, 274-289), with the flags saying
//   which side is rotated by R(alpha) (the T0 map, src/cyclic_solver.cpp:291-327);
// * a nested-dissection ordering of the free-node graph (geometric cuts, each
//   separator a minimum vertex cover of the cut edges) and the supernodal
//   assembly tree it induces (each separator and each leaf is one dense
//   front).  The reference orders with Eigen's AMD; the ordering changes
//   rounding only, never the eigenpairs.
#include <algorithm>
#include <functional>
#include <cmath>
#include <numeric>

#include "engine.hpp"

namespace tme {

namespace {

constexpr int kLeafNodes = 32;

struct NDTree {
  std::vector<std::vector<int>> nodes;
  std::vector<int> c0, c1;
  int add() {
    nodes.emplace_back();
    c0.push_back(-1);
    c1.push_back(-1);
    return static_cast<int>(nodes.size()) - 1;
  }
};

void jacobi3(double a[3][3], double v[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) v[i][j] = (i == j);
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = std::fabs(a[0][1]) + std::fabs(a[0][2]) + std::fabs(a[1][2]);
    if (off < 1e-300) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (std::fabs(a[p][q]) < 1e-300) continue;
        double th = 0.5 * (a[q][q] - a[p][p]) / a[p][q];
        double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {
          double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          double vkp = v[k][p], vkq = v[k][q];
          v[k][p] = c * vkp - s * vkq;
          v[k][q] = s * vkp + c * vkq;
        }
      }
  }
}

class NestedDissection {
 public:
  NestedDissection(const std::vector<std::vector<int>>& adj, const std::vector<double>& xyz,
                   int c1, int c2, int cax)
      : adj_(adj), xyz_(xyz), c1_(c1), c2_(c2), cax_(cax),
        stamp_(adj.size(), 0), side_(adj.size(), 0) {}

  NDTree run() {
    std::vector<int> all(adj_.size());
    std::iota(all.begin(), all.end(), 0);
    root_ = dissect(all);
    return std::move(tree_);
  }
  int root() const { return root_; }

 private:
  const std::vector<std::vector<int>>& adj_;
  const std::vector<double>& xyz_;
  int c1_, c2_, cax_;
  std::vector<int> stamp_;
  std::vector<char> side_;
  int cur_ = 0;
  NDTree tree_;
  int root_ = -1;

  double proj(int v, int dir, const double* pca) const {
    const double* p = &xyz_[3 * v];
    switch (dir) {
      case 0: return p[0];
      case 1: return p[1];
      case 2: return p[2];
      case 3: return std::hypot(p[c1_], p[c2_]);
      case 4: return std::atan2(p[c2_], p[c1_]);
      default: {
        const double* d = pca + 3 * (dir - 5);
        return p[0] * d[0] + p[1] * d[1] + p[2] * d[2];
      }
    }
  }

  // Minimum vertex cover of the edges between the boundary sets sa (side 0) and
  // sb (side 1) of the current split (stamp_ == cur_, side_ set): a maximum
  // matching by augmenting paths, then Koenig's construction.
  std::vector<int> min_cover(const std::vector<int>& sa, const std::vector<int>& sb) {
    const int na = (int)sa.size(), nb = (int)sb.size();
    if (mpos_.size() < adj_.size()) mpos_.assign(adj_.size(), -1);
    for (int i = 0; i < nb; ++i) mpos_[sb[i]] = i;
    std::vector<std::vector<int>> e(na);
    for (int i = 0; i < na; ++i)
      for (int u : adj_[sa[i]])
        if (stamp_[u] == cur_ && side_[u] == 1 && mpos_[u] >= 0 && mpos_[u] < nb && sb[mpos_[u]] == u)
          e[i].push_back(mpos_[u]);
    std::vector<int> ma(na, -1), mb(nb, -1), seen(nb, -1);
    int round = 0;
    std::function<bool(int)> aug = [&](int a) -> bool {
      for (int b : e[a]) {
        if (seen[b] == round) continue;
        seen[b] = round;
        if (mb[b] < 0 || aug(mb[b])) {
          ma[a] = b;
          mb[b] = a;
          return true;
        }
      }
      return false;
    };
    for (int a = 0; a < na; ++a) {
      ++round;
      aug(a);
    }
    // Z: reachable from unmatched side-0 vertices by alternating paths
    std::vector<char> za(na, 0), zb(nb, 0);
    std::vector<int> st;
    for (int a = 0; a < na; ++a)
      if (ma[a] < 0) {
        za[a] = 1;
        st.push_back(a);
      }
    while (!st.empty()) {
      const int a = st.back();
      st.pop_back();
      for (int b : e[a])
        if (!zb[b]) {
          zb[b] = 1;
          const int a2 = mb[b];
          if (a2 >= 0 && !za[a2]) {
            za[a2] = 1;
            st.push_back(a2);
          }
        }
    }
    std::vector<int> cover;
    for (int a = 0; a < na; ++a)
      if (!za[a]) cover.push_back(sa[a]);
    for (int b = 0; b < nb; ++b)
      if (zb[b]) cover.push_back(sb[b]);
    for (int i = 0; i < nb; ++i) mpos_[sb[i]] = -1;
    return cover;
  }
  std::vector<int> mpos_;

  int dissect(std::vector<int>& sub) {
    int t = tree_.add();
    if (static_cast<int>(sub.size()) <= kLeafNodes) {
      tree_.nodes[t] = sub;
      return t;
    }
    // principal axes of the subset
    double mean[3] = {0, 0, 0};
    for (int v : sub)
      for (int k = 0; k < 3; ++k) mean[k] += xyz_[3 * v + k];
    for (double& m : mean) m /= sub.size();
    double cov[3][3] = {{0}};
    for (int v : sub)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          cov[i][j] += (xyz_[3 * v + i] - mean[i]) * (xyz_[3 * v + j] - mean[j]);
    double evec[3][3];
    jacobi3(cov, evec);
    double pca[9];
    for (int e = 0; e < 3; ++e)
      for (int k = 0; k < 3; ++k) pca[3 * e + k] = evec[k][e];

    // Candidate cuts: the subset ordered along each of eight directions (x, y, z,
    // radius, angle, three principal axes) and cut at 40 %, 42.5 %, ..., 60 %;
    // the separator is a minimum vertex cover of the cut edges.  The cut kept
    // minimises |S| n^2 / (4 |A| |B|): the separator size, weighed against the
    // imbalance of the two parts.
    const int n = static_cast<int>(sub.size());
    std::vector<std::pair<double, int>> key(n);
    std::vector<int> best_sep, best_a, best_b;
    double best = 1e300;
    for (int dir = 0; dir < 8; ++dir) {
      for (int i = 0; i < n; ++i) key[i] = {proj(sub[i], dir, pca), sub[i]};
      std::sort(key.begin(), key.end());
      for (int f = 0; f <= 8; ++f) {
        const int half = std::max(1, std::min(n - 1, (int)std::lround((0.4 + 0.025 * f) * n)));
        ++cur_;
        for (int i = 0; i < n; ++i) {
          stamp_[key[i].second] = cur_;
          side_[key[i].second] = i < half ? 0 : 1;
        }
        std::vector<int> sa, sb;  // the boundary nodes of each side
        for (int i = 0; i < n; ++i) {
          const int v = key[i].second;
          for (int u : adj_[v])
            if (stamp_[u] == cur_ && side_[u] != side_[v]) {
              (side_[v] == 0 ? sa : sb).push_back(v);
              break;
            }
        }
        std::vector<int> sep = min_cover(sa, sb);
        int in_a = 0;
        for (int v : sep) in_a += side_[v] == 0;
        const double na = half - in_a, nb = (n - half) - ((double)sep.size() - in_a);
        const double score = sep.size() * (double)n * n / std::max(1.0, 4.0 * na * nb);
        if (score < best) {
          best = score;
          best_sep = sep;
          best_a.clear();
          best_b.clear();
          ++cur_;
          for (int v : sep) stamp_[v] = cur_;
          for (int i = 0; i < n; ++i) {
            const int v = key[i].second;
            if (stamp_[v] == cur_) continue;
            (i < half ? best_a : best_b).push_back(v);
          }
        }
      }
    }
    if (best_sep.empty()) {  // disconnected halves: any single node separates
      std::vector<int>& big = best_a.size() >= best_b.size() ? best_a : best_b;
      best_sep.push_back(big.back());
      big.pop_back();
    }
    std::sort(best_sep.begin(), best_sep.end());
    std::sort(best_a.begin(), best_a.end());
    std::sort(best_b.begin(), best_b.end());
    std::vector<int>().swap(sub);
    int ca = best_a.empty() ? -1 : dissect(best_a);
    int cb = best_b.empty() ? -1 : dissect(best_b);
    tree_.c0[t] = ca;
    tree_.c1[t] = cb;
    tree_.nodes[t] = std::move(best_sep);
    return t;
  }
};

}  // namespace

std::vector<char> shared_fronts(const Symbolic& S) {
  const int nsn = static_cast<int>(S.sn.size());
  std::vector<char> shared(nsn, 1);
  auto phase_slot = [&](int sl) { return S.kp_ptr[sl + 1] > S.kp_ptr[sl]; };
  for (const auto& lvl : S.levels)  // leaves first
    for (int id : lvl) {
      const Snode& s = S.sn[id];
      for (int q = s.asm0; q < s.asm1 && shared[id]; ++q)
        if (phase_slot(S.asm_slot[q]) || phase_slot(S.tslot[S.asm_slot[q]])) shared[id] = 0;
      for (int c : s.child)
        if (c >= 0 && !shared[c]) shared[id] = 0;
    }
  return shared;
}

void build_symbolic(Symbolic& S) {
  const int nn = S.nn;
  if (S.axis == 0) { S.c1 = 1; S.c2 = 2; S.cax = 0; }
  else if (S.axis == 1) { S.c1 = 0; S.c2 = 2; S.cax = 1; }
  else { S.c1 = 0; S.c2 = 1; S.cax = 2; }

  // ---- reduced / free nodes (temporary numbering) --------------------------
  std::vector<int> tmp_of(nn, -1);
  std::vector<int> tmp_to_full;
  S.n_free_dofs = 0;
  S.n_reduced_nodes = 0;
  for (int i = 0; i < nn; ++i) {
    if (S.node_class[i] == TM_NODE_RIGHT) continue;
    ++S.n_reduced_nodes;
    int nc = S.cmask[3 * i] + S.cmask[3 * i + 1] + S.cmask[3 * i + 2];
    S.n_free_dofs += 3 - nc;
    if (nc == 3) continue;
    tmp_of[i] = static_cast<int>(tmp_to_full.size());
    tmp_to_full.push_back(i);
  }
  for (int i = 0; i < nn; ++i) {
    if (S.node_class[i] != TM_NODE_RIGHT) continue;
    int l = S.partner[i];
    if (l < 0 || l >= nn || S.node_class[l] == TM_NODE_RIGHT)
      throw std::runtime_error("Right boundary node " + std::to_string(i) +
                               " has no matched left partner");
    tmp_of[i] = tmp_of[l];
  }
  const int nf = static_cast<int>(tmp_to_full.size());
  S.nf = nf;
  if (nf == 0) throw std::runtime_error("no free DOFs in the cyclic-reduced system");

  // ---- free-node graph ------------------------------------------------------
  std::vector<std::vector<int>> adj(nf);
  for (int e = 0; e < S.ne; ++e) {
    const int* el = &S.elements[10 * e];
    int f[10];
    for (int a = 0; a < 10; ++a) f[a] = tmp_of[el[a]];
    for (int a = 0; a < 10; ++a) {
      if (f[a] < 0) continue;
      for (int b = 0; b < 10; ++b)
        if (f[b] >= 0 && f[b] != f[a]) adj[f[a]].push_back(f[b]);
    }
  }
  for (auto& a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  std::vector<double> txyz(3 * nf);
  for (int t = 0; t < nf; ++t)
    for (int k = 0; k < 3; ++k) txyz[3 * t + k] = S.coords[3 * tmp_to_full[t] + k];

  // ---- nested dissection -> postorder supernodes ----------------------------
  NestedDissection nd(adj, txyz, S.c1, S.c2, S.cax);
  NDTree tree = nd.run();
  int root = nd.root();
  std::vector<int> perm(nf, -1);  // tmp -> permuted
  S.sn.clear();
  std::vector<int> tree_to_sn(tree.nodes.size(), -1);
  {
    // iterative postorder
    std::vector<std::pair<int, int>> st;
    st.push_back({root, 0});
    int next = 0;
    while (!st.empty()) {
      auto& [t, state] = st.back();
      if (state == 0) {
        state = 1;
        if (tree.c0[t] >= 0) st.push_back({tree.c0[t], 0});
        continue;
      }
      if (state == 1) {
        state = 2;
        if (tree.c1[t] >= 0) st.push_back({tree.c1[t], 0});
        continue;
      }
      Snode s;
      s.col0 = next;
      s.ncol = static_cast<int>(tree.nodes[t].size());
      for (int v : tree.nodes[t]) perm[v] = next++;
      int id = static_cast<int>(S.sn.size());
      tree_to_sn[t] = id;
      int k = 0;
      for (int c : {tree.c0[t], tree.c1[t]})
        if (c >= 0) {
          s.child[k++] = tree_to_sn[c];
          S.sn[tree_to_sn[c]].parent = id;
        }
      S.sn.push_back(s);
      st.pop_back();
    }
  }
  S.free_to_full.assign(nf, -1);
  for (int t = 0; t < nf; ++t) S.free_to_full[perm[t]] = tmp_to_full[t];
  S.free_of_full.assign(nn, -1);
  for (int i = 0; i < nn; ++i)
    if (tmp_of[i] >= 0) S.free_of_full[i] = perm[tmp_of[i]];
  S.pinned.assign(3 * nf, 0);
  for (int p = 0; p < nf; ++p)
    for (int c = 0; c < 3; ++c) S.pinned[3 * p + c] = S.cmask[3 * S.free_to_full[p] + c];

  // ---- node-block pattern in permuted numbering -----------------------------
  std::vector<std::vector<int>> padj(nf);
  for (int t = 0; t < nf; ++t) {
    auto& row = padj[perm[t]];
    row.reserve(adj[t].size() + 1);
    row.push_back(perm[t]);
    for (int u : adj[t]) row.push_back(perm[u]);
    std::sort(row.begin(), row.end());
  }
  S.rowptr.assign(nf + 1, 0);
  for (int p = 0; p < nf; ++p) S.rowptr[p + 1] = S.rowptr[p] + static_cast<int>(padj[p].size());
  S.colidx.resize(S.rowptr[nf]);
  for (int p = 0; p < nf; ++p)
    std::copy(padj[p].begin(), padj[p].end(), S.colidx.begin() + S.rowptr[p]);
  auto slot_of = [&](int I, int J) -> int {
    auto b = S.colidx.begin() + S.rowptr[I], e = S.colidx.begin() + S.rowptr[I + 1];
    auto it = std::lower_bound(b, e, J);
    if (it == e || *it != J) return -1;
    return static_cast<int>(it - S.colidx.begin());
  };
  const int nblk = S.rowptr[nf];
  S.tslot.resize(nblk);
  for (int I = 0; I < nf; ++I)
    for (int s = S.rowptr[I]; s < S.rowptr[I + 1]; ++s) S.tslot[s] = slot_of(S.colidx[s], I);

  // ---- segmented assembly lists -------------------------------------------
  {
    std::vector<int> ckc(nblk + 1, 0), ckp(nblk + 1, 0);
    auto visit = [&](auto&& fn) {
      for (int e = 0; e < S.ne; ++e) {
        const int* el = &S.elements[10 * e];
        for (int a = 0; a < 10; ++a) {
          int i = el[a], I = S.free_of_full[i];
          if (I < 0) continue;
          bool ri = S.node_class[i] == TM_NODE_RIGHT;
          for (int b = 0; b < 10; ++b) {
            int j = el[b], J = S.free_of_full[j];
            if (J < 0) continue;
            bool rj = S.node_class[j] == TM_NODE_RIGHT;
            if (ri && !rj) continue;  // implied by the adjoint term
            int s = slot_of(I, J);
            int code = (e * 100 + a * 10 + b) * 4 + (ri ? 1 : 0) + (rj ? 2 : 0);
            fn(ri == rj, s, code);
          }
        }
      }
    };
    visit([&](bool same, int s, int) { (same ? ckc : ckp)[s + 1]++; });
    for (int s = 0; s < nblk; ++s) { ckc[s + 1] += ckc[s]; ckp[s + 1] += ckp[s]; }
    S.kc_ptr = ckc;
    S.kp_ptr = ckp;
    S.kc_src.resize(ckc[nblk]);
    S.kp_src.resize(ckp[nblk]);
    std::vector<int> fc(ckc.begin(), ckc.end() - 1), fp(ckp.begin(), ckp.end() - 1);
    visit([&](bool same, int s, int code) {
      if (same) S.kc_src[fc[s]++] = code;
      else S.kp_src[fp[s]++] = code;
    });
    std::vector<int> cf(nf + 1, 0);
    for (int e = 0; e < S.ne; ++e)
      for (int a = 0; a < 10; ++a) {
        int I = S.free_of_full[S.elements[10 * e + a]];
        if (I >= 0) cf[I + 1]++;
      }
    for (int p = 0; p < nf; ++p) cf[p + 1] += cf[p];
    S.f_ptr = cf;
    S.f_src.resize(cf[nf]);
    std::vector<int> ff(cf.begin(), cf.end() - 1);
    for (int e = 0; e < S.ne; ++e)
      for (int a = 0; a < 10; ++a) {
        int i = S.elements[10 * e + a], I = S.free_of_full[i];
        if (I < 0) continue;
        S.f_src[ff[I]++] = (e * 10 + a) * 2 + (S.node_class[i] == TM_NODE_RIGHT ? 1 : 0);
      }
  }

  // ---- fronts: struct, relmap, assembly lists, levels -------------------------
  const int nsn = static_cast<int>(S.sn.size());
  S.struct_nodes.clear();
  S.relmap.clear();
  S.asm_slot.clear();
  S.asm_pos.clear();
  std::vector<int> mark(nf, -1);
  std::vector<std::vector<int>> structs(nsn);
  for (int id = 0; id < nsn; ++id) {
    Snode& s = S.sn[id];
    const int lim = s.col0 + s.ncol;
    std::vector<int>& st = structs[id];
    for (int v = s.col0; v < lim; ++v)
      for (int q = S.rowptr[v]; q < S.rowptr[v + 1]; ++q) {
        int u = S.colidx[q];
        if (u >= lim && mark[u] != id) { mark[u] = id; st.push_back(u); }
      }
    for (int c : s.child)
      if (c >= 0)
        for (int u : structs[c])
          if (u >= lim && mark[u] != id) { mark[u] = id; st.push_back(u); }
    std::sort(st.begin(), st.end());
    s.nstr = static_cast<int>(st.size());
    s.height = 0;
    for (int c : s.child)
      if (c >= 0) s.height = std::max(s.height, S.sn[c].height + 1);
  }
  int64_t foff = 0, toff = 0;
  S.nnz_L = 0;
  S.flops = 0;
  S.max_front = 0;
  int max_h = 0;
  for (int id = 0; id < nsn; ++id) {
    Snode& s = S.sn[id];
    s.str0 = static_cast<int>(S.struct_nodes.size());
    S.struct_nodes.insert(S.struct_nodes.end(), structs[id].begin(), structs[id].end());
    const int64_t m = s.m(), w = s.w();
    s.foff = foff;
    foff += m * m;
    s.toff = toff;
    toff += s.r();
    S.nnz_L += m * w - w * (w - 1) / 2;
    for (int64_t j = 0; j < w; ++j) S.flops += 4.0 * double(m - j) * double(m - j);
    S.max_front = std::max<int>(S.max_front, static_cast<int>(m));
    max_h = std::max(max_h, s.height);
  }
  S.front_entries = foff;
  S.tvec_entries = toff;
  // relmap: struct node of child -> node position in parent's front
  for (int id = 0; id < nsn; ++id) {
    Snode& s = S.sn[id];
    s.rel0 = static_cast<int>(S.relmap.size());
    if (s.parent < 0) {
      if (s.nstr != 0) throw std::logic_error("root front with non-empty struct");
      continue;
    }
    const Snode& p = S.sn[s.parent];
    const int* pst = &S.struct_nodes[p.str0];
    for (int q = 0; q < s.nstr; ++q) {
      int u = S.struct_nodes[s.str0 + q];
      int pos;
      if (u < p.col0 + p.ncol) {
        if (u < p.col0) throw std::logic_error("child struct below parent pivots");
        pos = u - p.col0;
      } else {
        const int* it = std::lower_bound(pst, pst + p.nstr, u);
        if (it == pst + p.nstr || *it != u) throw std::logic_error("child struct not in parent front");
        pos = p.ncol + static_cast<int>(it - pst);
      }
      S.relmap.push_back(pos);
    }
  }
  // assembly lists: entries (I, J), I >= J, J a pivot node of the front
  for (int id = 0; id < nsn; ++id) {
    Snode& s = S.sn[id];
    s.asm0 = static_cast<int>(S.asm_slot.size());
    const int* st = S.struct_nodes.data() + s.str0;
    for (int J = s.col0; J < s.col0 + s.ncol; ++J)
      for (int q = S.rowptr[J]; q < S.rowptr[J + 1]; ++q) {
        int I = S.colidx[q];
        if (I < J) continue;
        int rowpos;
        if (I < s.col0 + s.ncol) rowpos = I - s.col0;
        else rowpos = s.ncol + static_cast<int>(std::lower_bound(st, st + s.nstr, I) - st);
        S.asm_slot.push_back(S.tslot[q]);
        S.asm_pos.push_back(rowpos * 65536 + (J - s.col0));
      }
    s.asm1 = static_cast<int>(S.asm_slot.size());
  }
  S.levels.assign(max_h + 1, {});
  for (int id = 0; id < nsn; ++id) S.levels[S.sn[id].height].push_back(id);

  // ---- reference free-DOF ordering -> system index ---------------------------
  S.ref_free_to_sys.clear();
  for (int pass = 0; pass < 2; ++pass)
    for (int i = 0; i < nn; ++i) {
      int cls = S.node_class[i];
      if (cls == TM_NODE_RIGHT) continue;
      if ((pass == 0) != (cls == TM_NODE_INTERIOR)) continue;
      for (int c = 0; c < 3; ++c) {
        if (S.cmask[3 * i + c]) continue;
        S.ref_free_to_sys.push_back(3 * S.free_of_full[i] + c);
      }
    }
}

}  // namespace tme
