// Host-side checker of the symbolic analysis (turbomodal_amd/csrc/symbolic.cpp),
// test infrastructure only.  Reads a cyclic sector mesh from a binary file
// (int32 n_nodes, n_elems, num_sectors, axis; float64 coords[3 n]; int32
// elements[10 e]; int32 node_class[n]; int32 right_partner[n]; uint8
// constrained[3 n]), runs build_symbolic, and re-derives from the K(k) block
// pattern, independently of the analysis' own bookkeeping:
//   * the supernodes' pivot ranges tile the free nodes in postorder;
//   * nested dissection separates: no pattern edge joins the subtrees of the two
//     children of any front;
//   * every front's structure is that of symbolic elimination: the later nodes
//     adjacent to its pivots or in its children's structures;
//   * nnz(L) and the factor flop count from those structures;
//   * the shared fronts: a front is shared exactly when no slot of the pattern whose
//     earlier node lies in its subtree's node range has phase contributions (its own
//     or its transpose's), compared with shared_fronts().
// Prints one JSON line.
#include <algorithm>
#include <cstdio>
#include <set>
#include <vector>

#include "engine.hpp"

using namespace tme;

int main(int argc, char** argv) {
  if (argc != 2) return 2;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int hdr[4];
  bool ok_read = std::fread(hdr, 4, 4, f) == 4;
  Symbolic S;
  S.nn = hdr[0];
  S.ne = hdr[1];
  S.N = hdr[2];
  S.axis = hdr[3];
  S.coords.resize(3 * (size_t)S.nn);
  S.elements.resize(10 * (size_t)S.ne);
  S.node_class.resize(S.nn);
  S.partner.resize(S.nn);
  S.cmask.resize(3 * (size_t)S.nn);
  ok_read = ok_read && std::fread(S.coords.data(), 8, S.coords.size(), f) == S.coords.size();
  ok_read = ok_read && std::fread(S.elements.data(), 4, S.elements.size(), f) == S.elements.size();
  ok_read = ok_read && std::fread(S.node_class.data(), 4, S.nn, f) == (size_t)S.nn;
  ok_read = ok_read && std::fread(S.partner.data(), 4, S.nn, f) == (size_t)S.nn;
  ok_read = ok_read && std::fread(S.cmask.data(), 1, S.cmask.size(), f) == S.cmask.size();
  std::fclose(f);
  if (!ok_read) return 2;
  build_symbolic(S);
  const int nsn = (int)S.sn.size(), nf = S.nf;

  // pivot ranges tile [0, nf) in supernode order (postorder)
  bool tile_ok = true;
  int next = 0;
  for (const Snode& s : S.sn) {
    tile_ok = tile_ok && s.col0 == next && s.ncol > 0;
    next = s.col0 + s.ncol;
  }
  tile_ok = tile_ok && next == nf;
  // subtree node range of each front: [lo, col0 + ncol), children before parents
  std::vector<int> lo(nsn);
  bool tree_ok = true;
  for (int id = 0; id < nsn; ++id) {
    const Snode& s = S.sn[id];
    lo[id] = s.col0;
    for (int c : s.child)
      if (c >= 0) {
        tree_ok = tree_ok && c < id && S.sn[c].parent == id;
        lo[id] = std::min(lo[id], lo[c]);
      }
  }
  for (int id = 0; id < nsn; ++id) {  // the subtree range is contiguous: children adjacent, in order
    const Snode& s = S.sn[id];
    int end = lo[id];
    for (int c : s.child)
      if (c >= 0) {
        tree_ok = tree_ok && lo[c] == end;
        end = S.sn[c].col0 + S.sn[c].ncol;
      }
    tree_ok = tree_ok && end == s.col0;
  }
  // separators: no pattern edge between the two children's subtrees
  bool sep_ok = true;
  long long sep_checked = 0;
  for (int id = 0; id < nsn && sep_ok; ++id) {
    const Snode& s = S.sn[id];
    if (s.child[0] < 0 || s.child[1] < 0) continue;
    const int a0 = lo[s.child[0]], a1 = S.sn[s.child[0]].col0 + S.sn[s.child[0]].ncol;
    const int b0 = lo[s.child[1]], b1 = S.sn[s.child[1]].col0 + S.sn[s.child[1]].ncol;
    for (int u = a0; u < a1 && sep_ok; ++u)
      for (int q = S.rowptr[u]; q < S.rowptr[u + 1]; ++q) {
        const int v = S.colidx[q];
        ++sep_checked;
        if (v >= b0 && v < b1) sep_ok = false;
      }
  }
  // structures by symbolic elimination, compared with the analysis'
  std::vector<std::vector<int>> st(nsn);
  bool struct_ok = true;
  long long nnz = 0;
  double flops = 0;
  for (int id = 0; id < nsn; ++id) {
    const Snode& s = S.sn[id];
    const int hi = s.col0 + s.ncol;
    std::set<int> set;
    for (int u = s.col0; u < hi; ++u)
      for (int q = S.rowptr[u]; q < S.rowptr[u + 1]; ++q)
        if (S.colidx[q] >= hi) set.insert(S.colidx[q]);
    for (int c : s.child)
      if (c >= 0)
        for (int v : st[c])
          if (v >= hi) set.insert(v);
    st[id].assign(set.begin(), set.end());
    std::vector<int> got(S.struct_nodes.begin() + s.str0, S.struct_nodes.begin() + s.str0 + s.nstr);
    struct_ok = struct_ok && got == st[id];
    const long long w = 3LL * s.ncol, m = w + 3LL * (long long)st[id].size();
    nnz += m * w - w * (w - 1) / 2;
    for (long long j = 0; j < w; ++j) flops += 4.0 * double(m - j) * double(m - j);
  }
  const bool root_ok = nsn > 0 && S.sn[nsn - 1].parent < 0 && st[nsn - 1].empty();
  // shared fronts from the node ranges: entry (u, v) of the pattern is assembled in the
  // front that pivots min(u, v); a subtree is phase-free when none of its nodes is the
  // earlier node of a phase-coupled slot
  std::vector<char> phase_node(nf, 0);
  for (int u = 0; u < nf; ++u)
    for (int q = S.rowptr[u]; q < S.rowptr[u + 1]; ++q) {
      const int v = S.colidx[q], t = S.tslot[q];
      if (S.kp_ptr[q + 1] > S.kp_ptr[q] || S.kp_ptr[t + 1] > S.kp_ptr[t]) phase_node[std::min(u, v)] = 1;
    }
  const std::vector<char> sh = shared_fronts(S);
  bool shared_ok = (int)sh.size() == nsn;
  int n_shared = 0;
  double flops_shared = 0;
  for (int id = 0; id < nsn && shared_ok; ++id) {
    const Snode& s = S.sn[id];
    bool free = true;
    for (int u = lo[id]; u < s.col0 + s.ncol && free; ++u) free = !phase_node[u];
    shared_ok = shared_ok && (sh[id] != 0) == free;
    if (free) {
      ++n_shared;
      const long long w = 3LL * s.ncol, m = w + 3LL * s.nstr;
      for (long long j = 0; j < w; ++j) flops_shared += 4.0 * double(m - j) * double(m - j);
    }
  }
  std::printf(
      "{\"nf\": %d, \"n_free\": %d, \"nsn\": %d, \"max_front\": %d, \"nnz_L\": %lld, \"flops\": %.6e, "
      "\"nnz_L_recomputed\": %lld, \"flops_recomputed\": %.6e, \"tile_ok\": %s, \"tree_ok\": %s, \"sep_ok\": %s, "
      "\"sep_edges_checked\": %lld, \"struct_ok\": %s, \"root_ok\": %s, \"shared_ok\": %s, \"n_shared\": %d, "
      "\"flops_shared\": %.6e}\n",
      nf, S.n_free_dofs, nsn, S.max_front, (long long)S.nnz_L, S.flops, nnz, flops, tile_ok ? "true" : "false",
      tree_ok ? "true" : "false", sep_ok ? "true" : "false", sep_checked, struct_ok ? "true" : "false",
      root_ok ? "true" : "false", shared_ok ? "true" : "false", n_shared, flops_shared);
  return 0;
}
