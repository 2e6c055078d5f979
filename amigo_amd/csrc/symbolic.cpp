// Host symbolic analysis (see symbolic.h).  Every routine that must agree
// bit for bit with the reference cites the reference lines it restates; the
// storage moves of the quotient graph are kept identical on purpose because
// the reference's tie-breaking reads slack slots of that storage.
#include "symbolic.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <queue>

namespace amigo_amd {

namespace {

// Sort, drop negative entries and duplicates; returns the new length.
// (block_amd.h:459-478)
int sort_unique_nonneg(int len, int* a) {
  std::sort(a, a + len);
  int* first = std::lower_bound(a, a + len, 0);
  int* last = std::unique(first, a + len);
  int m = int(last - first);
  if (first != a) std::copy(first, last, a);
  return m;
}

// Rows must be strictly increasing in both the element and variable parts
// (block_amd.h:487-514).
int check_rows(int n, const int* rowp, const int* cols, const int* elen,
               const int* alen) {
  for (int i = 0; i < n; i++) {
    for (int j = rowp[i]; j < rowp[i] + elen[i] - 1; j++)
      if (cols[j] < 0 || cols[j + 1] < 0 || cols[j + 1] <= cols[j]) return i + 1;
    for (int j = rowp[i] + elen[i]; j < rowp[i] + alen[i] - 1; j++)
      if (cols[j] < 0 || cols[j + 1] < 0 || cols[j + 1] <= cols[j]) return i + 1;
  }
  return 0;
}

// Make `need` free slots at the end of row r by compacting neighbouring rows:
// first the rows before r (walking down), then the rows after it (walking
// up).  (block_amd.h:532-581)
void make_room(int n, int* rowp, int* cols, const int* alen, int r, int need) {
  int got = 0;
  if (r > 0) {
    int s = r - 1;
    while (got < need && s >= 0) {
      got += (rowp[s + 1] - rowp[s]) - alen[s];
      s--;
    }
    s++;
    int dst = rowp[s] + alen[s];
    for (int i = s + 1; i <= r; i++) {
      int src = rowp[i];
      int stop = src + alen[i];
      int begin = dst;
      while (src < stop) cols[dst++] = cols[src++];
      rowp[i] = begin;
    }
  }
  if (got < need) {
    int s = r + 1;
    while (got < need && s < n) {
      got += (rowp[s + 1] - rowp[s]) - alen[s];
      s++;
    }
    s--;
    if (s >= n) {
      s = n - 1;
      std::fprintf(stderr, "amigo_amd: quotient graph out of space\n");
    }
    int dst = rowp[s + 1] - 1;
    for (int i = s; i > r; i--) {
      int lo = rowp[i];
      for (int src = lo + alen[i] - 1; src >= lo; src--) cols[dst--] = cols[src];
      rowp[i] = rowp[i + 1] - alen[i];
    }
  }
}

// Variables a and b are indistinguishable when {adj(a), a} == {adj(b), b}.
// The scan tolerates one skip per side exactly like block_amd.h:603-630.
bool indistinguishable(int a, int b, const int* elen, const int* alen,
                       const int* rowp, const int* cols) {
  if (a == b) return false;
  if (alen[a] != alen[b] || elen[a] != elen[b]) return false;
  int pa = rowp[a], ea = pa + alen[a];
  int pb = rowp[b], eb = pb + alen[b];
  for (int t = 0; t < alen[a]; t++, pa++, pb++) {
    if (cols[pa] == b) pa++;
    if (cols[pb] == a) pb++;
    if (cols[pa] != cols[pb]) break;
  }
  return pa == ea && pb == eb;
}

// Detach variable v from every element and variable row that lists it
// (block_amd.h:637-678).
void detach_variable(int v, int* elen, int* alen, const int* rowp, int* cols) {
  int i = rowp[v];
  for (; i < rowp[v] + elen[v]; i++) {
    int e = cols[i];
    int j = rowp[e], last = rowp[e] + alen[e] - 1;
    while (j < last && cols[j] != v) j++;
    for (; j < last; j++) cols[j] = cols[j + 1];
    alen[e]--;
  }
  for (; i < rowp[v] + alen[v]; i++) {
    int u = cols[i];
    int j = rowp[u] + elen[u], last = rowp[u] + alen[u] - 1;
    while (j < last && cols[j] != v) j++;
    for (; j < last; j++) cols[j] = cols[j + 1];
    alen[u]--;
  }
  elen[v] = 0;
  alen[v] = 0;
}

// Degree buckets with LIFO insertion (block_amd.h:687-812).  min_deg only
// ever decreases, as in the reference.
class DegreeBuckets {
 public:
  DegreeBuckets(int n, const int* rowp, const int* is_mult)
      : n_(n),
        is_mult_(is_mult),
        deg_(n),
        head_(n + 1, -1),
        next_(n + 1, -1),
        prev_(n + 1, -1) {
    min_deg_ = 0;
    for (int i = 0; i < n; i++) {
      deg_[i] = rowp[i + 1] - rowp[i];
      if (i == 0 || deg_[i] < min_deg_) min_deg_ = deg_[i];
    }
    for (int i = 0; i < n; i++) push(i, deg_[i]);
  }
  int degree(int i) const { return deg_[i]; }
  void remove(int i) {
    int d = deg_[i];
    if (head_[d] == i) head_[d] = next_[i];
    if (next_[i] != -1) prev_[next_[i]] = prev_[i];
    if (prev_[i] != -1) next_[prev_[i]] = next_[i];
    prev_[i] = next_[i] = -1;
  }
  void set_degree(int i, int d) {
    if (d < min_deg_) min_deg_ = d;
    remove(i);
    push(i, d);
    deg_[i] = d;
  }
  int pick(bool skip_mult) const {
    for (int d = min_deg_; d <= n_; d++) {
      int h = head_[d];
      if (h != -1 && (!skip_mult || !is_mult_[h])) return h;
    }
    return -1;
  }

 private:
  void push(int i, int d) {
    if (head_[d] != -1) {
      prev_[head_[d]] = i;
      next_[i] = head_[d];
    }
    head_[d] = i;
  }
  int n_;
  const int* is_mult_;
  int min_deg_;
  std::vector<int> deg_, head_, next_, prev_;
};

}  // namespace

// Approximate minimum degree (block_amd.h:27-452).
int block_amd(int n, int* rowp, int* cols, int nmult, int* mult, int* perm) {
  std::vector<int> alen(n), elen(n, 0), slen(n, 1), edeg(n, -1), state(n, 1);
  std::vector<int> Lp(n + 1);
  for (int i = 0; i < n; i++) {
    perm[i] = i;
    alen[i] = rowp[i + 1] - rowp[i];
  }
  if (check_rows(n, rowp, cols, elen.data(), alen.data()) != 0) return 1;

  nmult = sort_unique_nonneg(nmult, mult);
  std::vector<int> is_mult(n, 0);
  for (int i = 0; i < nmult; i++) is_mult[mult[i]] = 1;

  DegreeBuckets buckets(n, rowp, is_mult.data());

  int nsv = 0;
  for (int i = 0; i < n; nsv++) {
    const bool early = i < n - nmult;
    const int piv = buckets.pick(early);
    buckets.remove(piv);
    perm[nsv] = piv;
    state[piv] = 0;

    // Lp = {piv} U (variables adjacent to piv) U (variables of piv's
    // elements), sorted and unique.
    int len = 0;
    Lp[len++] = piv;
    for (int j = rowp[piv] + elen[piv]; j < rowp[piv] + alen[piv]; j++)
      if (cols[j] != piv) Lp[len++] = cols[j];
    for (int j = rowp[piv]; j < rowp[piv] + elen[piv]; j++) {
      const int e = cols[j];
      for (int k = rowp[e] + elen[e]; k < rowp[e] + alen[e]; k++) {
        if (len >= n) len = sort_unique_nonneg(len, Lp.data());
        if (cols[k] != piv) Lp[len++] = cols[k];
      }
    }
    len = sort_unique_nonneg(len, Lp.data());

    // Absorb piv's elements and drop Lp from every other row of Lp.
    for (int j = 0; j < len; j++) {
      const int v = Lp[j];
      if (v == piv) continue;
      int nre = 0;
      const int pend = rowp[piv] + elen[piv];
      for (int k = rowp[piv], c = rowp[v]; k < pend && c < rowp[v] + elen[v];
           c++) {
        while (k < pend && cols[k] < cols[c]) k++;
        if (k < pend && cols[k] == cols[c]) {
          cols[c] = -1;
          nre++;
        }
      }
      int nrv = 0;
      for (int k = 0, c = rowp[v] + elen[v]; k < len && c < rowp[v] + alen[v];
           c++) {
        while (k < len && Lp[k] < cols[c]) k++;
        if (k < len && cols[c] == Lp[k]) {
          cols[c] = -1;
          nrv++;
        }
      }
      if (nre > 0 || nrv > 0) {
        // In-place compaction; the slots past the new end keep their old
        // contents, which later scans may read.
        const int end = rowp[v] + alen[v];
        int k = rowp[v], w = k;
        for (; k < end; k++, w++) {
          while (k < end && cols[k] == -1) k++;
          if (k < end) cols[w] = cols[k];
        }
        elen[v] -= nre;
        alen[v] -= nre + nrv;
      }
    }

    // Insert piv as a new element into each row of Lp \ {piv}.
    for (int j = 0; j < len; j++) {
      const int v = Lp[j];
      if (v == piv) continue;
      if (alen[v] == rowp[v + 1] - rowp[v])
        make_room(n, rowp, cols, alen.data(), v, 1);
      int k = rowp[v];
      const int end = k + elen[v];
      while (k < end && cols[k] < piv) k++;
      if (cols[k] != piv) {
        for (int p = rowp[v] + alen[v] - 1; p >= k; p--) cols[p + 1] = cols[p];
        cols[k] = piv;
        alen[v]++;
        elen[v]++;
      }
    }

    // piv's old elements are absorbed.
    for (int j = rowp[piv]; j < rowp[piv] + elen[piv]; j++) {
      const int e = cols[j];
      alen[e] = 0;
      elen[e] = 0;
      state[e] = -1;
    }

    // Row piv becomes the element Lp \ {piv}.
    const int cap = rowp[piv + 1] - rowp[piv];
    if (len - 1 > cap)
      make_room(n, rowp, cols, alen.data(), piv,
                len - 1 - (rowp[piv + 1] - rowp[piv]));
    elen[piv] = 0;
    alen[piv] = len - 1;
    for (int j = rowp[piv], k = 0; k < len; k++, j++) {
      if (Lp[k] == piv) {
        k++;
        if (k == len) break;
      }
      cols[j] = Lp[k];
    }

    // Approximate external degrees (block_amd.h:275-374).
    for (int j = 0; j < len; j++) {
      const int v = Lp[j];
      for (int k = rowp[v]; k < rowp[v] + elen[v]; k++) {
        const int e = cols[k];
        if (edeg[e] < 0) {
          edeg[e] = 0;
          for (int p = rowp[e]; p < rowp[e] + alen[e]; p++) edeg[e] += slen[cols[p]];
        }
        edeg[e] -= slen[v];
      }
    }
    int deg_lp = 0;
    for (int j = 0; j < len; j++)
      if (Lp[j] != piv) deg_lp += slen[Lp[j]];

    for (int j = 0; j < len; j++) {
      const int v = Lp[j];
      if (v == piv) continue;
      int est = deg_lp;
      for (int k = rowp[v] + elen[v]; k < rowp[v] + alen[v]; k++) est += slen[cols[k]];
      const int vstart = rowp[v] + elen[v];
      const int vlen = alen[v] - elen[v];
      if (vlen > 0 && std::binary_search(cols + vstart, cols + vstart + vlen, v))
        est -= slen[v];
      if (std::binary_search(Lp.data(), Lp.data() + len, v)) est -= slen[v];
      for (int k = rowp[v]; k < rowp[v] + elen[v]; k++) {
        const int e = cols[k];
        if (e == piv) continue;
        if (edeg[e] < 0) {
          edeg[e] = 0;
          for (int p = rowp[e]; p < rowp[e] + alen[e]; p++) edeg[e] += slen[cols[p]];
        }
        est += edeg[e];
      }
      est = std::min(est, n - i);
      const int d = buckets.degree(v);
      buckets.set_degree(v, std::min(est, d + deg_lp));
    }
    for (int j = 0; j < len; j++) {
      const int v = Lp[j];
      for (int k = rowp[v]; k < rowp[v] + elen[v]; k++) edeg[cols[k]] = -1;
    }

    // Supervariable detection (block_amd.h:377-408).
    for (int j = 0; j < len; j++) {
      const int a = Lp[j];
      if (a == piv || slen[a] < 0 || (early && is_mult[a])) continue;
      for (int k = j + 1; k < len; k++) {
        const int b = Lp[k];
        if (b == piv || slen[b] < 0 || (early && is_mult[b])) continue;
        if (indistinguishable(a, b, elen.data(), alen.data(), rowp, cols)) {
          slen[a] += slen[b];
          buckets.set_degree(a, buckets.degree(a) - slen[b]);
          slen[b] = -(a + 1);
          state[b] = -1;
          detach_variable(b, elen.data(), alen.data(), rowp, cols);
          buckets.remove(b);
        }
      }
    }
    i += slen[piv];
  }

  // Expand supervariables: principal variables in pivot order, each
  // followed by the variables merged into it in increasing index order.
  int end = n;
  for (int j = nsv - 1; j >= 0; j--) {
    const int sv = perm[j];
    end -= slen[sv];
    perm[end] = sv;
    state[sv] = end + 1;
  }
  for (int k = 0; k < n; k++) {
    if (slen[k] < 0) {
      int j = k;
      while (slen[j] < 0) j = -(slen[j] + 1);
      perm[state[j]] = k;
      state[j]++;
    }
  }
  return 0;
}

// ordering_utils.h:812-838
static void sort_unique_rows(int n, std::vector<int>& rowp,
                             std::vector<int>& cols) {
  int start = rowp[0];
  for (int i = 0; i < n; i++) {
    int size = rowp[i + 1] - start;
    int* a = cols.data() + start;
    std::sort(a, a + size);
    int m = int(std::unique(a, a + size) - a);
    std::copy(a, a + m, cols.data() + rowp[i]);
    start = rowp[i + 1];
    rowp[i + 1] = rowp[i] + m;
  }
}

// ordering_utils.h:313-360
void copy_for_reorder(int order, int n, const int* rowp, const int* cols,
                      std::vector<int>& rp, std::vector<int>& cl) {
  rp.assign(n + 1, 0);
  if (order == ORDER_NESTED_DISSECTION) {
    cl.resize(rowp[n]);
    int ptr = 0;
    for (int i = 0; i < n; i++) {
      for (int jp = rowp[i]; jp < rowp[i + 1]; jp++)
        if (cols[jp] != i) cl[ptr++] = cols[jp];
      rp[i + 1] = ptr;
    }
  } else {
    cl.resize(rowp[n] + n);
    int ptr = 0;
    for (int i = 0; i < n; i++) {
      cl[ptr++] = i;
      for (int jp = rowp[i]; jp < rowp[i + 1]; jp++) cl[ptr++] = cols[jp];
      rp[i + 1] = ptr;
    }
  }
  sort_unique_rows(n, rp, cl);
}

// ---------------------------------------------------------------------------
// Nested dissection (this build's own; the reference calls METIS_NodeND or
// falls back to the identity, ordering_utils.h:177-197).  Recursive
// bisection by breadth-first level structures from a pseudo-peripheral
// vertex; the middle level (thinned to a vertex separator) is ordered last.
// Parts below `leaf` vertices are ordered by approximate minimum degree.
namespace {

// Path-like pieces (level width <= ND_CHAIN_WIDTH) are bisected down to
// ND_CHAIN_LEAF vertices.  On seeded cart-pole trajectories (N = 16..512
// stages) this cuts the elimination tree from up to 68 levels to 9 and the
// critical path (sum over levels of the largest front's pivots) from about
// 1,800 pivots to about 300, with 40% fewer flops; 2-D and 3-D grids keep
// their ordering (their level structures are wide).
constexpr int ND_CHAIN_WIDTH = 2;
constexpr int ND_CHAIN_LEAF = 2;

struct NDWork {
  int n;
  const int* rowp;
  const int* cols;
  std::vector<int> mark, level, local;
  int stamp = 0;
};

// BFS over the vertices with mark == tag; returns the last vertex visited.
int bfs_levels(NDWork& w, const std::vector<int>& verts, int tag, int root,
               std::vector<int>& order, int& nlev) {
  for (int v : verts) w.level[v] = -1;
  order.clear();
  order.push_back(root);
  w.level[root] = 0;
  nlev = 1;
  for (size_t h = 0; h < order.size(); h++) {
    int v = order[h];
    for (int p = w.rowp[v]; p < w.rowp[v + 1]; p++) {
      int u = w.cols[p];
      if (w.mark[u] == tag && w.level[u] < 0) {
        w.level[u] = w.level[v] + 1;
        nlev = std::max(nlev, w.level[u] + 1);
        order.push_back(u);
      }
    }
  }
  return order.back();
}

void amd_leaf(NDWork& w, const std::vector<int>& verts, int* out) {
  int m = int(verts.size());
  if (m <= 2) {
    for (int i = 0; i < m; i++) out[i] = verts[i];
    return;
  }
  for (int i = 0; i < m; i++) w.local[verts[i]] = i;
  int tag = w.mark[verts[0]];
  std::vector<int> rp(m + 1, 0), cl;
  for (int i = 0; i < m; i++) {
    int v = verts[i];
    cl.push_back(i);
    for (int p = w.rowp[v]; p < w.rowp[v + 1]; p++) {
      int u = w.cols[p];
      if (w.mark[u] == tag && u != v) cl.push_back(w.local[u]);
    }
    rp[i + 1] = int(cl.size());
  }
  sort_unique_rows(m, rp, cl);
  std::vector<int> perm(m);
  block_amd(m, rp.data(), cl.data(), 0, nullptr, perm.data());
  for (int i = 0; i < m; i++) out[i] = verts[perm[i]];
}

void dissect(NDWork& w, std::vector<int> verts, int* out, int leaf) {
  int m = int(verts.size());
  int tag = ++w.stamp;
  for (int v : verts) w.mark[v] = tag;
  if (m <= std::min(leaf, ND_CHAIN_LEAF)) {
    amd_leaf(w, verts, out);
    return;
  }
  // Components first: order each connected component independently.
  std::vector<int> order;
  int nlev = 0;
  int far = bfs_levels(w, verts, tag, verts[0], order, nlev);
  if (m <= leaf && int(order.size()) < m) {
    amd_leaf(w, verts, out);
    return;
  }
  if (int(order.size()) < m) {
    std::vector<int> a = order, b;
    for (int v : verts)
      if (w.level[v] < 0) b.push_back(v);
    int na = int(a.size());
    dissect(w, a, out, leaf);
    dissect(w, b, out + na, leaf);
    return;
  }
  // Pseudo-peripheral root: a few sweeps to the farthest vertex.
  int root = far;
  for (int sweep = 0; sweep < 3; sweep++) {
    int prev_nlev = nlev;
    int nf = bfs_levels(w, verts, tag, root, order, nlev);
    if (nlev <= prev_nlev && sweep > 0) break;
    root = nf;
  }
  bfs_levels(w, verts, tag, root, order, nlev);
  if (nlev < 3) {
    amd_leaf(w, verts, out);
    return;
  }
  // Choose the level that balances the halves.
  std::vector<int> cnt(nlev, 0);
  for (int v : verts) cnt[w.level[v]]++;
  if (m <= leaf) {
    // A piece of leaf size is ordered by minimum degree unless it is
    // path-like (every level of its level structure holds at most
    // ND_CHAIN_WIDTH vertices, e.g. the time stages of a trajectory):
    // minimum degree eliminates a path from its ends inwards, one long
    // chain of supernodes, so such a piece is bisected further.
    const int width = *std::max_element(cnt.begin(), cnt.end());
    if (width > ND_CHAIN_WIDTH) {
      amd_leaf(w, verts, out);
      return;
    }
  }
  int acc = 0, sep = 1;
  for (int l = 0; l < nlev; l++) {
    if (acc + cnt[l] > m / 2) {
      sep = std::max(1, std::min(l, nlev - 2));
      break;
    }
    acc += cnt[l];
  }
  // Separator = vertices of level `sep` adjacent to level sep+1.
  std::vector<int> A, B, S;
  for (int v : verts) {
    int l = w.level[v];
    if (l < sep) {
      A.push_back(v);
    } else if (l > sep) {
      B.push_back(v);
    } else {
      bool touches = false;
      for (int p = w.rowp[v]; p < w.rowp[v + 1] && !touches; p++) {
        int u = w.cols[p];
        if (w.mark[u] == tag && w.level[u] == sep + 1) touches = true;
      }
      (touches ? S : A).push_back(v);
    }
  }
  int na = int(A.size()), nb = int(B.size());
  if (na == 0 || nb == 0) {
    amd_leaf(w, verts, out);
    return;
  }
  dissect(w, A, out, leaf);
  dissect(w, B, out + na, leaf);
  for (size_t i = 0; i < S.size(); i++) out[na + nb + i] = S[i];
}

}  // namespace

void nested_dissection(int n, const int* rowp, const int* cols, int* perm,
                       const int* groups) {
  if (groups != nullptr) {
    // Dissect the quotient graph of the variable groups (e.g. the grid or
    // time nodes of a model), then list each group's variables in increasing
    // index order, those coupled only to their own group last.  A multiplier
    // laid out after its primal partners is then eliminated in the same front
    // as them, where a 2x2 pivot is available; a group-internal variable
    // (a control coupled to its own constraint only) placed last is filled
    // into the clique of the earlier columns, so consecutive groups of a
    // separator chain into one supernode instead of one node per group.
    int G = 0;
    for (int i = 0; i < n; i++) G = std::max(G, groups[i] + 1);
    std::vector<int> gp(G + 1, 0), gcols;
    std::vector<std::vector<int>> adj(G);
    for (int i = 0; i < n; i++)
      for (int p = rowp[i]; p < rowp[i + 1]; p++) {
        int gi = groups[i], gj = groups[cols[p]];
        if (gi != gj) adj[gi].push_back(gj);
      }
    for (int g = 0; g < G; g++) {
      std::sort(adj[g].begin(), adj[g].end());
      adj[g].erase(std::unique(adj[g].begin(), adj[g].end()), adj[g].end());
      gp[g + 1] = gp[g] + int(adj[g].size());
      gcols.insert(gcols.end(), adj[g].begin(), adj[g].end());
    }
    std::vector<int> gperm(G);
    nested_dissection(G, gp.data(), gcols.data(), gperm.data(), nullptr);
    std::vector<int> mptr(G + 1, 0), members(n);
    for (int i = 0; i < n; i++) mptr[groups[i] + 1]++;
    for (int g = 0; g < G; g++) mptr[g + 1] += mptr[g];
    std::vector<int> fill(mptr.begin(), mptr.end() - 1);
    for (int i = 0; i < n; i++) members[fill[groups[i]]++] = i;
    std::vector<char> internal(n, 1);
    for (int i = 0; i < n; i++)
      for (int p = rowp[i]; p < rowp[i + 1]; p++)
        if (groups[cols[p]] != groups[i]) internal[i] = 0;
    int k = 0;
    for (int t = 0; t < G; t++) {
      int g = gperm[t];
      for (int pass = 0; pass < 2; pass++)
        for (int q = mptr[g]; q < mptr[g + 1]; q++)
          if (internal[members[q]] == pass) perm[k++] = members[q];
    }
    return;
  }
  NDWork w;
  w.n = n;
  w.rowp = rowp;
  w.cols = cols;
  w.mark.assign(n, 0);
  w.level.assign(n, -1);
  w.local.assign(n, 0);
  std::vector<int> verts(n);
  std::iota(verts.begin(), verts.end(), 0);
  dissect(w, verts, perm, 256);
}

void compute_ordering(int order, int n, const int* rowp, const int* cols,
                      std::vector<int>& perm, std::vector<int>& iperm,
                      const int* groups) {
  perm.assign(n, 0);
  iperm.assign(n, 0);
  if (order == ORDER_NATURAL || order == ORDER_MULTI_COLOR) {
    std::iota(perm.begin(), perm.end(), 0);
  } else {
    std::vector<int> rp, cl;
    copy_for_reorder(order, n, rowp, cols, rp, cl);
    if (order == ORDER_AMD) {
      block_amd(n, rp.data(), cl.data(), 0, nullptr, perm.data());
    } else {
      nested_dissection(n, rp.data(), cl.data(), perm.data(), groups);
    }
  }
  for (int i = 0; i < n; i++) iperm[perm[i]] = i;
}

// ---------------------------------------------------------------------------
// Elimination tree by path compression on ancestors (sparse_ldl.h:2375-2440).
void etree(int n, const int* colp, const int* rows, const int* perm,
           const int* iperm, int* parent) {
  std::vector<int> anc(n, -1);
  std::fill(parent, parent + n, -1);
  for (int k = 0; k < n; k++) {
    const int ka = perm ? perm[k] : k;
    for (int ip = colp[ka]; ip < colp[ka + 1]; ip++) {
      int i = iperm ? iperm[rows[ip]] : rows[ip];
      while (i < k) {
        const int t = anc[i];
        anc[i] = k;
        if (t == -1) {
          parent[i] = k;
          break;
        }
        i = t;
      }
    }
  }
}

// Depth-first postorder; children visited in increasing index order
// (sparse_ldl.h:2455-2518).
void postorder(int n, const int* parent, int* post) {
  std::vector<int> head(n, -1), next(n, -1), stack(n);
  for (int j = n - 1; j >= 0; j--) {
    if (parent[j] != -1) {
      next[j] = head[parent[j]];
      head[parent[j]] = j;
    }
  }
  int k = 0;
  for (int j = 0; j < n; j++) {
    if (parent[j] != -1) continue;
    int top = 0;
    stack[0] = j;
    while (top >= 0) {
      const int p = stack[top];
      const int c = head[p];
      if (c == -1) {
        post[k++] = p;
        top--;
      } else {
        head[p] = next[c];
        stack[++top] = c;
      }
    }
  }
}

// Column counts of L (including the diagonal) from the row-subtree leaf
// structure (sparse_ldl.h:2552-2701).
void column_counts(int n, const int* colp, const int* rows, const int* post,
                   const int* perm, const int* iperm, const int* parent,
                   int* cc) {
  std::vector<int> first(n, -1), prevleaf(n, -1), maxfirst(n, -1), anc(n);
  std::fill(cc, cc + n, 0);
  for (int k = 0; k < n; k++) {
    int j = post[k];
    if (first[j] == -1) cc[j] = 1;
    while (j != -1 && first[j] == -1) {
      first[j] = k;
      j = parent[j];
    }
  }
  std::iota(anc.begin(), anc.end(), 0);
  auto root_of = [&](int p) {
    int r = p;
    while (anc[r] != r) r = anc[r];
    while (anc[p] != p) {
      int nx = anc[p];
      anc[p] = r;
      p = nx;
    }
    return r;
  };
  for (int k = 0; k < n; k++) {
    const int j = post[k];
    const int ja = perm ? perm[j] : j;
    if (parent[j] != -1) cc[parent[j]]--;
    for (int ip = colp[ja]; ip < colp[ja + 1]; ip++) {
      const int i = iperm ? iperm[rows[ip]] : rows[ip];
      if (i <= j) continue;
      if (first[j] <= maxfirst[i]) continue;
      cc[j]++;
      maxfirst[i] = first[j];
      const int jprev = prevleaf[i];
      prevleaf[i] = j;
      if (jprev != -1) cc[root_of(jprev)]--;
    }
    if (parent[j] != -1) anc[j] = parent[j];
  }
  for (int k = 0; k < n; k++) {
    const int j = post[k];
    if (parent[j] != -1) cc[parent[j]] += cc[j];
  }
}

namespace {

// Fundamental supernodes: chains in the postorder whose column counts drop by
// one per step (sparse_ldl.h:2716-2747).
int fundamental_supernodes(int n, const int* post, const int* parent,
                           const int* cc, int* vtosn, int* sntov) {
  int sn = 0;
  for (int i = 0; i < n; sn++) {
    int var = post[i];
    vtosn[var] = sn;
    sntov[i++] = var;
    int nxt = i < n ? post[i] : n;
    while (i < n && parent[var] == nxt && cc[nxt] == cc[var] - 1) {
      var = nxt;
      vtosn[var] = sn;
      sntov[i++] = var;
      if (i < n) nxt = post[i];
    }
  }
  return sn;
}

// Children lists with the roots stored after the last node
// (sparse_ldl.h:2960-2996).
void child_lists(int K, const int* par, int* cptr, int* kids) {
  std::fill(cptr, cptr + K + 2, 0);
  for (int i = 0; i < K; i++) cptr[(par[i] != -1 ? par[i] : K) + 1]++;
  for (int i = 0; i < K + 1; i++) cptr[i + 1] += cptr[i];
  for (int i = 0; i < K; i++) {
    int p = par[i] != -1 ? par[i] : K;
    kids[cptr[p]++] = i;
  }
  for (int i = K; i >= 0; i--) cptr[i + 1] = cptr[i];
  cptr[0] = 0;
}

// Relaxed supernode amalgamation with a width limit of 64 and at most 256
// explicit zeros added per merge chain (sparse_ldl.h:2768-2949).  Children
// are visited in an order sorted by the width of the root of their *slot
// index* in the child array, exactly as the reference's comparator does.
int relax_supernodes(int n, int ns, const int* sparent, const int* scptr,
                     const int* skids, const int* cc, const int* sptr,
                     const int* stov, std::vector<int>& node_ptr,
                     std::vector<int>& node_to_var,
                     std::vector<int>& var_to_node, std::vector<int>& ncc,
                     std::vector<int>& npar) {
  const int max_width = 64, max_zeros = 256;
  std::vector<int> root(ns), rcc(ns), width(ns), zeros(ns, 0);
  for (int s = 0; s < ns; s++) {
    width[s] = sptr[s + 1] - sptr[s];
    rcc[s] = cc[stov[sptr[s]]];
    root[s] = s;
  }
  auto find = [&](int s) {
    while (root[s] != s) {
      root[s] = root[root[s]];
      s = root[s];
    }
    return s;
  };
  std::vector<int> order;
  for (int s = 0; s < ns; s++) {
    order.clear();
    for (int ip = scptr[s]; ip < scptr[s + 1]; ip++) order.push_back(ip);
    std::sort(order.begin(), order.end(),
              [&](int a, int b) { return width[find(a)] < width[find(b)]; });
    for (int ip : order) {
      const int child = skids[ip];
      const int rp = find(s), rc = find(child);
      const int w = width[rc] + width[rp];
      if (rc != rp && w <= max_width) {
        const int extra = std::max(0, rcc[rp] - (rcc[rc] - width[rp]));
        const int added = extra * width[rc];
        if (zeros[rp] + zeros[rc] + added <= max_zeros) {
          rcc[rp] = w + rcc[rp] - width[rp];
          width[rp] = w;
          root[rc] = rp;
          zeros[rp] += added + zeros[rc];
        }
      }
    }
  }
  std::vector<int> newid(ns, -1);
  int K = 0;
  for (int s = 0; s < ns; s++)
    if (find(s) == s) newid[s] = K++;
  std::vector<int> s2n(ns);
  for (int s = 0; s < ns; s++) s2n[s] = newid[find(s)];

  node_ptr.assign(K + 1, 0);
  node_to_var.assign(n, 0);
  var_to_node.assign(n, 0);
  ncc.assign(K, 0);
  for (int s = 0; s < ns; s++) node_ptr[s2n[s] + 1] += sptr[s + 1] - sptr[s];
  for (int k = 0; k < K; k++) node_ptr[k + 1] += node_ptr[k];
  std::vector<int> fill(node_ptr.begin(), node_ptr.end() - 1);
  for (int s = 0; s < ns; s++) {
    int k = s2n[s];
    for (int ip = sptr[s]; ip < sptr[s + 1]; ip++) {
      node_to_var[fill[k]++] = stov[ip];
      var_to_node[stov[ip]] = k;
    }
  }
  for (int s = 0; s < ns; s++) ncc[s2n[s]] = rcc[find(s)];
  npar.assign(K, -1);
  for (int s = 0; s < ns; s++) {
    if (find(s) != s) continue;
    int ps = sparent[s];
    if (ps >= 0) {
      int kp = s2n[find(ps)];
      if (kp != s2n[s]) npar[s2n[s]] = kp;
    }
  }
  return K;
}

// Contribution rows of every node, found by walking the node tree from each
// off-diagonal entry (sparse_ldl.h:3018-3104).  Rows are recorded in the
// original numbering, in increasing permuted order.
void contribution_rows(int n, const int* colp, const int* rows,
                       const int* perm, const int* iperm, const int* parent,
                       int K, const int* nptr, const int* vtosn,
                       const int* sntov, const int* cptr, int* cvars) {
  std::vector<int> cnt(K, 0), flag(K, -1), snpar(K);
  for (int s = 0; s < K; s++) {
    int last = sntov[nptr[s + 1] - 1];
    int p = parent[last];
    snpar[s] = p >= 0 ? vtosn[p] : -1;
  }
  for (int k = 0; k < n; k++) {
    const int ks = vtosn[k];
    const int ka = perm ? perm[k] : k;
    for (int ip = colp[ka]; ip < colp[ka + 1]; ip++) {
      const int i = iperm ? iperm[rows[ip]] : rows[ip];
      if (i >= k) continue;
      for (int s = vtosn[i]; s >= 0 && s != ks; s = snpar[s]) {
        if (flag[s] == k) break;
        flag[s] = k;
        cvars[cptr[s] + cnt[s]++] = ka;
      }
    }
  }
}

}  // namespace

// SparseLDL::symbolic_analysis (sparse_ldl.h:2204-2358).
void symbolic_analysis(int order, int n, const int* rowp, const int* cols,
                       SymbolicPlan& P, const int* groups, const int* user_perm) {
  P.n = n;
  P.order = order;
  const int* perm = nullptr;
  const int* iperm = nullptr;
  if (user_perm) {
    P.perm.assign(user_perm, user_perm + n);
    P.iperm.assign(n, 0);
    for (int i = 0; i < n; i++) P.iperm[P.perm[i]] = i;
    perm = P.perm.data();
    iperm = P.iperm.data();
  } else if (order != ORDER_NATURAL) {
    compute_ordering(order, n, rowp, cols, P.perm, P.iperm, groups);
    perm = P.perm.data();
    iperm = P.iperm.data();
  } else {
    P.perm.clear();
    P.iperm.clear();
  }

  std::vector<int> parent(n), post(n), cc(n);
  etree(n, rowp, cols, perm, iperm, parent.data());
  postorder(n, parent.data(), post.data());
  column_counts(n, rowp, cols, post.data(), perm, iperm, parent.data(),
                cc.data());

  std::vector<int> vtosn(n), sntov(n);
  int ns = fundamental_supernodes(n, post.data(), parent.data(), cc.data(),
                                  vtosn.data(), sntov.data());
  std::vector<int> sptr(ns + 1, 0);
  for (int i = 0; i < n; i++) sptr[vtosn[i] + 1]++;
  for (int s = 0; s < ns; s++) sptr[s + 1] += sptr[s];
  std::vector<int> spar(ns);
  for (int s = 0; s < ns; s++) {
    int last = sntov[sptr[s + 1] - 1];
    int p = parent[last];
    spar[s] = p >= 0 ? vtosn[p] : -1;
  }
  std::vector<int> scptr(ns + 2), skids(ns);
  child_lists(ns, spar.data(), scptr.data(), skids.data());

  std::vector<int> var_to_node, ncc;
  int K = relax_supernodes(n, ns, spar.data(), scptr.data(), skids.data(),
                           cc.data(), sptr.data(), sntov.data(), P.node_ptr,
                           P.node_to_var, var_to_node, ncc, P.node_parent);
  P.num_nodes = K;
  P.children_ptr.assign(K + 2, 0);
  P.children.assign(K, 0);
  child_lists(K, P.node_parent.data(), P.children_ptr.data(),
              P.children.data());

  P.contrib_ptr.assign(K + 1, 0);
  for (int k = 0; k < K; k++)
    P.contrib_ptr[k + 1] = ncc[k] - (P.node_ptr[k + 1] - P.node_ptr[k]);
  for (int k = 0; k < K; k++) P.contrib_ptr[k + 1] += P.contrib_ptr[k];
  P.max_contrib = 0;
  for (int k = 0; k < K; k++)
    P.max_contrib =
        std::max(P.max_contrib, P.contrib_ptr[k + 1] - P.contrib_ptr[k]);
  P.contrib_rows.assign(P.contrib_ptr[K], 0);
  contribution_rows(n, rowp, cols, perm, iperm, parent.data(), K,
                    P.node_ptr.data(), var_to_node.data(),
                    P.node_to_var.data(), P.contrib_ptr.data(),
                    P.contrib_rows.data());
  if (perm) {
    for (int i = 0; i < n; i++) P.node_to_var[i] = perm[P.node_to_var[i]];
  }
  P.cholesky_factor_nnz = 0;
  P.cholesky_int_nnz = 0;
  for (int k = 0; k < K; k++) {
    int64_t w = P.node_ptr[k + 1] - P.node_ptr[k];
    int64_t c = P.contrib_ptr[k + 1] - P.contrib_ptr[k];
    P.cholesky_int_nnz += w;
    P.cholesky_factor_nnz += w * (w + c);
  }
}

void concat_plans(const std::vector<const SymbolicPlan*>& blocks,
                  const std::vector<int>& var_offset, SymbolicPlan& P) {
  int n = 0, K = 0;
  size_t nc = 0;
  for (const SymbolicPlan* b : blocks) {
    n += b->n;
    K += b->num_nodes;
    nc += b->contrib_rows.size();
  }
  P.n = n;
  P.order = blocks.empty() ? ORDER_NATURAL : blocks[0]->order;
  P.num_nodes = K;
  P.perm.assign(n, 0);
  P.iperm.assign(n, 0);
  P.node_ptr.assign(K + 1, 0);
  P.node_to_var.assign(n, 0);
  P.node_parent.assign(K, -1);
  P.contrib_ptr.assign(K + 1, 0);
  P.contrib_rows.assign(nc, 0);
  P.max_contrib = 0;
  P.cholesky_factor_nnz = 0;
  P.cholesky_int_nnz = 0;
  int k0 = 0;
  size_t c0 = 0;
  for (size_t b = 0; b < blocks.size(); b++) {
    const SymbolicPlan& B = *blocks[b];
    const int v0 = var_offset[b];
    for (int i = 0; i < B.n; i++) {
      P.perm[v0 + i] = v0 + (B.perm.empty() ? i : B.perm[i]);
      P.node_to_var[v0 + i] = v0 + B.node_to_var[i];
    }
    for (int k = 0; k < B.num_nodes; k++) {
      P.node_ptr[k0 + k + 1] = v0 + B.node_ptr[k + 1];
      P.node_parent[k0 + k] = B.node_parent[k] >= 0 ? k0 + B.node_parent[k] : -1;
      P.contrib_ptr[k0 + k + 1] = int(c0) + B.contrib_ptr[k + 1];
    }
    for (size_t q = 0; q < B.contrib_rows.size(); q++) P.contrib_rows[c0 + q] = v0 + B.contrib_rows[q];
    P.max_contrib = std::max(P.max_contrib, B.max_contrib);
    P.cholesky_factor_nnz += B.cholesky_factor_nnz;
    P.cholesky_int_nnz += B.cholesky_int_nnz;
    k0 += B.num_nodes;
    c0 += B.contrib_rows.size();
  }
  for (int i = 0; i < n; i++) P.iperm[P.perm[i]] = i;
  P.children_ptr.assign(K + 2, 0);
  P.children.assign(K, 0);
  child_lists(K, P.node_parent.data(), P.children_ptr.data(), P.children.data());
}

}  // namespace amigo_amd
