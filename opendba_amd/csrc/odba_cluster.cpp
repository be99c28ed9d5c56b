// Clustering of the all-vs-all distances and per-cluster medoids, as
// approximateMedoidIndices does it (dba.hpp:219-351):
//
//   * distances divided by their maximum (dba.hpp:186-226),
//   * complete-linkage hierarchical clustering (hclust_fast with
//     HCLUST_METHOD_COMPLETE from the hclust-cpp submodule, dba.hpp:234),
//   * one of three cuts: cdist > 1 -> cutree_k with K raised until cdist
//     clusters have two or more members (dba.hpp:237-264); cdist == 1 -> one
//     cluster (dba.hpp:266-272); 0 <= cdist < 1 -> cutree_cdist at that
//     dendrogram height (dba.hpp:273-277),
//   * per cluster, the medoid with the reference's rules (dba.hpp:297-351),
//     including its indexing: sums of squares are accumulated from the
//     distance rows of the cluster-member ordinals (not the members' global
//     rows) into one array that is not reset between clusters.
//
// hclust-cpp is not part of the reference tree (empty submodule, no pinned
// revision). What follows restates its published algorithms: Muellner's
// nearest-neighbour chain for complete linkage (reciprocal nearest
// neighbours, the previous chain element kept on ties, merged cluster stored
// at the larger index), merges stably sorted by height and relabelled with a
// union-find into R's merge convention (singletons -1..-n, clusters by
// step, the smaller label first), and the cutree_k / cutree_cdist cuts.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "../../include/odba/odba.h"

namespace {

inline long long tri_row(long long i, long long n) { return (n - 1) * n / 2 - (n - i - 1) * (n - i) / 2; }

// Complete linkage. d: condensed upper triangle (row-major, i < j), modified.
// merge: (n-1) x 2 column-major R matrix, height: n-1 merge heights.
void hclust_complete(int n, std::vector<double>& d, int* merge, double* height) {
  auto D = [&](int i, int j) -> double& { return i < j ? d[tri_row(i, n) + j - i - 1] : d[tri_row(j, n) + i - j - 1]; };
  // active nodes as a doubly linked list in index order (succ n = end)
  std::vector<int> succ(n), pred(n);
  for (int i = 0; i < n; i++) { succ[i] = i + 1; pred[i] = i - 1; }
  int start = 0;
  auto remove = [&](int x) {
    if (pred[x] >= 0) succ[pred[x]] = succ[x];
    else start = succ[x];
    if (succ[x] < n) pred[succ[x]] = pred[x];
  };
  struct Step { int a, b; double h; };
  std::vector<Step> steps;
  steps.reserve(n - 1);
  std::vector<int> chain(n + 1);
  int tip = 0;
  for (int s = 0; s < n - 1; s++) {
    int i1, i2;
    double mn;
    if (tip <= 3) {
      chain[0] = i1 = start;
      tip = 1;
      i2 = succ[i1];
      mn = D(i1, i2);
      for (int j = succ[i2]; j < n; j = succ[j])
        if (D(i1, j) < mn) { mn = D(i1, j); i2 = j; }
    } else {
      tip -= 3;
      i1 = chain[tip - 1];
      i2 = chain[tip];
      mn = D(i1, i2);
    }
    do {
      chain[tip] = i2;
      for (int j = start; j < n; j = succ[j]) {
        if (j == i2) continue;
        if (D(j, i2) < mn) { mn = D(j, i2); i1 = j; }
      }
      i2 = i1;
      i1 = chain[tip++];
    } while (i2 != chain[tip - 2]);
    if (i1 > i2) std::swap(i1, i2);
    steps.push_back({i1, i2, mn});
    for (int j = start; j < n; j = succ[j])
      if (j != i1 && j != i2) D(j, i2) = std::max(D(j, i1), D(j, i2));
    remove(i1);
  }
  std::stable_sort(steps.begin(), steps.end(), [](const Step& x, const Step& y) { return x.h < y.h; });
  // union-find over 2n-1 node ids; clusters get ids n, n+1, ...
  std::vector<int> parent(2 * n - 1, 0);
  int next = n;
  auto find = [&](int x) {
    int r = x;
    while (parent[r]) r = parent[r];
    while (parent[x] && parent[x] != r) { int p = parent[x]; parent[x] = r; x = p; }
    return r;
  };
  for (int s = 0; s < n - 1; s++) {
    int a = find(steps[s].a), b = find(steps[s].b);
    parent[a] = parent[b] = next++;
    if (a > b) std::swap(a, b);
    merge[s] = a < n ? -(a + 1) : a - n + 1;
    merge[s + n - 1] = b < n ? -(b + 1) : b - n + 1;
    height[s] = steps[s].h;
  }
}

// Labels 0.. for a cut of the dendrogram into nclust clusters: labels follow
// the order in which observations are first seen.
void cutree_k(int n, const int* merge, int nclust, int* labels) {
  if (nclust > n || nclust < 2) {
    for (int j = 0; j < n; j++) labels[j] = 0;
    return;
  }
  std::vector<int> last(n, 0);  // step (1-based) of each observation's last merge
  for (int k = 1; k <= n - nclust; k++) {
    int m1 = merge[k - 1], m2 = merge[n - 1 + k - 1];
    if (m1 < 0 && m2 < 0) {
      last[-m1 - 1] = last[-m2 - 1] = k;
    } else if (m1 < 0 || m2 < 0) {
      int j;
      if (m1 < 0) { j = -m1; m1 = m2; } else j = -m2;
      for (int l = 0; l < n; l++)
        if (last[l] == m1) last[l] = k;
      last[j - 1] = k;
    } else {
      for (int l = 0; l < n; l++)
        if (last[l] == m1 || last[l] == m2) last[l] = k;
    }
  }
  int label = 0;
  std::vector<int> z(n, -1);
  for (int j = 0; j < n; j++) {
    if (last[j] == 0) labels[j] = label++;
    else {
      if (z[last[j]] < 0) z[last[j]] = label++;
      labels[j] = z[last[j]];
    }
  }
}

void cutree_cdist(int n, const int* merge, const double* height, double cdist, int* labels) {
  int k = 0;
  for (; k < n - 1; k++)
    if (height[k] >= cdist) break;
  cutree_k(n, merge, n - k, labels);
}

template <typename T>
int cluster_t(const T* dists, int n, const uint64_t* lens, double cdist, int* memb, int* medoids, int* nclusters) {
  const long long np = (long long)n * (n - 1) / 2;
  T maxd = T(0);
  for (long long k = 0; k < np; k++)
    if (maxd < dists[k]) maxd = dists[k];
  if (maxd == T(0)) maxd = T(1);
  if (cdist > 1 || (cdist >= 0 && cdist < 1)) {
    std::vector<double> dd(np);
    for (long long k = 0; k < np; k++) dd[k] = (double)dists[k] / (double)maxd;
    std::vector<int> merge(2 * (n - 1) + 2);
    std::vector<double> height(n);
    if (n >= 2) hclust_complete(n, dd, merge.data(), height.data());
    if (cdist > 1) {
      int new_k = (int)cdist;
      if (new_k > n) new_k = n;
      int multi;
      do {
        cutree_k(n, merge.data(), new_k, memb);
        std::vector<int> cnt(new_k, 0);
        for (int i = 0; i < n; i++) cnt[memb[i]]++;
        multi = 0;
        for (int i = 0; i < new_k; i++)
          if (cnt[i] > 1) multi++;
        new_k += (int)cdist - multi;
      } while (multi < (int)cdist && new_k < n);
    } else {
      cutree_cdist(n, merge.data(), height.data(), cdist, memb);
    }
  } else {
    for (int i = 0; i < n; i++) memb[i] = 0;  // cdist == 1 (or negative: not implemented there either)
  }
  int nc = 1;
  for (int i = 0; i < n; i++) nc = std::max(nc, memb[i] + 1);
  // medoids (dba.hpp:297-351)
  std::vector<T> sos_all(n, T(0));
  if (nc == 1) {  // the whole-set sums of squares (dba.hpp:189-207)
    for (int r = 0; r + 1 < n; r++)
      for (int c = r + 1; c < n; c++) {
        T v = dists[tri_row(r, n) + c - r - 1];
        v *= v;
        sos_all[r] += v;
        sos_all[c] += v;
      }
  }
  std::vector<T> cl_sos(n, T(0));
  T* sos = nc == 1 ? sos_all.data() : cl_sos.data();
  for (int cc = 0; cc < nc; cc++) {
    std::vector<int> idx;
    for (int i = 0; i < n; i++)
      if (memb[i] == cc) idx.push_back(i);
    const int m = (int)idx.size();
    if (nc > 1) {
      for (int i = 0; i + 1 < m; i++) {
        const long long off = tri_row(i, n);  // member ordinal as the row
        for (int j = i + 1; j < m; j++) {
          const T pd = dists[off + j - i - 1];
          sos[i] += pd * pd;
          sos[j] += pd * pd;
        }
      }
    }
    int med = -1;
    if (m > 2) {
      T low = std::numeric_limits<T>::max();
      for (int i = 0; i < m; i++)
        if (sos[i] < low) { med = idx[i]; low = sos[i]; }
    } else if (m == 2) {
      med = lens[idx[0]] > lens[idx[1]] ? idx[0] : idx[1];
    } else if (m == 1) {
      med = idx[0];
    }
    if (med < 0) return ODBA_E_MEDOID;
    medoids[cc] = med;
  }
  *nclusters = nc;
  return ODBA_OK;
}

}  // namespace

extern "C" {

int odba_cluster(int dtype, const void* dists, int n, const uint64_t* lens, double cdist, int* memberships,
                 int* medoids, int* num_clusters) {
  if (!dists || !lens || !memberships || !medoids || !num_clusters || n < 1) return ODBA_E_INVALID;
  switch (dtype) {
    case ODBA_I32: return cluster_t<int32_t>((const int32_t*)dists, n, lens, cdist, memberships, medoids, num_clusters);
    case ODBA_U32: return cluster_t<uint32_t>((const uint32_t*)dists, n, lens, cdist, memberships, medoids, num_clusters);
    case ODBA_U64: return cluster_t<uint64_t>((const uint64_t*)dists, n, lens, cdist, memberships, medoids, num_clusters);
    case ODBA_F32: return cluster_t<float>((const float*)dists, n, lens, cdist, memberships, medoids, num_clusters);
    case ODBA_F64: return cluster_t<double>((const double*)dists, n, lens, cdist, memberships, medoids, num_clusters);
  }
  return ODBA_E_INVALID;
}

int odba_hclust_complete(int n, const double* condensed, int* merge, double* height) {
  if (!condensed || !merge || !height || n < 2) return ODBA_E_INVALID;
  std::vector<double> d(condensed, condensed + (size_t)n * (n - 1) / 2);
  hclust_complete(n, d, merge, height);
  return ODBA_OK;
}

int odba_cutree_k(int n, const int* merge, int nclust, int* labels) {
  if (!merge || !labels || n < 1) return ODBA_E_INVALID;
  cutree_k(n, merge, nclust, labels);
  return ODBA_OK;
}

}  // extern "C"
