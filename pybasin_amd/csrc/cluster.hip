// cluster.hip -- the label layer on the GPU: k nearest neighbours, HDBSCAN
// (core distances, minimum spanning tree of the mutual-reachability graph,
// condensed tree, excess-of-mass selection) and silhouette scores.
//
// Replaces the sklearn calls behind the reference's HDBSCANClusterer
// (src/pybasin/predictors/hdbscan_clusterer.py:16-166): HDBSCAN.fit_predict,
// NearestNeighbors.kneighbors (noise assignment) and silhouette_score
// (min_cluster_size auto-tuning).  The algorithm is HDBSCAN* as published
// (Campello, Moulavi, Sander 2013; McInnes, Healy 2017): core distance =
// distance to the min_samples-th nearest point counting the point itself,
// mutual reachability max(core_a, core_b, d(a, b)), its minimum spanning
// tree, the single-linkage hierarchy of that tree, the hierarchy condensed
// with min_cluster_size, cluster stability sum((lambda_p - lambda_birth) *
// size), excess-of-mass selection without the root, and noise for points
// that leave the tree outside a selected cluster.
//
// Distances are Euclidean in double, accumulated feature by feature as
// sum((a_f - b_f)^2) with separate multiplies and adds; neighbour searches
// compare squared distances.  The spanning tree is built by Boruvka rounds:
// every point finds its cheapest edge to another component on the GPU
// (O(n^2) per round, O(log n) rounds), the host merges components.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <numeric>
#include <vector>

namespace bh {
namespace clu {

constexpr int TILE = 256;   // reference points staged per LDS tile
constexpr int MAXF = 16;    // features (rows are padded to 2, 4, 8 or 16 slots)
constexpr int MAXK = 32;    // neighbours
constexpr int MST_CHUNKS = 8;  // reference-row chunks per Boruvka round

// Kernels are instantiated for FC = 2, 4, 8, 16 feature slots; rows hold F
// <= FC features and the slots beyond F are zero (adding (0 - 0)^2 = +0 to a
// sum leaves it unchanged, so padded distances equal unpadded ones).
template <int FC>
__device__ __forceinline__ double sqdist(const double *a, const double *b) {
    double s = 0.0;
#pragma unroll
    for (int f = 0; f < FC; f++) {
        const double t = __dsub_rn(a[f], b[f]);
        s = __dadd_rn(s, __dmul_rn(t, t));
    }
    return s;
}

// one row of F features into FC slots
template <int FC>
__device__ __forceinline__ void load_row(double (&x)[FC], const double *src, int F, bool valid) {
#pragma unroll
    for (int f = 0; f < FC; f++) x[f] = (valid && f < F) ? src[f] : 0.0;
}

// rows [r0, r0 + cnt) of an (n, F) array into an LDS tile of FC-wide rows
template <int FC>
__device__ __forceinline__ void stage_tile(double *tile, const double *src, int64_t r0, int cnt,
                                           int F) {
    for (int e = threadIdx.x; e < cnt * FC; e += TILE) {
        const int j = e / FC, f = e % FC;
        tile[e] = f < F ? src[(r0 + j) * F + f] : 0.0;
    }
}

// k smallest squared distances from each query row to the reference rows
// (ascending; equal distances keep the lower reference index).  Outputs are
// the distances (not squared) and the indices, (nq, k) row-major.
template <int K, int FC>
__global__ void __launch_bounds__(TILE) knn_kernel(const double *Q, int64_t nq, const double *R,
                                                   int64_t nr, int F, int k_out, double *out_d,
                                                   int64_t *out_i) {
    __shared__ double tile[TILE * FC];
    const int64_t q = (int64_t)blockIdx.x * TILE + threadIdx.x;
    double x[FC];
    load_row<FC>(x, Q + q * F, F, q < nq);
    double bd[K];
    int64_t bi[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        bd[k] = INFINITY;
        bi[k] = -1;
    }
    for (int64_t r0 = 0; r0 < nr; r0 += TILE) {
        const int cnt = (int)min((int64_t)TILE, nr - r0);
        __syncthreads();
        stage_tile<FC>(tile, R, r0, cnt, F);
        __syncthreads();
        if (q >= nq) continue;
        for (int j = 0; j < cnt; j++) {
            const double d = sqdist<FC>(x, tile + j * FC);
            if (d < bd[K - 1]) {
                // insertion into the sorted list (strict <: earlier index wins ties)
                double cd = d;
                int64_t ci = r0 + j;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    if (cd < bd[k]) {
                        const double td = bd[k];
                        const int64_t ti = bi[k];
                        bd[k] = cd;
                        bi[k] = ci;
                        cd = td;
                        ci = ti;
                    }
                }
            }
        }
    }
    if (q < nq) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (k < k_out) {
                out_d[q * k_out + k] = sqrt(bd[k]);
                out_i[q * k_out + k] = bi[k];
            }
        }
    }
}

// One Boruvka round: for each point the cheapest mutual-reachability edge
// to a point of another component.  The comparison key is
// max(core_a^2, core_b^2, d^2) (monotone in the edge weight); ties go to the
// lower index.  The reference rows are split into gridDim.y chunks (more
// waves in flight than one row per lane gives); chunk y writes its best to
// best_key / best_j[y * n + i] (best_j = -1: no row of another component in
// the chunk), and the host takes the minimum over chunks.
template <int FC>
__global__ void __launch_bounds__(TILE) boruvka_kernel(const double *X, int64_t n, int F,
                                                       const double *core_sq, const int32_t *comp,
                                                       int64_t chunk, double *best_key,
                                                       int64_t *best_j) {
    __shared__ double tile[TILE * FC];
    __shared__ double tcore[TILE];
    __shared__ int32_t tcomp[TILE];
    const int64_t i = (int64_t)blockIdx.x * TILE + threadIdx.x;
    const int64_t j0 = (int64_t)blockIdx.y * chunk, j1 = min(n, j0 + chunk);
    double x[FC];
    load_row<FC>(x, X + i * F, F, i < n);
    const double ci_sq = i < n ? core_sq[i] : 0.0;
    const int32_t ci = i < n ? comp[i] : -1;
    double bk = i < n ? INFINITY : -1.0;  // an idle lane never updates
    int64_t bj = -1;
    for (int64_t r0 = j0; r0 < j1; r0 += TILE) {
        const int cnt = (int)min((int64_t)TILE, j1 - r0);
        __syncthreads();
        stage_tile<FC>(tile, X, r0, cnt, F);
        for (int e = threadIdx.x; e < cnt; e += TILE) {
            tcore[e] = core_sq[r0 + e];
            tcomp[e] = comp[r0 + e];
        }
        __syncthreads();
        for (int j = 0; j < cnt; j++) {
            if (tcomp[j] == ci) continue;
            const double lo = fmax(ci_sq, tcore[j]);
            if (lo > bk) continue;  // cannot beat the current best
            const double key = fmax(lo, sqdist<FC>(x, tile + j * FC));
            if (key < bk) {
                bk = key;
                bj = r0 + j;
            }
        }
    }
    if (i < n) {
        best_key[blockIdx.y * n + i] = bk;
        best_j[blockIdx.y * n + i] = bj;
    }
}

// exact weights max(core_a, core_b, d(a, b)) of the chosen edges
template <int FC>
__global__ void edge_weight_kernel(const double *X, int F, const double *core, const int64_t *ea,
                                   const int64_t *eb, int64_t m, double *w) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const int64_t a = ea[e], b = eb[e];
    double xa[FC], xb[FC];
    load_row<FC>(xa, X + a * F, F, true);
    load_row<FC>(xb, X + b * F, F, true);
    const double d = sqrt(sqdist<FC>(xa, xb));
    w[e] = fmax(fmax(core[a], core[b]), d);
}

// Silhouette: rows sorted by label, labels 0..C-1 in contiguous blocks
// [start[c], start[c+1]).  silhouette_sums_kernel: block (x, c) sums the
// distances from its rows to every row of label c into sums[c * n + i];
// silhouette_value_kernel: a = own-label mean (excluding the row itself),
// b = smallest other-label mean, s = (b - a) / max(a, b) (0 for a label of
// one row).
template <int FC>
__global__ void __launch_bounds__(TILE) silhouette_sums_kernel(const double *X, int64_t n, int F,
                                                               const int64_t *start,
                                                               double *sums) {
    __shared__ double tile[TILE * FC];
    const int64_t i = (int64_t)blockIdx.x * TILE + threadIdx.x;
    const int c = blockIdx.y;
    double x[FC];
    load_row<FC>(x, X + i * F, F, i < n);
    const int64_t c0 = start[c], c1 = start[c + 1];
    double sum = 0.0;
    for (int64_t r0 = c0; r0 < c1; r0 += TILE) {
        const int cnt = (int)min((int64_t)TILE, c1 - r0);
        __syncthreads();
        stage_tile<FC>(tile, X, r0, cnt, F);
        __syncthreads();
        if (i < n)
            for (int j = 0; j < cnt; j++) sum += sqrt(sqdist<FC>(x, tile + j * FC));
    }
    if (i < n) sums[(int64_t)c * n + i] = sum;
}

__global__ void silhouette_value_kernel(const double *sums, int64_t n, const int32_t *label,
                                        const int64_t *start, int C, double *s_out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int li = label[i];
    double a = 0.0, b = INFINITY;
    for (int c = 0; c < C; c++) {
        const int64_t size = start[c + 1] - start[c];
        const double sum = sums[(int64_t)c * n + i];
        if (c == li)
            a = size > 1 ? sum / (double)(size - 1) : 0.0;
        else if (size > 0)
            b = fmin(b, sum / (double)size);
    }
    const int64_t own = start[li + 1] - start[li];
    s_out[i] = (own > 1 && isfinite(b)) ? (b - a) / fmax(a, b) : 0.0;
}

// ---- host side: spanning tree rounds and the cluster hierarchy -------------

struct UnionFind {
    std::vector<int64_t> parent, size;
    explicit UnionFind(int64_t n) : parent(n), size(n, 1) { std::iota(parent.begin(), parent.end(), 0); }
    int64_t find(int64_t x) {
        int64_t r = x;
        while (parent[r] != r) r = parent[r];
        while (parent[x] != r) {
            const int64_t nx = parent[x];
            parent[x] = r;
            x = nx;
        }
        return r;
    }
};

// Single-linkage hierarchy of a spanning tree: merges in increasing weight
// (stable for equal weights), node n + m for merge m, children and sizes.
struct Hierarchy {
    std::vector<int64_t> left, right, size;
    std::vector<double> dist;
};

inline Hierarchy single_linkage(int64_t n, const int64_t *ea, const int64_t *eb, const double *w) {
    // increasing weight, equal weights in edge order (a stable sort), sorted
    // as contiguous (weight, index) pairs; edges that arrive sorted (the
    // device sorts them) are taken as they are
    std::vector<int64_t> order(n - 1);
    if (std::is_sorted(w, w + (n - 1))) {
        std::iota(order.begin(), order.end(), 0);
    } else {
        std::vector<std::pair<double, int64_t>> keyed(n - 1);
        for (int64_t e = 0; e < n - 1; e++) keyed[e] = {w[e], e};
        std::sort(keyed.begin(), keyed.end());
        for (int64_t e = 0; e < n - 1; e++) order[e] = keyed[e].second;
    }
    Hierarchy h;
    h.left.resize(n - 1);
    h.right.resize(n - 1);
    h.size.resize(n - 1);
    h.dist.resize(n - 1);
    // union-find over points whose root carries the hierarchy node id
    UnionFind uf(n);
    std::vector<int64_t> node_of(n);
    std::iota(node_of.begin(), node_of.end(), 0);
    for (int64_t m = 0; m < n - 1; m++) {
        const int64_t e = order[m];
        const int64_t ra = uf.find(ea[e]), rb = uf.find(eb[e]);
        h.left[m] = node_of[ra];
        h.right[m] = node_of[rb];
        h.dist[m] = w[e];
        h.size[m] = uf.size[ra] + uf.size[rb];
        // union by size, the new root names node n + m
        int64_t big = ra, small = rb;
        if (uf.size[big] < uf.size[small]) std::swap(big, small);
        uf.parent[small] = big;
        uf.size[big] += uf.size[small];
        node_of[big] = n + m;
    }
    return h;
}

struct Condensed {
    std::vector<int64_t> parent, child, csize;
    std::vector<double> lambda;
};

// Condense the hierarchy with min_cluster_size (children visited breadth
// first from the root; cluster ids n, n + 1, ... in that order).
inline Condensed condense(const Hierarchy &h, int64_t n, int64_t mcs) {
    Condensed c;
    c.parent.reserve(n + 64);
    c.child.reserve(n + 64);
    c.csize.reserve(n + 64);
    c.lambda.reserve(n + 64);
    const int64_t root = 2 * n - 2;
    auto count = [&](int64_t node) { return node >= n ? h.size[node - n] : (int64_t)1; };
    std::vector<int64_t> relabel(2 * n - 1, -1);
    relabel[root] = n;
    int64_t next_label = n + 1;
    // points below a pruned node fall out at the pruning lambda
    std::vector<int64_t> stack;
    stack.reserve(64);
    auto fall_out = [&](int64_t top, int64_t cluster, double lam) {
        stack.clear();
        stack.push_back(top);
        while (!stack.empty()) {
            const int64_t v = stack.back();
            stack.pop_back();
            if (v < n) {
                c.parent.push_back(cluster);
                c.child.push_back(v);
                c.lambda.push_back(lam);
                c.csize.push_back(1);
            } else {
                stack.push_back(h.left[v - n]);
                stack.push_back(h.right[v - n]);
            }
        }
    };
    std::vector<int64_t> queue{root};
    for (size_t qi = 0; qi < queue.size(); qi++) {
        const int64_t node = queue[qi];
        if (node < n || relabel[node] < 0) continue;
        const int64_t m = node - n;
        const int64_t l = h.left[m], r = h.right[m];
        const double lam = h.dist[m] > 0.0 ? 1.0 / h.dist[m] : INFINITY;
        const int64_t lc = count(l), rc = count(r), cl = relabel[node];
        if (lc >= mcs && rc >= mcs) {
            relabel[l] = next_label++;
            relabel[r] = next_label++;
            c.parent.push_back(cl);
            c.child.push_back(relabel[l]);
            c.lambda.push_back(lam);
            c.csize.push_back(lc);
            c.parent.push_back(cl);
            c.child.push_back(relabel[r]);
            c.lambda.push_back(lam);
            c.csize.push_back(rc);
            queue.push_back(l);
            queue.push_back(r);
        } else if (lc < mcs && rc < mcs) {
            fall_out(l, cl, lam);
            fall_out(r, cl, lam);
        } else if (lc < mcs) {
            relabel[r] = cl;
            fall_out(l, cl, lam);
            queue.push_back(r);
        } else {
            relabel[l] = cl;
            fall_out(r, cl, lam);
            queue.push_back(l);
        }
    }
    return c;
}

// Excess-of-mass selection (root excluded) and labels: -1 noise, selected
// clusters numbered 0.. in increasing cluster id.  Returns the cluster count.
inline int32_t select_and_label(const Condensed &c, int64_t n, int32_t *labels) {
    int64_t n_clusters_total = n + 1;
    for (int64_t x : c.child)
        if (x >= n) n_clusters_total = std::max(n_clusters_total, x + 1);
    const int64_t nc = n_clusters_total - n;
    std::vector<double> birth(nc, 0.0), stab(nc, 0.0);
    std::vector<int64_t> cparent(nc, -1);
    for (size_t e = 0; e < c.child.size(); e++)
        if (c.child[e] >= n) {
            birth[c.child[e] - n] = c.lambda[e];
            cparent[c.child[e] - n] = c.parent[e] - n;
        }
    for (size_t e = 0; e < c.child.size(); e++) {
        const int64_t p = c.parent[e] - n;
        stab[p] += (c.lambda[e] - birth[p]) * (double)c.csize[e];
    }
    std::vector<std::vector<int64_t>> kids(nc);
    for (int64_t k = 1; k < nc; k++) kids[cparent[k]].push_back(k);
    std::vector<char> selected(nc, 0);
    // children carry larger ids than parents: decreasing id is bottom-up
    for (int64_t k = nc - 1; k >= 1; k--) {
        double sub = 0.0;
        for (int64_t ch : kids[k]) sub += stab[ch];
        if (sub > stab[k]) {
            selected[k] = 0;
            stab[k] = sub;
        } else {
            selected[k] = 1;
            std::vector<int64_t> stack(kids[k].begin(), kids[k].end());
            while (!stack.empty()) {
                const int64_t v = stack.back();
                stack.pop_back();
                selected[v] = 0;
                for (int64_t ch : kids[v]) stack.push_back(ch);
            }
        }
    }
    std::vector<int32_t> label_of(nc, -1);
    int32_t next = 0;
    for (int64_t k = 1; k < nc; k++)
        if (selected[k]) label_of[k] = next++;
    // a point's label: the nearest selected cluster on its way to the root
    std::vector<int64_t> point_parent(n, -1);
    for (size_t e = 0; e < c.child.size(); e++)
        if (c.child[e] < n) point_parent[c.child[e]] = c.parent[e] - n;
    std::vector<int32_t> resolved(nc, -2);
    std::vector<int64_t> path;
    auto cluster_label = [&](int64_t k) {
        path.clear();
        int32_t lab = -1;
        while (k >= 0) {
            if (resolved[k] != -2) {
                lab = resolved[k];
                break;
            }
            path.push_back(k);
            if (selected[k]) {
                lab = label_of[k];
                break;
            }
            k = cparent[k];
        }
        for (int64_t v : path) resolved[v] = lab;
        return lab;
    };
    for (int64_t p = 0; p < n; p++) labels[p] = point_parent[p] >= 0 ? cluster_label(point_parent[p]) : -1;
    return next;
}

}  // namespace clu
}  // namespace bh
