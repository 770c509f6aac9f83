// cluster.hip -- the label layer on the GPU: k nearest neighbours, HDBSCAN
// (core distances, minimum spanning tree of the mutual-reachability graph,
// condensed tree, excess-of-mass selection) and silhouette scores.
//
// Replaces the sklearn calls behind the reference's HDBSCANClusterer
// (src/pybasin/predictors/hdbscan_clusterer.py:12-148): HDBSCAN.fit_predict,
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
// (tiles of rows in spatial order skipped by bounding boxes; O(log n)
// rounds), the host merges components.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <numeric>
#include <vector>

namespace bh {
namespace clu {

constexpr int TILE = 256;   // reference points staged per LDS tile
constexpr int MAXF = 32;    // features (rows are padded to 2, 4, 8, 16 or 32 slots)
constexpr int MAXK = 32;    // neighbours

// Kernels are instantiated for FC = 2, 4, 8, 16, 32 feature slots; rows hold F
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

// Boruvka rounds over tiles of QT consecutive rows.  Rows should arrive in
// a spatially coherent order (the caller sorts them along a space-filling
// curve): then a tile's bounding box is small, and a query can skip every
// tile whose lower bound max(core_a^2, min core^2 of the tile, squared
// distance to the box) already loses to its best edge, or whose rows all
// lie in the query's own component.
constexpr int QT = 64;

// per-tile bounding box (FC slots; padded slots are 0) and, when core_sq is
// given, the smallest squared core distance, once per spanning tree
template <int FC>
__global__ void tile_bounds_kernel(const double *X, int64_t n, int F, const double *core_sq,
                                   int64_t n_tiles, double *lo, double *hi, double *mincore) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tiles) return;
    double l[FC], h[FC], x[FC];
#pragma unroll
    for (int f = 0; f < FC; f++) {
        l[f] = INFINITY;
        h[f] = -INFINITY;
    }
    double mc = INFINITY;
    const int64_t r1 = min(n, (t + 1) * QT);
    for (int64_t r = t * QT; r < r1; r++) {
        load_row<FC>(x, X + r * F, F, true);
#pragma unroll
        for (int f = 0; f < FC; f++) {
            l[f] = fmin(l[f], x[f]);
            h[f] = fmax(h[f], x[f]);
        }
        if (core_sq) mc = fmin(mc, core_sq[r]);
    }
#pragma unroll
    for (int f = 0; f < FC; f++) {
        lo[t * FC + f] = l[f];
        hi[t * FC + f] = h[f];
    }
    if (mincore) mincore[t] = mc;
}

// the component all rows of a tile share, or -1, once per round
__global__ void tile_comp_kernel(const int32_t *comp, int64_t n, int64_t n_tiles, int32_t *tcomp) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tiles) return;
    int32_t c = comp[t * QT];
    const int64_t r1 = min(n, (t + 1) * QT);
    for (int64_t r = t * QT + 1; r < r1 && c >= 0; r++)
        if (comp[r] != c) c = -1;
    tcomp[t] = c;
}

// One Boruvka round: for each row the cheapest mutual-reachability edge to
// a row of another component, key max(core_a^2, core_b^2, d^2) (monotone in
// the edge weight); among equal keys the partner nearest in row order, then
// the lower index (dense regions have many equal keys, and this order lets
// a tie found nearby prune every farther tile).  One wave per query tile;
// reference tiles are visited outward from the query's own, so good edges
// are found early.  The box bound is a true lower bound of the computed key:
// each difference to the box is at most the difference to any row in it,
// and rounding to nearest is monotone.
template <int FC>
__global__ void __launch_bounds__(QT) boruvka_kernel(const double *X, int64_t n, int F,
                                                     const double *core_sq, const int32_t *comp,
                                                     const double *lo, const double *hi,
                                                     const double *mincore, const int32_t *tcomp,
                                                     int64_t n_tiles, double *best_key,
                                                     int64_t *best_j) {
    const int64_t qt = blockIdx.x;
    const int64_t i = qt * QT + threadIdx.x;
    const bool valid = i < n;
    double x[FC];
    load_row<FC>(x, X + (valid ? i : 0) * F, F, valid);
    const double ci_sq = valid ? core_sq[i] : 0.0;
    const int32_t ci = valid ? comp[i] : -2;  // -2 matches no tile component
    double bk = valid ? INFINITY : -1.0;        // an idle lane never needs a tile
    int64_t bj = -1, bd = INT64_MAX;           // partner, its distance in row order
    for (int64_t s = 0; s < 2 * n_tiles; s++) {
        // tiles qt, qt + 1, qt - 1, qt + 2, qt - 2, ...
        const int64_t d = (s + 1) / 2;
        const int64_t t = (s & 1) ? qt + d : qt - d;
        if (t < 0 || t >= n_tiles) continue;
        double g = 0.0;
#pragma unroll
        for (int f = 0; f < FC; f++) {
            const double v = fmax(fmax(__dsub_rn(lo[t * FC + f], x[f]), __dsub_rn(x[f], hi[t * FC + f])), 0.0);
            g = __dadd_rn(g, __dmul_rn(v, v));
        }
        const double lb = fmax(fmax(ci_sq, mincore[t]), g);
        const int64_t t0 = t * QT, t1 = min(n, t0 + QT) - 1;
        const int64_t near = i < t0 ? t0 - i : (i > t1 ? i - t1 : 0);  // nearest row index distance
        const bool need = tcomp[t] != ci && !(lb > bk) && !(lb == bk && near > bd);
        if (!__any(need)) continue;
        const int64_t r1 = min(n, (t + 1) * QT);
        for (int64_t j = t * QT; j < r1; j++) {
            if (comp[j] == ci) continue;
            const double lo2 = fmax(ci_sq, core_sq[j]);
            if (lo2 > bk) continue;
            double y[FC];
            load_row<FC>(y, X + j * F, F, true);
            const double key = fmax(lo2, sqdist<FC>(x, y));
            const int64_t dj = j > i ? j - i : i - j;
            if (key < bk || (key == bk && (dj < bd || (dj == bd && j < bj)))) {
                bk = key;
                bj = j;
                bd = dj;
            }
        }
    }
    if (valid) {
        best_key[i] = bk;
        best_j[i] = bj;
    }
}

// Core distances: the k-th smallest distance from each row to the rows of X
// (itself included), for rows in spatial order.  One wave per query tile,
// reference tiles outward from the query's own; a tile is skipped when its
// box is no closer than every lane's current k-th distance.  The k smallest
// squared distances are kept ascending in K registers (slots past k hold
// the next ones and are not read); only the k-th value leaves the kernel,
// and it does not depend on the visiting order.
template <int K, int FC>
__global__ void __launch_bounds__(QT) core_distance_kernel(const double *X, int64_t n, int F, int k,
                                                           const double *lo, const double *hi,
                                                           int64_t n_tiles, double *core) {
    const int64_t qt = blockIdx.x;
    const int64_t i = qt * QT + threadIdx.x;
    const bool valid = i < n;
    double x[FC];
    load_row<FC>(x, X + (valid ? i : 0) * F, F, valid);
    double bd[K];
#pragma unroll
    for (int j = 0; j < K; j++) bd[j] = INFINITY;
    double kd = INFINITY;  // bd[k - 1], kept beside the list (no runtime index into it)
    for (int64_t s = 0; s < 2 * n_tiles; s++) {
        const int64_t d = (s + 1) / 2;
        const int64_t t = (s & 1) ? qt + d : qt - d;
        if (t < 0 || t >= n_tiles) continue;
        double g = 0.0;
#pragma unroll
        for (int f = 0; f < FC; f++) {
            const double v = fmax(fmax(__dsub_rn(lo[t * FC + f], x[f]), __dsub_rn(x[f], hi[t * FC + f])), 0.0);
            g = __dadd_rn(g, __dmul_rn(v, v));
        }
        const bool need = valid && g < kd;
        if (!__any(need)) continue;
        const int64_t r1 = min(n, (t + 1) * QT);
        for (int64_t j = t * QT; j < r1; j++) {
            double y[FC];
            load_row<FC>(y, X + j * F, F, true);
            const double v = sqdist<FC>(x, y);
            if (v < bd[K - 1]) {
                // insert into the ascending list
#pragma unroll
                for (int m = K - 1; m > 0; m--) bd[m] = (v < bd[m - 1]) ? bd[m - 1] : fmin(bd[m], v);
                bd[0] = fmin(bd[0], v);
                double nk = bd[0];
#pragma unroll
                for (int m = 1; m < K; m++) nk = (m == k - 1) ? bd[m] : nk;
                kd = nk;
            }
        }
    }
    if (valid) core[i] = sqrt(kd);
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
// [start[c], start[c+1]), each row standing for w rows (w = 1 when no
// weights are given: equal rows collapse into one weighted row).
// silhouette_sums_kernel: block (x, c) sums the weighted distances from its
// rows to every row of label c into sums[c * n + i];
// silhouette_value_kernel: with N_c the weighted size of label c, a = own
// sum / (N_own - 1) (the row's own copies add distance 0), b = smallest
// other sum / N_c, s = (b - a) / max(a, b) (0 for a label of one row).
template <int FC>
__global__ void __launch_bounds__(TILE) silhouette_sums_kernel(const double *X, const double *w,
                                                               int64_t n, int F,
                                                               const int64_t *start,
                                                               double *sums) {
    __shared__ double tile[TILE * FC];
    __shared__ double tile_w[TILE];
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
        if (w)
            for (int e = threadIdx.x; e < cnt; e += TILE) tile_w[e] = w[r0 + e];
        __syncthreads();
        if (i < n) {
            if (w)
                for (int j = 0; j < cnt; j++) sum += tile_w[j] * sqrt(sqdist<FC>(x, tile + j * FC));
            else
                for (int j = 0; j < cnt; j++) sum += sqrt(sqdist<FC>(x, tile + j * FC));
        }
    }
    if (i < n) sums[(int64_t)c * n + i] = sum;
}

__global__ void silhouette_value_kernel(const double *sums, int64_t n, const int32_t *label,
                                        const int64_t *start, const double *counts, int C,
                                        double *s_out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int li = label[i];
    double a = 0.0, b = INFINITY;
    for (int c = 0; c < C; c++) {
        const double size = counts ? counts[c] : (double)(start[c + 1] - start[c]);
        const double sum = sums[(int64_t)c * n + i];
        if (c == li)
            a = size > 1.0 ? sum / (size - 1.0) : 0.0;
        else if (size > 0.0)
            b = fmin(b, sum / size);
    }
    const double own = counts ? counts[li] : (double)(start[li + 1] - start[li]);
    s_out[i] = (own > 1.0 && isfinite(b)) ? (b - a) / fmax(a, b) : 0.0;
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
