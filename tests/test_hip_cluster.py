"""GPU label layer (libbasinhip.so bh_knn / bh_hdbscan_* / bh_silhouette)
against scikit-learn, the library the reference's HDBSCANClusterer and
KNeighborsClassifier call (src/pybasin/predictors/hdbscan_clusterer.py:12-148).

sklearn 1.7.2 is installed here (the reference pins 1.8.0; the HDBSCAN,
NearestNeighbors and silhouette code paths used are unchanged between them).
The mutual-reachability graph max(core_a, core_b, d(a, b)) has many equal
edge weights (every edge shorter than both core distances weighs a core
distance), so its minimum spanning tree is not unique: sklearn's Prim order
and this Boruvka order can attach a point that sits exactly at a split
height to either side, and number the clusters differently.  Labels are
therefore compared up to a renaming of the clusters, with a handful of
tie-placed points allowed to differ (measured: 1-2 points in 1,500-4,000).
"""

import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _blobs(n, seed, noise_frac=0.05, F=2):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-10, 10, size=(4, F))
    scales = np.array([0.3, 0.6, 1.0, 0.4])
    k = rng.integers(0, 4, n)
    X = centers[k] + rng.normal(size=(n, F)) * scales[k, None]
    m = int(noise_frac * n)
    X[:m] = rng.uniform(-14, 14, size=(m, F))
    return X


def _mismatch(a, b):
    """rows whose labels differ after the best renaming of a's clusters onto
    b's (noise -1 maps to noise only)."""
    from scipy.optimize import linear_sum_assignment

    a, b = np.asarray(a), np.asarray(b)
    ua, ub = np.unique(a), np.unique(b)
    C = np.array([[np.sum((a == x) & (b == y)) for y in ub] for x in ua], dtype=float)
    for i, x in enumerate(ua):  # noise only matches noise
        for j, y in enumerate(ub):
            if (x == -1) != (y == -1):
                C[i, j] = -1e9
    r, c = linear_sum_assignment(-C)
    return len(a) - int(sum(max(C[i, j], 0) for i, j in zip(r, c)))


@pytest.mark.parametrize("F,k", [(2, 1), (3, 5), (8, 10), (16, 32), (20, 10), (32, 4)])
def test_knn_matches_sklearn(cuda_device, F, k):
    from sklearn.neighbors import NearestNeighbors

    from pybasin_amd.cluster import knn

    rng = np.random.default_rng(F)
    R = rng.normal(size=(3000, F))
    Q = rng.normal(size=(700, F))
    d, i = knn(torch.tensor(Q, device=cuda_device), torch.tensor(R, device=cuda_device), k)
    dr, ir = NearestNeighbors(n_neighbors=k, algorithm="brute").fit(R).kneighbors(Q)
    np.testing.assert_allclose(d.cpu().numpy(), dr, rtol=1e-12, atol=1e-12)
    assert np.array_equal(i.cpu().numpy(), ir)


@pytest.mark.parametrize("F,k,clustered", [(2, 10, True), (3, 5, False), (8, 16, True), (2, 32, False),
                                            (20, 10, True)])
def test_pruned_core_distances_equal_knn(cuda_device, F, k, clustered):
    """The tile-pruned core distances equal column k - 1 of the brute-force
    k-NN bit for bit, in spatial and in arbitrary row order (duplicates
    included, as dense attractor features have them)."""
    from pybasin_amd.cluster import core_distances, knn, spatial_order

    rng = np.random.default_rng(F + k)
    if clustered:
        centers = rng.normal(size=(5, F)) * 3
        X = centers[rng.integers(0, 5, 5000)] + rng.normal(size=(5000, F)) * 0.05
        X[:300] = X[300:600]  # exact duplicates
    else:
        X = rng.normal(size=(4321, F))
    Xd = torch.tensor(X, device=cuda_device)
    want = knn(Xd, Xd, k)[0][:, -1]
    assert torch.equal(core_distances(Xd, k), want)
    perm = spatial_order(Xd)
    got = core_distances(Xd[perm].contiguous(), k, ordered=True)
    assert torch.equal(got, want[perm])


def test_mst_weights_equal_dense_reference(cuda_device):
    """The Boruvka tree has the weight multiset of a minimum spanning tree of
    the dense mutual-reachability matrix (scipy)."""
    from scipy.sparse.csgraph import minimum_spanning_tree
    from scipy.spatial.distance import cdist
    from sklearn.neighbors import NearestNeighbors

    from pybasin_amd.cluster import core_distances, mutual_reachability_mst

    X = _blobs(1500, 3)
    Xd = torch.tensor(X, device=cuda_device)
    core = core_distances(Xd, 10)
    core_ref = NearestNeighbors(n_neighbors=10).fit(X).kneighbors(X)[0][:, -1]
    np.testing.assert_allclose(core.cpu().numpy(), core_ref, rtol=1e-12)
    a, b, w = mutual_reachability_mst(Xd, core)
    D = cdist(X, X)
    M = np.maximum(np.maximum(D, core_ref[:, None]), core_ref[None, :])
    np.fill_diagonal(M, 0.0)
    T = minimum_spanning_tree(M).tocoo()
    np.testing.assert_allclose(np.sort(w.cpu().numpy()), np.sort(T.data), rtol=1e-12)
    # a tree: n - 1 edges that connect every row
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    A = coo_matrix((np.ones(len(a)), (a.cpu().numpy(), b.cpu().numpy())), shape=(1500, 1500))
    assert connected_components(A, directed=False)[0] == 1


@pytest.mark.parametrize("mcs,ms,seed", [(50, 10, 1), (15, 5, 2), (100, 10, 4), (30, None, 5)])
def test_hdbscan_labels_match_sklearn(cuda_device, mcs, ms, seed):
    from sklearn.cluster import HDBSCAN

    from pybasin_amd.cluster import HipHDBSCANClusterer

    X = _blobs(4000, seed)
    want = HDBSCAN(min_cluster_size=mcs, min_samples=ms).fit_predict(X)
    got = HipHDBSCANClusterer(min_cluster_size=mcs, min_samples=ms,
                              device=str(cuda_device)).fit_predict(X)
    miss = _mismatch(got, want)
    print(f"[result] mcs {mcs} ms {ms}: clusters {len(set(want)) - (-1 in want)}, noise "
          f"{(want == -1).mean():.3f}, rows differing after renaming {miss}", flush=True)
    assert len(set(got)) == len(set(want))
    assert miss <= 0.001 * len(X)


def test_silhouette_matches_sklearn(cuda_device):
    from sklearn.metrics import silhouette_score as sk

    from pybasin_amd.cluster import silhouette_score

    X = _blobs(3000, 7, noise_frac=0.0, F=3)
    lab = np.random.default_rng(0).integers(0, 5, 3000)
    lab[:3] = 9  # a small cluster
    lab[3] = 11  # a cluster of one (silhouette 0)
    got = silhouette_score(torch.tensor(X, device=cuda_device), lab)
    assert got == pytest.approx(sk(X, lab), rel=1e-10, abs=1e-12)


def test_silhouette_with_repeated_rows_matches_sklearn(cuda_device):
    """Rows repeated many times (as dense attractor features are) collapse
    into weighted rows; the score still equals sklearn's."""
    from sklearn.metrics import silhouette_score as sk

    from pybasin_amd.cluster import silhouette_score

    rng = np.random.default_rng(4)
    base = rng.normal(size=(60, 2))
    X = base[rng.integers(0, 60, 4000)]
    lab = (X[:, 0] > 0).astype(np.int64) + 2 * (X[:, 1] > 0.5)
    lab[:2] = 7  # one label shared by a repeated row and its copy
    X[1] = X[0]
    got = silhouette_score(torch.tensor(X, device=cuda_device), lab)
    # the definition with direct differences (sklearn forms distances from
    # dot products, so equal rows come out a little above 0 there)
    D = np.sqrt(((X[:, None, :] - X[None, :, :]) ** 2).sum(-1))
    s = np.zeros(len(X))
    for i in range(len(X)):
        own = lab == lab[i]
        if own.sum() > 1:
            a = D[i, own].sum() / (own.sum() - 1)
            b = min(D[i, lab == c].mean() for c in np.unique(lab) if c != lab[i])
            s[i] = (b - a) / max(a, b)
    assert got == pytest.approx(s.mean(), rel=1e-12, abs=1e-14)
    assert got == pytest.approx(sk(X, lab), rel=1e-8)


def _reference_clusterer(X):
    """The reference's HDBSCANClusterer(auto_tune=True, assign_noise=True)
    steps (hdbscan_clusterer.py:50-148) on sklearn objects."""
    from sklearn.cluster import HDBSCAN
    from sklearn.metrics import silhouette_score
    from sklearn.neighbors import NearestNeighbors

    n = len(X)
    sizes = [max(10, int(0.005 * n)), max(25, int(0.01 * n)), max(50, int(0.02 * n)),
             max(100, int(0.03 * n)), max(150, int(0.05 * n))]
    best, best_size = -1.0, sizes[0]
    for size in sizes:
        lab = HDBSCAN(min_cluster_size=size, min_samples=min(10, size // 5)).fit_predict(X)
        m = lab != -1
        if len(np.unique(lab[m])) >= 2 and m.sum() > 1:
            sc = silhouette_score(X[m], lab[m])
            if sc > best:
                best, best_size = sc, size
    lab = HDBSCAN(min_cluster_size=best_size, min_samples=min(10, best_size // 5)).fit_predict(X)
    noise = lab == -1
    if noise.any() and not noise.all():
        nn = NearestNeighbors(n_neighbors=min(5, (~noise).sum())).fit(X[~noise])
        _, idx = nn.kneighbors(X[noise])
        ll = lab[~noise]
        out = lab.copy()
        for r, nb in zip(np.nonzero(noise)[0], idx):
            u, c = np.unique(ll[nb], return_counts=True)
            out[r] = u[np.argmax(c)]
        lab = out
    return lab, best_size


def test_auto_tuned_clusterer_with_noise_assignment_matches_reference_steps(cuda_device):
    from pybasin_amd.cluster import HipHDBSCANClusterer

    X = _blobs(5000, 11)
    want, size = _reference_clusterer(X)
    c = HipHDBSCANClusterer(auto_tune=True, assign_noise=True, device=str(cuda_device))
    got = c.fit_predict(X)
    print(f"[result] auto-tuned size {c.min_cluster_size} (reference steps {size}); "
          f"scores {c.scores_}", flush=True)
    assert c.min_cluster_size == size
    miss = _mismatch(got, want)
    print(f"[result] rows differing after renaming {miss}", flush=True)
    assert len(set(got)) == len(set(want))
    assert miss <= 0.002 * len(X)


def test_knn_classifier_matches_sklearn(cuda_device):
    from sklearn.neighbors import KNeighborsClassifier

    from pybasin_amd.cluster import HipKNeighborsClassifier

    rng = np.random.default_rng(3)
    T = rng.normal(size=(40, 4))
    y = rng.choice(["a", "b", "c"], 40)
    Q = rng.normal(size=(5000, 4))
    for k in (1, 5):
        want = KNeighborsClassifier(n_neighbors=k, algorithm="brute").fit(T, y).predict(Q)
        got = HipKNeighborsClassifier(k, device=str(cuda_device)).fit(T, y).predict(Q)
        assert np.array_equal(got, want)


def test_hdbscan_on_duffing_features_at_scale(cuda_device):
    """Duffing (max x, std x) features of 20,000 integrated ICs: the GPU
    clusterer reproduces sklearn HDBSCAN's labels; times printed."""
    from sklearn.cluster import HDBSCAN

    from oracle import oracle as O
    from pybasin_amd.cluster import HipHDBSCANClusterer
    from pybasin_amd.features import feature_from_stats
    from pybasin_amd.solvers import HipDopri5Solver
    from pybasin_amd.systems import DuffingODE

    y0 = torch.tensor(O.sample_uniform(20_000, [-1.0, -0.5], [1.0, 1.0], 42)).to(cuda_device)
    solver = HipDopri5Solver(t_span=(0, 1000), t_steps=1000, t_eval=(900.0, 1000.0), device="cuda",
                             cache_dir=None)
    res = solver.integrate_fused(DuffingODE({"delta": 0.08, "k3": 1.0, "A": 0.2}), y0,
                                 time_steady=900.0)
    X = torch.stack([feature_from_stats("maximum", res.stats)[:, 0],
                     feature_from_stats("standard_deviation", res.stats)[:, 0]], 1)
    X = X.to(torch.float32).double().cpu().numpy()  # the extractor's float32 columns
    t0 = time.perf_counter()
    want = HDBSCAN(min_cluster_size=200, min_samples=10).fit_predict(X)
    t_cpu = time.perf_counter() - t0
    t0 = time.perf_counter()
    got = HipHDBSCANClusterer(min_cluster_size=200, min_samples=10,
                              device=str(cuda_device)).fit_predict(X)
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t0
    miss = _mismatch(got, want)
    print(f"[result] duffing 20k: sklearn {t_cpu:.2f} s, HIP {t_gpu:.3f} s, clusters "
          f"{len(set(want)) - (-1 in want)}, rows differing after renaming {miss}", flush=True)
    assert len(set(got)) == len(set(want))
    assert miss <= 0.005 * len(X)


def test_unsupervised_estimator_on_pendulum_features(cuda_device):
    """BasinStabilityEstimator with the GPU HDBSCAN clusterer (auto-tuned,
    noise assigned: the reference's default predictor) on the pendulum
    case-1 ICs and log_delta(omega) features: the labels are those of the
    reference's clusterer steps on sklearn for the same features, and the
    fixed-point basin comes out as one cluster with the reference's
    supervised fraction (main_pendulum_case1.json: FP 0.1505).  (The
    limit-cycle windows differ only by sampling phase, and HDBSCAN splits
    that band by density -- in sklearn as here.)"""
    import json
    import os

    from oracle import oracle as O
    from pybasin_amd.cluster import HipHDBSCANClusterer
    from pybasin_amd.estimator import BasinStabilityEstimator
    from pybasin_amd.features import HipFusedFeatureExtractor
    from pybasin_amd.solvers import HipDopri5Solver
    from pybasin_amd.systems import PendulumODE

    with open(os.path.join(os.path.dirname(__file__), "golden", "label_fixtures.json")) as f:
        fx = json.load(f)["pendulum_case1"]
    y0 = torch.tensor(O.sample_uniform(fx["n"], fx["min_limits"], fx["max_limits"], fx["seed"]))
    solver = HipDopri5Solver(t_span=(0, 1000), t_steps=1000, device="cuda", cache_dir=None)
    ext = HipFusedFeatureExtractor(time_steady=950.0, features=None,
                                   features_per_state={1: {"log_delta": None}})
    bse = BasinStabilityEstimator(PendulumODE({"alpha": 0.1, "T": 0.5, "K": 1.0}), None, fx["n"],
                                  solver, ext, predictor=HipHDBSCANClusterer(
                                      auto_tune=True, assign_noise=True, device=str(cuda_device)))
    res = bse.run(y0)
    want, _ = _reference_clusterer(res.features.double().numpy())
    miss = _mismatch(res.labels.astype(np.int64), want)
    fp = [b["basinStability"] for b in fx["bs"] if b["label"] == "FP"][0]
    print(f"[result] unsupervised pendulum: {res.basin_stability}; rows differing from the "
          f"reference steps on sklearn {miss}", flush=True)
    assert miss <= 0.002 * fx["n"]
    assert fp in res.basin_stability.values()
