"""The label layer on the GPU: nearest neighbours and HDBSCAN.

``HipHDBSCANClusterer`` mirrors the reference's HDBSCANClusterer
(src/pybasin/predictors/hdbscan_clusterer.py:12-148): sklearn HDBSCAN
semantics (min_cluster_size, min_samples counting the point itself,
excess-of-mass selection, no single cluster, Euclidean metric), the
silhouette-based min_cluster_size auto-tuning over the same five candidate
sizes with min_samples = min(10, size // 5), and the k-NN majority vote that
assigns noise points to clusters.  The O(n^2) parts (neighbour lists, the
Boruvka rounds of the mutual-reachability spanning tree, silhouette sums)
run as HIP kernels of libbasinhip.so; the O(n log n) tree logic runs in its
host code.  The spanning tree depends only on min_samples, so the
auto-tuning candidates that share it reuse one tree.

``HipKNeighborsClassifier`` is the case studies' KNeighborsClassifier on
the GPU (fit on template features, predict by majority of the k nearest).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib

MAX_FEATURES = 32


def _as_rows(X, device) -> torch.Tensor:
    t = torch.as_tensor(X)
    if t.ndim != 2:
        raise ValueError(f"expected (n, F) features, got shape {tuple(t.shape)}")
    if t.shape[1] > MAX_FEATURES:
        raise ValueError(f"at most {MAX_FEATURES} features, got {t.shape[1]}")
    return t.to(device=device, dtype=torch.float64).contiguous()


def knn(Q: torch.Tensor, R: torch.Tensor, k: int, stream=None):
    """(distances, indices) of the k nearest rows of R for every row of Q
    (device tensors, float64; ascending, equal distances by lower index)."""
    nq, F = Q.shape
    nr = R.shape[0]
    d = torch.empty((nq, k), dtype=torch.float64, device=Q.device)
    i = torch.empty((nq, k), dtype=torch.int64, device=Q.device)
    if nq == 0:
        return d, i
    s = stream or torch.cuda.current_stream(Q.device)
    with torch.cuda.device(Q.device):
        _lib.check(_lib.lib().bh_knn(Q.data_ptr(), nq, R.data_ptr(), nr, F, int(k), d.data_ptr(),
                                     i.data_ptr(), s.cuda_stream), "bh_knn")
    return d, i


def _workspace(n: int, device) -> torch.Tensor:
    return torch.empty((int(_lib.lib().bh_hdbscan_workspace_bytes(n)),), dtype=torch.uint8,
                       device=device)


def core_distances(X: torch.Tensor, min_samples: int, ordered: bool = False) -> torch.Tensor:
    """Distance to the min_samples-th nearest row, the row itself included
    (sklearn HDBSCAN's core distance).  The kernel prunes by tiles of
    consecutive rows; unless ordered is set, the rows are put in spatial
    order first (the result is in the caller's row order either way)."""
    n, F = X.shape
    perm = None if ordered else spatial_order(X)
    Xs = X if perm is None else X[perm].contiguous()
    core = torch.empty((n,), dtype=torch.float64, device=X.device)
    ws = _workspace(n, X.device)
    s = torch.cuda.current_stream(X.device)
    with torch.cuda.device(X.device):
        _lib.check(_lib.lib().bh_core_distances(Xs.data_ptr(), n, F, int(min_samples), core.data_ptr(),
                                                ws.data_ptr(), ws.numel(), s.cuda_stream),
                   "bh_core_distances")
    if perm is None:
        return core
    out = torch.empty_like(core)
    out[perm] = core
    return out


def spatial_order(X: torch.Tensor) -> torch.Tensor:
    """Row order along a Z-order (Morton) curve of the first one to three
    features, each quantised to 21 bits over its range: rows close in the
    order are close in space, which lets the spanning-tree rounds skip
    whole tiles of rows."""
    d = min(X.shape[1], 3)
    lo = X[:, :d].min(0).values
    span = (X[:, :d].max(0).values - lo).clamp_min(1e-300)
    u = ((X[:, :d] - lo) / span * (2**21 - 1)).clamp(0, 2**21 - 1).to(torch.int64)
    code = torch.zeros(X.shape[0], dtype=torch.int64, device=X.device)
    for b in range(21):
        for f in range(d):
            code |= ((u[:, f] >> b) & 1) << (b * d + f)
    return torch.argsort(code, stable=True)


def mutual_reachability_mst(X: torch.Tensor, core: torch.Tensor, ordered: bool = False):
    """Spanning-tree edges (a, b, w) of the mutual-reachability graph (device),
    in increasing weight (equal weights in the order the rounds found them).
    The rounds run on the rows in spatial order (put so here unless ordered
    is set); edges come back in the caller's row indices."""
    n, F = X.shape
    if ordered:
        perm = torch.arange(n, device=X.device)
    else:
        perm = spatial_order(X)
        X = X[perm].contiguous()
        core = core[perm].contiguous()
    a = torch.empty((n - 1,), dtype=torch.int64, device=X.device)
    b = torch.empty_like(a)
    w = torch.empty((n - 1,), dtype=torch.float64, device=X.device)
    ws = _workspace(n, X.device)
    s = torch.cuda.current_stream(X.device)
    with torch.cuda.device(X.device):
        _lib.check(_lib.lib().bh_hdbscan_mst(X.data_ptr(), n, F, core.data_ptr(), a.data_ptr(),
                                             b.data_ptr(), w.data_ptr(), ws.data_ptr(), ws.numel(),
                                             s.cuda_stream), "bh_hdbscan_mst")
    # increasing weight (stable), so the host hierarchy build skips its sort
    w, order = torch.sort(w, stable=True)
    return perm[a[order]], perm[b[order]], w


def labels_from_mst(a: np.ndarray, b: np.ndarray, w: np.ndarray, n: int,
                    min_cluster_size) -> np.ndarray:
    """Flat HDBSCAN labels (-1 noise) from a spanning tree given on the host;
    min_cluster_size may be a list (one label row per size, one hierarchy)."""
    a = np.ascontiguousarray(a, dtype=np.int64)
    b = np.ascontiguousarray(b, dtype=np.int64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    sizes = np.ascontiguousarray(np.atleast_1d(min_cluster_size), dtype=np.int32)
    lab = np.empty((len(sizes), n), dtype=np.int32)
    nc = np.zeros(len(sizes), dtype=np.int32)
    _lib.check(_lib.lib().bh_hdbscan_labels(a.ctypes.data, b.ctypes.data, w.ctypes.data, n,
                                            sizes.ctypes.data, len(sizes), lab.ctypes.data,
                                            nc.ctypes.data), "bh_hdbscan_labels")
    lab = lab.astype(np.int64)
    return lab if np.ndim(min_cluster_size) else lab[0]


def silhouette_score(X: torch.Tensor, labels) -> float:
    """Mean silhouette value over the rows (sklearn.metrics.silhouette_score,
    Euclidean; labels any integers, at least two distinct).  Equal rows with
    equal labels are collapsed into one weighted row first (dense attractor
    features repeat exactly), which cuts the O(n^2) pair sums without
    changing the score beyond rounding."""
    lab = torch.as_tensor(np.asarray(labels), device=X.device)
    uniq, inv = torch.unique(lab, return_inverse=True)
    if len(uniq) < 2:
        raise ValueError("silhouette needs at least two labels")
    # (label, row) pairs, sorted by label first: contiguous label blocks
    keyed = torch.cat([inv.to(torch.float64)[:, None], X.to(torch.float64)], 1)
    rows, counts = torch.unique(keyed, dim=0, return_counts=True)
    ls = rows[:, 0].to(torch.int32).contiguous()
    Xs = rows[:, 1:].contiguous()
    w = counts.to(torch.float64).contiguous()
    m = Xs.shape[0]
    lc = torch.zeros(len(uniq), dtype=torch.float64, device=X.device).index_add_(0, ls.long(), w)
    starts = torch.zeros(len(uniq) + 1, dtype=torch.int64, device=X.device)
    starts[1:] = torch.cumsum(torch.bincount(ls.long(), minlength=len(uniq)), 0)
    s = torch.empty((m,), dtype=torch.float64, device=X.device)
    sums = torch.empty((len(uniq), m), dtype=torch.float64, device=X.device)
    st = torch.cuda.current_stream(X.device)
    with torch.cuda.device(X.device):
        _lib.check(_lib.lib().bh_silhouette(Xs.data_ptr(), m, Xs.shape[1], ls.data_ptr(),
                                            starts.data_ptr(), len(uniq), w.data_ptr(), lc.data_ptr(),
                                            sums.data_ptr(), s.data_ptr(), st.cuda_stream),
                   "bh_silhouette")
    return float((s * w).sum() / w.sum())


class HipHDBSCANClusterer:
    """HDBSCANClusterer on the GPU (hdbscan_clusterer.py:12-148)."""

    def __init__(self, min_cluster_size: int = 50, min_samples: int | None = 10,
                 assign_noise: bool = False, n_neighbors: int = 5, auto_tune: bool = False,
                 device: str = "cuda:0"):
        self.min_cluster_size = int(min_cluster_size)
        self.min_samples = min_samples
        self.assign_noise = assign_noise
        self.n_neighbors = int(n_neighbors)
        self.auto_tune = auto_tune
        self.device = torch.device(device)
        self._trees: dict[int, tuple] = {}

    # one spanning tree per min_samples value
    def _tree(self, X: torch.Tensor, min_samples: int):
        if min_samples not in self._trees:
            if min_samples > X.shape[0]:
                raise ValueError(f"min_samples ({min_samples}) must be at most the number of "
                                 f"samples ({X.shape[0]})")
            # rows in spatial order once for both the core distances and the tree
            perm = spatial_order(X)
            Xs = X[perm].contiguous()
            a, b, w = mutual_reachability_mst(Xs, core_distances(Xs, min_samples, ordered=True),
                                              ordered=True)
            self._trees[min_samples] = (perm[a].cpu().numpy(), perm[b].cpu().numpy(), w.cpu().numpy())
        return self._trees[min_samples]

    def _hdbscan(self, X: torch.Tensor, min_cluster_size: int, min_samples: int | None) -> np.ndarray:
        ms = min_cluster_size if min_samples is None else min_samples
        a, b, w = self._tree(X, ms)
        return labels_from_mst(a, b, w, X.shape[0], min_cluster_size)

    def _find_optimal_min_cluster_size(self, X: torch.Tensor) -> int:
        n = X.shape[0]
        sizes = [max(10, int(0.005 * n)), max(25, int(0.01 * n)), max(50, int(0.02 * n)),
                 max(100, int(0.03 * n)), max(150, int(0.05 * n))]
        best_score, best = -1.0, sizes[0]
        self.scores_ = {}
        # candidates sharing min_samples share a tree: label them together
        by_ms: dict[int, list[int]] = {}
        for size in sizes:
            by_ms.setdefault(min(10, size // 5), []).append(size)
        labels = {}
        for ms, group in by_ms.items():
            a, b, w = self._tree(X, ms)
            for size, lab in zip(group, labels_from_mst(a, b, w, X.shape[0], group)):
                labels[size] = lab
        self._tuned_labels = labels
        seen: dict[bytes, float] = {}  # candidate sizes often give the same clustering
        for size in sizes:
            lab = labels[size]
            keep = lab != -1
            if len(np.unique(lab[keep])) >= 2 and keep.sum() > 1:
                key = lab.tobytes()
                if key not in seen:
                    seen[key] = silhouette_score(X[torch.as_tensor(keep, device=X.device)], lab[keep])
                score = seen[key]
                self.scores_[size] = score
                if score > best_score:
                    best_score, best = score, size
        return best

    def _assign_noise(self, X: torch.Tensor, lab: np.ndarray) -> np.ndarray:
        out = lab.copy()
        noise = lab == -1
        if not noise.any() or noise.all():
            return out
        lab_idx = np.nonzero(~noise)[0]
        k = min(self.n_neighbors, len(lab_idx))
        R = X[torch.as_tensor(lab_idx, device=X.device)].contiguous()
        Q = X[torch.as_tensor(np.nonzero(noise)[0], device=X.device)].contiguous()
        _, idx = knn(Q, R, k)
        neigh = lab[lab_idx][idx.cpu().numpy()]            # (n_noise, k)
        # majority vote, ties to the smallest label (np.unique order + argmax)
        classes = np.unique(lab[lab_idx])
        votes = (neigh[:, :, None] == classes[None, None, :]).sum(1)
        out[noise] = classes[np.argmax(votes, axis=1)]
        return out

    def fit_predict(self, X, y=None) -> np.ndarray:
        Xd = _as_rows(X, self.device)
        self._trees = {}
        if self.auto_tune:
            size = self._find_optimal_min_cluster_size(Xd)
            self.min_cluster_size = size
            self.min_samples = min(10, size // 5)
            lab = self._tuned_labels[size]  # same tree, same size: the labels are these
        else:
            lab = self._hdbscan(Xd, self.min_cluster_size, self.min_samples)
        if self.assign_noise:
            lab = self._assign_noise(Xd, lab)
        self.labels_ = lab
        return lab


class HipKNeighborsClassifier:
    """KNeighborsClassifier(n_neighbors=k) on the GPU: fit on (template)
    features and labels, predict the majority label of the k nearest
    (distance ties by lower training index, vote ties by smallest label)."""

    def __init__(self, n_neighbors: int = 1, device: str = "cuda:0"):
        self.n_neighbors = int(n_neighbors)
        self.device = torch.device(device)

    def fit(self, X, y) -> "HipKNeighborsClassifier":
        self._X = _as_rows(X, self.device)
        self.classes_, self._y = np.unique(np.asarray(y), return_inverse=True)
        return self

    def predict(self, X) -> np.ndarray:
        Q = _as_rows(X, self.device)
        k = min(self.n_neighbors, self._X.shape[0])
        _, idx = knn(Q, self._X, k)
        neigh = self._y[idx.cpu().numpy()]
        votes = (neigh[:, :, None] == np.arange(len(self.classes_))[None, None, :]).sum(1)
        return self.classes_[np.argmax(votes, axis=1)]
