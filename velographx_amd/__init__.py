"""VeloGraphX traversal path on AMD Instinct MI355X (gfx950).

Drop-in for the reference's ``velographx`` Python module on the traversal
path: ``Graph``, ``UpdateBatch``, ``IncrementalBFS``, ``IncrementalPageRank``,
``PageRankValidation`` and the ``from_*`` ingestion helpers, plus
the static ``CsrGraph`` with ``bfs_distances`` and ``pagerank``, the
consolidation policy, controller and device snapshot
(``ConsolidationController``, ``consolidate_to_csr_snapshot``),
``MultiSourceBFS`` (up to 64 roots per pass) and, in ``velographx_amd.autograd``,
differentiable weighted personalized PageRank, backed by
hand-written HIP kernels (``libvgx.so``) through a C++ facade.

``import torch`` comes first on purpose: torch ships its own HIP runtime,
and loading it before ``libvgx.so`` makes both share one runtime (same
SONAME), so device pointers from torch tensors are valid in the engine.
The extension must be built (``python velographx_amd/build.py``); there is
no CPU fallback.
"""
import torch  # noqa: F401  (see module docstring)

from ._velographx import (  # noqa: F401
    AdaptivePolicy,
    AdaptiveSelector,
    ConsolidationController,
    ConsolidationControllerConfig,
    ConsolidationPolicy,
    ConsolidationSignal,
    ConsolidationSnapshot,
    CsrGraph,
    Graph,
    IncrementalBFS,
    IncrementalComponents,
    IncrementalKCore,
    IncrementalPageRank,
    IncrementalSSSP,
    MultiSourceBFS,
    PageRankValidation,
    UpdateBatch,
    bfs_distances,
    consolidate_to_csr_snapshot,
    device_count,
    evaluate_consolidation,
    from_arrow_columns,
    from_arrow_table,
    from_numpy_edges,
    from_scipy_csr,
    pagerank,
    rmat_edges,
    scale_aware_consolidation_policy,
    version,
)

__all__ = [
    "Graph", "UpdateBatch", "IncrementalBFS", "IncrementalPageRank", "PageRankValidation",
    "from_numpy_edges", "from_scipy_csr", "from_arrow_columns", "from_arrow_table",
    "rmat_edges", "device_count", "version", "MultiSourceBFS", "AdaptivePolicy", "AdaptiveSelector",
    "ConsolidationPolicy", "ConsolidationSignal", "ConsolidationControllerConfig", "ConsolidationController",
    "ConsolidationSnapshot", "evaluate_consolidation", "scale_aware_consolidation_policy",
    "consolidate_to_csr_snapshot", "CsrGraph", "bfs_distances", "pagerank", "IncrementalSSSP", "IncrementalKCore", "IncrementalComponents",
]


# ---------------------------------------------------------------- zero-copy
# Update batches already in device memory and DLPack views of the results
# (python/bindings.cpp:21-36 takes host arrays only; these add the device
# path): a batch in torch device tensors is applied without a host copy, and
# distances or ranks come back as torch tensors over the engine's own
# device arrays.


def _device_batch(src, dst, add):
    if not all(isinstance(t, torch.Tensor) and t.is_cuda for t in (src, dst, add)):
        raise TypeError("src, dst and add must be device tensors")
    if src.dtype not in (torch.int32, torch.uint32) or dst.dtype not in (torch.int32, torch.uint32):
        raise TypeError("src and dst must be int32 or uint32 tensors")
    if add.dtype not in (torch.uint8, torch.bool):
        raise TypeError("add must be a uint8 or bool tensor")
    if not (src.numel() == dst.numel() == add.numel()):
        raise ValueError("src, dst and add must have equal length")
    if not (src.device == dst.device == add.device):
        raise ValueError("src, dst and add must be on one device")
    src, dst, add = src.contiguous(), dst.contiguous(), add.contiguous()
    if add.dtype == torch.bool:
        add = add.view(torch.uint8)
    stream = torch.cuda.current_stream(src.device).cuda_stream
    return src, dst, add, stream


def _graph_apply_tensors(self, src, dst, add):
    """DynamicGraph::apply for a batch held in device tensors (int32/uint32
    ids, uint8/bool add flags) written on the current stream."""
    s, d, a, stream = _device_batch(src, dst, add)
    if s.device.index != self.device:
        raise ValueError("the batch must be on the graph's device")
    self._apply_device(s.data_ptr(), d.data_ptr(), a.data_ptr(), s.numel(), stream)


def _bfs_apply_tensors(self, src, dst, add):
    """IncrementalBFS::apply for a batch held in device tensors."""
    s, d, a, stream = _device_batch(src, dst, add)
    self._apply_device(s.data_ptr(), d.data_ptr(), a.data_ptr(), s.numel(), stream)


def _bfs_distances_tensor(self):
    """The current distances as a uint32 device tensor sharing the engine's
    array (no copy), for reading on the current stream: the object's next
    apply or recompute waits on the device for the work queued there before
    it writes the distances (the tensor then shows the new ones, or, where
    the array had to grow, keeps the old ones), and the memory stays valid
    while the tensor lives."""
    return torch.from_dlpack(self.distances_dlpack(torch.cuda.current_stream().cuda_stream))


def _pr_values_tensor(self):
    """The current ranks as a float64 (float32 at precision 32) device tensor
    sharing the engine's array, for reading on the current stream; as
    IncrementalBFS.distances_tensor."""
    return torch.from_dlpack(self.values_dlpack(torch.cuda.current_stream().cuda_stream))


Graph.apply_tensors = _graph_apply_tensors
IncrementalBFS.apply_tensors = _bfs_apply_tensors
IncrementalBFS.distances_tensor = _bfs_distances_tensor
IncrementalPageRank.values_tensor = _pr_values_tensor


# ------------------------------------------------ edge lists in device memory
# from_numpy_edges / from_scipy_csr / from_arrow_columns / from_arrow_table
# (python/bindings.cpp:21-36) read host arrays.  Given arrays that live in
# device memory (torch tensors on the GPU, or any DLPack producer on a ROCm
# device), the same functions build the graph from them where they lie: the
# ids are checked and sorted into rows on the device and only two 8-byte
# results cross PCIe.  Host inputs keep the host path unchanged.

_host_from_numpy_edges = from_numpy_edges
_host_from_scipy_csr = from_scipy_csr
_host_from_arrow_columns = from_arrow_columns
_KDL_ROCM = 10


def _device_tensor(x):
    """x as a torch device tensor when it lives in device memory, else None."""
    if isinstance(x, torch.Tensor):
        return x if x.is_cuda else None
    dev = getattr(x, "__dlpack_device__", None)
    if dev is not None and not isinstance(x, type):
        kind, _ = dev()
        if kind == _KDL_ROCM:
            return torch.from_dlpack(x)
    return None


def _id_tensor(t, what):
    """t with a dtype the device build reads (int32, uint32, int64); narrower
    integer types are widened, anything else is refused."""
    if t.dtype in (torch.int32, torch.uint32, torch.int64):
        return t
    if t.dtype in (torch.int8, torch.int16, torch.uint8, torch.uint16):
        return t.to(torch.int32)
    raise TypeError(f"{what} must hold integer ids, not {t.dtype}")


def _edges_on_device(src, dst, directed):
    """Graph(0, directed) + one UpdateBatch of additions (src[i], dst[i]), the
    lists in device memory: bulk-loaded on the device (same adjacency, vertex
    count and version as the batch; version stays 0 for an empty list)."""
    if src.device != dst.device:
        raise ValueError("source and destination ids must be on one device")
    if src.dtype != dst.dtype:
        src, dst = src.to(torch.int64), dst.to(torch.int64)
    with torch.cuda.device(src.device):
        g = Graph(0, directed)
        m = src.shape[0]
        if m == 0:
            return g
        stride = src.stride(0)
        if stride <= 0 or dst.stride(0) != stride:
            src, dst = src.contiguous(), dst.contiguous()
            stride = 1
        stream = torch.cuda.current_stream(src.device).cuda_stream
        g._bulk_load_device(src.data_ptr(), dst.data_ptr(), m, stride, src.element_size(), stream)
    return g


def from_numpy_edges(edges, directed=False):
    """python/bindings.cpp:21-25: a Graph from an (N, 2) array of edges.
    Device arrays (torch GPU tensors, DLPack producers on a ROCm device) are
    read in place on the device."""
    t = _device_tensor(edges)
    if t is None:
        return _host_from_numpy_edges(edges, directed)
    if t.dim() != 2 or t.shape[1] != 2:
        raise ValueError("edge array must have shape (N, 2)")
    t = _id_tensor(t, "edge array")
    return _edges_on_device(t[:, 0], t[:, 1], directed)


def from_arrow_columns(src, dst, directed=False):
    """python/bindings.cpp:34: a Graph from source and destination columns;
    device columns are read in place."""
    s, d = _device_tensor(src), _device_tensor(dst)
    if s is None and d is None:
        return _host_from_arrow_columns(src, dst, directed)
    if s is None or d is None:
        raise ValueError("source and destination columns must both be in device memory or both on the host")
    if s.dim() != 1 or d.dim() != 1:
        raise ValueError("columns must be one-dimensional")
    if s.shape[0] != d.shape[0]:
        raise ValueError("Arrow source and destination columns must have equal length")
    return _edges_on_device(_id_tensor(s, "source column"), _id_tensor(d, "destination column"), directed)


def from_arrow_table(table, src="src", dst="dst", directed=False):
    """python/bindings.cpp:35: from_arrow_columns over two named columns."""
    if not hasattr(table, "column"):
        raise ValueError("expected an Arrow-like table exposing column(name)")
    return from_arrow_columns(table.column(src), table.column(dst), directed)


def from_scipy_csr(csr, directed=False):
    """python/bindings.cpp:26-32: a Graph from a square CSR matrix (an
    undirected graph takes the entries with v >= u).  A torch sparse CSR
    tensor on the GPU, or a CSR-like object whose indptr and indices are
    device arrays, is read in place."""
    if isinstance(csr, torch.Tensor) and csr.layout == torch.sparse_csr and csr.is_cuda:
        indptr, indices, shape = csr.crow_indices(), csr.col_indices(), tuple(csr.shape)
    else:
        if not (hasattr(csr, "indptr") and hasattr(csr, "indices") and hasattr(csr, "shape")):
            raise ValueError("expected a SciPy CSR-like object with indptr, indices, and shape")
        indptr, indices = _device_tensor(csr.indptr), _device_tensor(csr.indices)
        if indptr is None and indices is None:
            return _host_from_scipy_csr(csr, directed)
        if indptr is None or indices is None:
            raise ValueError("indptr and indices must both be in device memory or both on the host")
        shape = tuple(csr.shape)
    if len(shape) != 2 or shape[0] != shape[1]:
        raise ValueError("CSR graph matrix must be square")
    n = int(shape[0])
    if indptr.dim() != 1 or indptr.shape[0] != n + 1:
        raise ValueError("invalid CSR indptr length")
    if indptr.device != indices.device:
        raise ValueError("indptr and indices must be on one device")
    if indptr.dtype == torch.int32 and indices.dtype == torch.int32:
        wide = torch.int32
    else:
        wide = torch.int64
    indptr = _id_tensor(indptr, "indptr").to(wide).contiguous()
    indices = _id_tensor(indices, "indices").to(wide).contiguous()
    with torch.cuda.device(indptr.device):
        g = Graph(n, directed)
        stream = torch.cuda.current_stream(indptr.device).cuda_stream
        g._bulk_load_csr_device(indptr.data_ptr(), indices.data_ptr() if indices.numel() else indptr.data_ptr(),
                                n, indices.numel(), indptr.element_size(), stream)
    return g
