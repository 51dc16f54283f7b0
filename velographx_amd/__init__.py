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
