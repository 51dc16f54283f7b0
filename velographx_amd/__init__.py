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
    IncrementalPageRank,
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
    "consolidate_to_csr_snapshot", "CsrGraph", "bfs_distances", "pagerank",
]
