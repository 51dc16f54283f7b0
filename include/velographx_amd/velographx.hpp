// velographx.hpp -- the whole C++20 facade over the gfx950 C ABI in vgx.h,
// with the reference engine's class names:
//   types.hpp, graph_access.hpp   VertexId, ReadableGraph / MutableGraph and
//                                 the access functions
//   dynamic_graph.hpp             EdgeUpdate, UpdateBatch, DynamicGraph
//   csr_graph.hpp, algorithms.hpp CsrGraph, bfs_distances, pagerank,
//                                 connected_components
//   bfs.hpp, pagerank.hpp, sssp.hpp, components.hpp, kcore.hpp
//                                 BasicIncremental*<Graph> and their
//                                 DynamicGraph aliases (IncrementalBFS, ...)
//   consolidation.hpp             snapshots and the consolidation controller
// The same headers are reachable under the reference's own paths
// (include/velographx/...).
#pragma once

#include "velographx_amd/algorithms.hpp"
#include "velographx_amd/bfs.hpp"
#include "velographx_amd/components.hpp"
#include "velographx_amd/csr_graph.hpp"
#include "velographx_amd/dynamic_graph.hpp"
#include "velographx_amd/graph_access.hpp"
#include "velographx_amd/kcore.hpp"
#include "velographx_amd/pagerank.hpp"
#include "velographx_amd/sssp.hpp"
#include "velographx_amd/types.hpp"
