// pybind11 module mirroring the reference's python/bindings.cpp:38-52 for
// the traversal path (Graph, UpdateBatch, IncrementalBFS,
// IncrementalPageRank, PageRankValidation, from_numpy_edges /
// from_scipy_csr / from_arrow_columns / from_arrow_table), backed by the
// gfx950 engine through the C++ facade.  Same names, argument names,
// defaults and return types; additions are marked as such.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "velographx_amd/consolidation.hpp"
#include "velographx_amd/csr_graph.hpp"
#include "velographx_amd/components.hpp"
#include "velographx_amd/kcore.hpp"
#include "velographx_amd/sssp.hpp"
#include "velographx_amd/pagerank.hpp"
#include "velographx_amd/selector.hpp"
#include "velographx_amd/velographx.hpp"

namespace py = pybind11;
using velographx::DynamicGraph;
using velographx::IncrementalBFS;
using velographx::IncrementalPageRank;
using velographx::PageRankValidation;
using velographx::UpdateBatch;
using velographx::VertexId;

namespace {

using U32Array = py::array_t<std::uint32_t, py::array::c_style | py::array::forcecast>;

py::dict trace_to_dict(const vgx_selector_trace& t) {
  py::dict d;
  d["update_fraction"] = t.update_fraction;
  d["shallow_parent_deletion_fraction"] = t.shallow_parent_deletion_fraction;
  d["predicted_incremental_us"] = t.predicted_incremental_us;
  d["predicted_full_us"] = t.predicted_full_us;
  d["execution_us"] = t.execution_us;
  d["decision_us"] = t.decision_us;
  d["affected_vertices"] = t.affected_vertices;
  d["first_batch"] = t.first_batch != 0;
  d["chose_full"] = t.chose_full != 0;
  d["internal_fallback"] = t.internal_fallback != 0;
  d["reason"] = velographx::policy_reason_name(t.reason);
  d["reason_code"] = t.reason;
  return d;
}

vgx_selector_trace trace_from_dict(const py::dict& d) {
  vgx_selector_trace t{};
  auto get = [&](const char* k, double def) { return d.contains(k) ? d[k].cast<double>() : def; };
  t.update_fraction = get("update_fraction", 0.0);
  t.shallow_parent_deletion_fraction = get("shallow_parent_deletion_fraction", 0.0);
  t.predicted_incremental_us = get("predicted_incremental_us", 0.0);
  t.predicted_full_us = get("predicted_full_us", 0.0);
  t.execution_us = get("execution_us", 0.0);
  t.decision_us = get("decision_us", 0.0);
  t.affected_vertices = d.contains("affected_vertices") ? d["affected_vertices"].cast<std::uint64_t>() : 0;
  t.first_batch = d.contains("first_batch") && d["first_batch"].cast<bool>() ? 1 : 0;
  t.chose_full = d.contains("chose_full") && d["chose_full"].cast<bool>() ? 1 : 0;
  t.internal_fallback = d.contains("internal_fallback") && d["internal_fallback"].cast<bool>() ? 1 : 0;
  t.reason = d.contains("reason_code") ? d["reason_code"].cast<int>() : 0;
  return t;
}

// bindings.cpp:21-25: an UpdateBatch of additions applied to Graph(0, directed).
DynamicGraph graph_from_numpy_edges(const U32Array& edges, bool directed) {
  const auto info = edges.request();
  if (info.ndim != 2 || info.shape[1] != 2) throw std::invalid_argument("edge array must have shape (N, 2)");
  DynamicGraph graph(0, directed);
  const auto m = static_cast<std::size_t>(info.shape[0]);
  const auto* data = static_cast<const std::uint32_t*>(info.ptr);
  std::vector<std::uint32_t> s(m), d(m);
  std::vector<std::uint8_t> a(m, 1);
  for (std::size_t i = 0; i < m; ++i) {
    s[i] = data[2 * i];
    d[i] = data[2 * i + 1];
  }
  {
    py::gil_scoped_release nogil;
    velographx::vgx_check(vgx_graph_apply(graph.handle(), s.data(), d.data(), a.data(), m));
  }
  return graph;
}

// bindings.cpp:26-32
DynamicGraph graph_from_scipy_csr(const py::object& csr, bool directed) {
  if (!py::hasattr(csr, "indptr") || !py::hasattr(csr, "indices") || !py::hasattr(csr, "shape"))
    throw std::invalid_argument("expected a SciPy CSR-like object with indptr, indices, and shape");
  const auto shape = csr.attr("shape").cast<py::tuple>();
  if (shape.size() != 2 || shape[0].cast<py::ssize_t>() != shape[1].cast<py::ssize_t>())
    throw std::invalid_argument("CSR graph matrix must be square");
  const auto vertices = static_cast<std::size_t>(shape[0].cast<py::ssize_t>());
  auto indptr = py::array_t<std::int64_t, py::array::c_style | py::array::forcecast>(csr.attr("indptr"));
  auto indices = py::array_t<std::int64_t, py::array::c_style | py::array::forcecast>(csr.attr("indices"));
  const auto p = indptr.unchecked<1>();
  const auto idx = indices.unchecked<1>();
  if (static_cast<std::size_t>(p.shape(0)) != vertices + 1) throw std::invalid_argument("invalid CSR indptr length");
  DynamicGraph graph(vertices, directed);
  std::vector<std::uint32_t> s, d;
  for (std::size_t u = 0; u < vertices; ++u) {
    const auto begin = p(static_cast<py::ssize_t>(u)), end = p(static_cast<py::ssize_t>(u + 1));
    if (begin < 0 || end < begin || end > idx.shape(0)) throw std::invalid_argument("invalid CSR offsets");
    for (auto pos = begin; pos < end; ++pos) {
      const auto v = idx(static_cast<py::ssize_t>(pos));
      if (v < 0 || static_cast<std::size_t>(v) >= vertices)
        throw std::invalid_argument("CSR contains out-of-range vertex id");
      if (!directed && static_cast<std::size_t>(v) < u) continue;
      s.push_back(static_cast<std::uint32_t>(u));
      d.push_back(static_cast<std::uint32_t>(v));
    }
  }
  std::vector<std::uint8_t> a(s.size(), 1);
  {
    py::gil_scoped_release nogil;
    velographx::vgx_check(vgx_graph_apply(graph.handle(), s.data(), d.data(), a.data(), s.size()));
  }
  return graph;
}

py::array arrow_column_to_numpy(const py::object& column) {
  if (!py::hasattr(column, "to_numpy")) throw std::invalid_argument("expected an Arrow-like column exposing to_numpy()");
  return py::array::ensure(column.attr("to_numpy")(py::arg("zero_copy_only") = false));
}

// bindings.cpp:34
DynamicGraph graph_from_arrow_columns(const py::object& src_column, const py::object& dst_column, bool directed) {
  U32Array src(arrow_column_to_numpy(src_column));
  U32Array dst(arrow_column_to_numpy(dst_column));
  const auto s = src.unchecked<1>();
  const auto d = dst.unchecked<1>();
  if (s.shape(0) != d.shape(0)) throw std::invalid_argument("Arrow source and destination columns must have equal length");
  DynamicGraph graph(0, directed);
  const auto m = static_cast<std::size_t>(s.shape(0));
  std::vector<std::uint32_t> sv(m), dv(m);
  for (std::size_t i = 0; i < m; ++i) {
    sv[i] = s(static_cast<py::ssize_t>(i));
    dv[i] = d(static_cast<py::ssize_t>(i));
  }
  std::vector<std::uint8_t> a(m, 1);
  velographx::vgx_check(vgx_graph_apply(graph.handle(), sv.data(), dv.data(), a.data(), m));
  return graph;
}

// bindings.cpp:35
DynamicGraph graph_from_arrow_table(const py::object& table, const std::string& src_name, const std::string& dst_name,
                                    bool directed) {
  if (!py::hasattr(table, "column")) throw std::invalid_argument("expected an Arrow-like table exposing column(name)");
  return graph_from_arrow_columns(table.attr("column")(src_name), table.attr("column")(dst_name), directed);
}

template <class T>
py::array_t<T> to_numpy(std::vector<T>&& v) {
  auto* heap = new std::vector<T>(std::move(v));
  py::capsule owner(heap, [](void* p) { delete static_cast<std::vector<T>*>(p); });
  return py::array_t<T>(static_cast<py::ssize_t>(heap->size()), heap->data(), owner);
}

// DLPack (the published v0 ABI of dlpack.h: DLDevice, DLDataType, DLTensor,
// DLManagedTensor) for zero-copy views of device arrays, consumed by
// torch.from_dlpack.  The view keeps the owning Python object alive until
// the consumer releases it, and holds the engine's view guard on the array
// (vgx_bfs_hold_view / vgx_pagerank_hold_view) until then.
struct DLDevice {
  std::int32_t device_type;  // 10 = kDLROCM
  std::int32_t device_id;
};
struct DLDataType {
  std::uint8_t code;  // 0 int, 1 uint, 2 float
  std::uint8_t bits;
  std::uint16_t lanes;
};
struct DLTensor {
  void* data;
  DLDevice device;
  std::int32_t ndim;
  DLDataType dtype;
  std::int64_t* shape;
  std::int64_t* strides;
  std::uint64_t byte_offset;
};
struct DLManagedTensor {
  DLTensor dl_tensor;
  void* manager_ctx;
  void (*deleter)(DLManagedTensor*);
};

struct DlpackCtx {
  std::int64_t shape[1];
  py::object owner;
  void (*release)(void*);  // drops the view's hold on the engine's array
  void* handle;
};

void dlpack_delete(DLManagedTensor* t) {
  auto* ctx = static_cast<DlpackCtx*>(t->manager_ctx);
  {
    py::gil_scoped_acquire gil;
    if (ctx->release) ctx->release(ctx->handle);
    delete ctx;
  }
  delete t;
}

// A 1-D view of `count` elements at device pointer `data` on `device`; the
// caller has taken the hold that `release(handle)` gives back.
py::capsule dlpack_view(void* data, std::uint64_t count, int device, DLDataType dtype, py::object owner,
                        void (*release)(void*), void* handle) {
  auto* ctx = new DlpackCtx{{static_cast<std::int64_t>(count)}, std::move(owner), release, handle};
  auto* t = new DLManagedTensor{};
  t->dl_tensor.data = data;
  t->dl_tensor.device = DLDevice{10, device};
  t->dl_tensor.ndim = 1;
  t->dl_tensor.dtype = dtype;
  t->dl_tensor.shape = ctx->shape;
  t->dl_tensor.strides = nullptr;
  t->dl_tensor.byte_offset = 0;
  t->manager_ctx = ctx;
  t->deleter = dlpack_delete;
  // An unconsumed capsule frees the view itself; a consumer renames it
  // "used_dltensor" and calls the deleter when done.
  return py::capsule(t, "dltensor", [](PyObject* cap) {
    if (PyCapsule_IsValid(cap, "dltensor")) {
      auto* m = static_cast<DLManagedTensor*>(PyCapsule_GetPointer(cap, "dltensor"));
      if (m && m->deleter) m->deleter(m);
    }
  });
}

int device_of(vgx_graph* g) {
  int dev = 0;
  velographx::vgx_check(vgx_graph_device(g, &dev));
  return dev;
}

}  // namespace

py::dict push_stats_dict(const vgx_pagerank_push_stats& st) {
  py::dict d;
  d["touched_rows"] = st.touched_rows;
  d["rounds"] = st.rounds;
  d["pushed_vertices"] = st.pushed_vertices;
  d["pushed_arcs"] = st.pushed_arcs;
  d["fallback_sweeps"] = st.fallback_sweeps;
  d["used_fallback"] = st.used_fallback != 0;
  d["skipped_push"] = st.skipped_push != 0;
  d["residual_l1"] = st.residual_l1;
  d["threshold"] = st.threshold;
  return d;
}

PYBIND11_MODULE(_velographx, m) {
  m.doc() = "VeloGraphX traversal path on AMD Instinct MI355X (gfx950)";

  py::class_<UpdateBatch>(m, "UpdateBatch")
      .def(py::init<>())
      .def("add", &UpdateBatch::add, py::arg("u"), py::arg("v"), py::arg("timestamp") = 0)
      .def("remove", &UpdateBatch::remove, py::arg("u"), py::arg("v"), py::arg("timestamp") = 0)
      // additions: size and bulk construction from arrays
      .def("__len__", [](const UpdateBatch& b) { return b.updates.size(); })
      .def_static(
          "from_arrays",
          [](const U32Array& src, const U32Array& dst, const py::array_t<std::uint8_t, py::array::forcecast>& add) {
            const auto s = src.unchecked<1>();
            const auto d = dst.unchecked<1>();
            const auto a = add.unchecked<1>();
            if (s.shape(0) != d.shape(0) || s.shape(0) != a.shape(0))
              throw std::invalid_argument("src, dst and add must have equal length");
            UpdateBatch b;
            b.updates.reserve(static_cast<std::size_t>(s.shape(0)));
            for (py::ssize_t i = 0; i < s.shape(0); ++i) b.updates.push_back({s(i), d(i), a(i) != 0, 0});
            return b;
          },
          py::arg("src"), py::arg("dst"), py::arg("add"));

  py::class_<DynamicGraph>(m, "Graph")
      .def(py::init<std::size_t, bool>(), py::arg("vertices") = 0, py::arg("directed") = false)
      .def("add_edge", &DynamicGraph::add_edge)
      .def("remove_edge", &DynamicGraph::remove_edge)
      .def("apply", &DynamicGraph::apply, py::call_guard<py::gil_scoped_release>())
      .def("neighbors", &DynamicGraph::neighbors)
      .def_property_readonly("version", &DynamicGraph::version)
      .def_property_readonly("vertex_count", &DynamicGraph::vertex_count)
      .def("compact", &DynamicGraph::compact)
      // additions
      .def_property_readonly("directed", &DynamicGraph::directed)
      .def_property_readonly("edge_count_directed", &DynamicGraph::edge_count_directed)
      .def("in_neighbors", &DynamicGraph::in_neighbors)
      .def("has_edge", &DynamicGraph::has_edge)
      .def("degrees", [](const DynamicGraph& g, bool in_dir) { return to_numpy(g.degrees(in_dir)); },
           py::arg("in_direction") = false)
      .def("bulk_load_edges",
           [](DynamicGraph& g, const U32Array& edges) {
             const auto info = edges.request();
             if (info.ndim != 2 || info.shape[1] != 2) throw std::invalid_argument("edge array must have shape (N, 2)");
             const auto m2 = static_cast<std::size_t>(info.shape[0]);
             const auto* data = static_cast<const std::uint32_t*>(info.ptr);
             py::gil_scoped_release nogil;
             g.bulk_load_pairs(data, m2);
           },
           py::arg("edges"))
      .def("bulk_load_arrays",
           [](DynamicGraph& g, const U32Array& src, const U32Array& dst) {
             if (src.size() != dst.size()) throw std::invalid_argument("src and dst must have equal length");
             const auto* s = src.data();
             const auto* d = dst.data();
             const auto m2 = static_cast<std::size_t>(src.size());
             py::gil_scoped_release nogil;
             g.bulk_load_edges(s, d, m2);
           },
           py::arg("src"), py::arg("dst"))
      .def("apply_arrays",
           [](DynamicGraph& g, const U32Array& src, const U32Array& dst,
              const py::array_t<std::uint8_t, py::array::c_style | py::array::forcecast>& add) {
             if (src.size() != dst.size() || src.size() != add.size())
               throw std::invalid_argument("src, dst and add must have equal length");
             const auto* s = src.data();
             const auto* d = dst.data();
             const auto* a = add.data();
             const auto k = static_cast<std::size_t>(src.size());
             py::gil_scoped_release nogil;
             velographx::vgx_check(vgx_graph_apply(g.handle(), s, d, a, k));
           },
           py::arg("src"), py::arg("dst"), py::arg("add"))
      .def("export_csr",
           [](const DynamicGraph& g, bool in_dir) {
             std::vector<std::uint64_t> off(g.vertex_count() + 1);
             std::vector<std::uint32_t> adj(g.edge_count_directed());
             velographx::vgx_check(vgx_graph_export_csr(g.handle(), in_dir ? 1 : 0, off.data(), adj.data()));
             return py::make_tuple(to_numpy(std::move(off)), to_numpy(std::move(adj)));
           },
           py::arg("in_direction") = false)
      .def("storage",
           [](const DynamicGraph& g) {
             auto s = g.storage();
             py::dict d;
             d["arena_entries"] = s.arena_entries;
             d["arena_used"] = s.arena_used;
             d["tombstones"] = s.tombstones;
             d["storage_bytes"] = s.storage_bytes;
             d["relocated_rows"] = s.relocated_rows;
             return d;
           })
      .def("storage_bytes", &DynamicGraph::storage_bytes)
      .def(
          "_apply_device",
          [](DynamicGraph& g, std::uintptr_t src, std::uintptr_t dst, std::uintptr_t add, std::uint64_t k,
             std::uintptr_t stream) {
            py::gil_scoped_release nogil;
            velographx::vgx_check(vgx_graph_apply_device(g.handle(), reinterpret_cast<const std::uint32_t*>(src),
                                                         reinterpret_cast<const std::uint32_t*>(dst),
                                                         reinterpret_cast<const std::uint8_t*>(add), k,
                                                         reinterpret_cast<void*>(stream)));
          },
          py::arg("src_ptr"), py::arg("dst_ptr"), py::arg("add_ptr"), py::arg("k"), py::arg("stream") = 0)
      .def(
          "_bulk_load_device",
          [](DynamicGraph& g, std::uintptr_t src, std::uintptr_t dst, std::uint64_t m, std::uint64_t stride,
             std::uint32_t id_bytes, std::uintptr_t stream) {
            py::gil_scoped_release nogil;
            velographx::vgx_check(vgx_graph_bulk_load_device(g.handle(), reinterpret_cast<const void*>(src),
                                                             reinterpret_cast<const void*>(dst), m, stride, id_bytes,
                                                             reinterpret_cast<void*>(stream)));
          },
          py::arg("src_ptr"), py::arg("dst_ptr"), py::arg("m"), py::arg("stride"), py::arg("id_bytes"),
          py::arg("stream") = 0)
      .def(
          "_bulk_load_csr_device",
          [](DynamicGraph& g, std::uintptr_t indptr, std::uintptr_t indices, std::uint64_t vertices,
             std::uint64_t nnz, std::uint32_t index_bytes, std::uintptr_t stream) {
            py::gil_scoped_release nogil;
            velographx::vgx_check(vgx_graph_bulk_load_csr_device(
                g.handle(), reinterpret_cast<const void*>(indptr), reinterpret_cast<const void*>(indices), vertices,
                nnz, index_bytes, reinterpret_cast<void*>(stream)));
          },
          py::arg("indptr_ptr"), py::arg("indices_ptr"), py::arg("vertices"), py::arg("nnz"),
          py::arg("index_bytes"), py::arg("stream") = 0)
      .def_property_readonly("device", [](const DynamicGraph& g) { return device_of(g.handle()); })
      .def("is_compact", &DynamicGraph::is_compact)
      .def("ensure_vertex", &DynamicGraph::ensure_vertex, py::arg("v"))
      .def_property_readonly("base_edge_count_directed", &DynamicGraph::base_edge_count_directed)
      .def_property_readonly("delta_edge_count", &DynamicGraph::delta_edge_count)
      .def_property_readonly("delta_ratio", &DynamicGraph::delta_ratio)
      .def("dirty_out_segment_count", &DynamicGraph::dirty_out_segment_count)
      .def("dirty_in_segment_count", &DynamicGraph::dirty_in_segment_count)
      .def("maybe_compact", &DynamicGraph::maybe_compact, py::arg("threshold") = 0.25,
           py::call_guard<py::gil_scoped_release>())
      .def(
          "probe_neighbors",
          [](const DynamicGraph& g, std::size_t count, std::uint32_t repeats) {
            velographx::NeighborProbe p;
            {
              py::gil_scoped_release nogil;
              p = velographx::probe_neighbors(g, count, repeats);
            }
            return py::make_tuple(p.ns_per_probe, p.digest);
          },
          py::arg("count") = 4096, py::arg("repeats") = 5)
      .def_property_readonly("_handle", [](const DynamicGraph& g) { return reinterpret_cast<std::uintptr_t>(g.handle()); });

  // incremental/sssp.hpp (unweighted SSSP on the BFS engine)
  py::class_<velographx::IncrementalSSSP>(m, "IncrementalSSSP")
      .def(py::init<DynamicGraph&, VertexId>(), py::arg("graph"), py::arg("source"), py::keep_alive<1, 2>())
      .def_property_readonly("distances", &velographx::IncrementalSSSP::distances)
      .def("distances_numpy", [](const velographx::IncrementalSSSP& x) { return to_numpy(x.distances()); })
      .def("apply", &velographx::IncrementalSSSP::apply, py::call_guard<py::gil_scoped_release>())
      .def("recompute", &velographx::IncrementalSSSP::recompute, py::call_guard<py::gil_scoped_release>())
      .def(
          "apply_arrays",
          [](velographx::IncrementalSSSP& x, const U32Array& src, const U32Array& dst,
             const py::array_t<std::uint8_t, py::array::c_style | py::array::forcecast>& add) {
            if (src.size() != dst.size() || src.size() != add.size())
              throw std::invalid_argument("src, dst and add must have equal length");
            const auto* s = src.data();
            const auto* d = dst.data();
            const auto* a = add.data();
            const auto k = static_cast<std::size_t>(src.size());
            py::gil_scoped_release nogil;
            x.apply_arrays(s, d, a, k);
          },
          py::arg("src"), py::arg("dst"), py::arg("add"))
      .def_property_readonly_static("infinity", [](py::object) { return velographx::IncrementalSSSP::infinity; });

  // incremental/connected_components.hpp (labels: smallest id per component)
  py::class_<velographx::IncrementalComponents>(m, "IncrementalComponents")
      .def(py::init<DynamicGraph&>(), py::arg("graph"), py::keep_alive<1, 2>())
      .def("component", &velographx::IncrementalComponents::component, py::arg("v"))
      .def("labels_numpy", [](const velographx::IncrementalComponents& x) { return to_numpy(x.labels()); })
      .def_property_readonly("last_repaired_vertices", &velographx::IncrementalComponents::last_repaired_vertices)
      .def("apply", &velographx::IncrementalComponents::apply, py::call_guard<py::gil_scoped_release>())
      .def("recompute", &velographx::IncrementalComponents::recompute, py::call_guard<py::gil_scoped_release>())
      .def(
          "apply_arrays",
          [](velographx::IncrementalComponents& x, const U32Array& src, const U32Array& dst,
             const py::array_t<std::uint8_t, py::array::c_style | py::array::forcecast>& add) {
            if (src.size() != dst.size() || src.size() != add.size())
              throw std::invalid_argument("src, dst and add must have equal length");
            const auto* s = src.data();
            const auto* d = dst.data();
            const auto* a = add.data();
            const auto k = static_cast<std::size_t>(src.size());
            py::gil_scoped_release nogil;
            x.apply_arrays(s, d, a, k);
          },
          py::arg("src"), py::arg("dst"), py::arg("add"))
      .def("stats", [](const velographx::IncrementalComponents& x) {
        std::uint64_t r = 0;
        std::uint32_t rounds = 0;
        velographx::vgx_check(vgx_components_stats(x.handle(), &r, &rounds));
        py::dict d;
        d["last_repaired_vertices"] = r;
        d["rounds"] = rounds;
        return d;
      });

  // incremental/kcore.hpp (core numbers by device peeling)
  py::class_<velographx::IncrementalKCore>(m, "IncrementalKCore")
      .def(py::init<DynamicGraph&>(), py::arg("graph"), py::keep_alive<1, 2>())
      .def_property_readonly("core", &velographx::IncrementalKCore::core)
      .def("core_numpy", [](const velographx::IncrementalKCore& x) { return to_numpy(x.core()); })
      .def_property_readonly("last_repaired_vertices", &velographx::IncrementalKCore::last_repaired_vertices)
      .def("apply", &velographx::IncrementalKCore::apply, py::call_guard<py::gil_scoped_release>())
      .def("recompute", &velographx::IncrementalKCore::recompute, py::call_guard<py::gil_scoped_release>())
      .def(
          "apply_arrays",
          [](velographx::IncrementalKCore& x, const U32Array& src, const U32Array& dst,
             const py::array_t<std::uint8_t, py::array::c_style | py::array::forcecast>& add) {
            if (src.size() != dst.size() || src.size() != add.size())
              throw std::invalid_argument("src, dst and add must have equal length");
            const auto* s = src.data();
            const auto* d = dst.data();
            const auto* a = add.data();
            const auto k = static_cast<std::size_t>(src.size());
            py::gil_scoped_release nogil;
            x.apply_arrays(s, d, a, k);
          },
          py::arg("src"), py::arg("dst"), py::arg("add"))
      .def("stats", [](const velographx::IncrementalKCore& x) {
        std::uint64_t r = 0;
        std::uint32_t levels = 0, rounds = 0;
        velographx::vgx_check(vgx_kcore_stats(x.handle(), &r, &levels, &rounds));
        py::dict d;
        d["last_repaired_vertices"] = r;
        d["levels"] = levels;
        d["rounds"] = rounds;
        return d;
      });

  // csr_graph.hpp, algorithms.hpp:8-10 (static graph, one-shot algorithms)
  py::class_<velographx::CsrGraph>(m, "CsrGraph")
      .def(py::init([](const U32Array& edges, bool directed) {
             const auto info = edges.request();
             if (info.size == 0) return std::make_unique<velographx::CsrGraph>(std::vector<velographx::CsrGraph::Edge>{}, directed);
             if (info.ndim != 2 || info.shape[1] != 2) throw std::invalid_argument("edge array must have shape (N, 2)");
             const auto m = static_cast<std::size_t>(info.shape[0]);
             const auto* data = static_cast<const std::uint32_t*>(info.ptr);
             py::gil_scoped_release nogil;
             return std::make_unique<velographx::CsrGraph>(velographx::CsrGraph::Pairs{}, data, m, directed);
           }),
           py::arg("edges"), py::arg("directed") = false)
      .def("vertex_count", &velographx::CsrGraph::vertex_count)
      .def("edge_entry_count", &velographx::CsrGraph::edge_entry_count)
      .def("directed", &velographx::CsrGraph::directed)
      .def("neighbors", &velographx::CsrGraph::neighbors, py::arg("v"))
      .def("in_neighbors", &velographx::CsrGraph::in_neighbors, py::arg("v"))
      .def("degree", &velographx::CsrGraph::degree, py::arg("v"))
      .def("has_edge", &velographx::CsrGraph::has_edge, py::arg("u"), py::arg("v"));
  m.def(
      "bfs_distances",
      [](const velographx::CsrGraph& g, VertexId source) {
        // algorithms.cpp:11-27; copied from the device straight into the
        // returned array
        if (source >= g.vertex_count()) throw std::out_of_range("source outside graph");
        py::array_t<std::uint32_t> out(static_cast<py::ssize_t>(g.vertex_count()));
        std::uint32_t* p = out.mutable_data();
        {
          py::gil_scoped_release nogil;
          g.bfs_distances_into(source, p);
        }
        return out;
      },
      py::arg("graph"), py::arg("source"));
  m.def(
      "pagerank",
      [](const velographx::CsrGraph& g, double damping, std::size_t max_iterations, double tolerance) {
        std::vector<double> r;
        {
          py::gil_scoped_release nogil;
          r = velographx::pagerank(g, damping, max_iterations, tolerance);
        }
        return to_numpy(std::move(r));
      },
      py::arg("graph"), py::arg("damping") = 0.85, py::arg("max_iterations") = 100, py::arg("tolerance") = 1e-10);

  // storage/consolidation.hpp:10-163
  py::class_<velographx::ConsolidationPolicy>(m, "ConsolidationPolicy")
      .def(py::init<>())
      .def(py::init([](double storage, double latency) { return velographx::ConsolidationPolicy{storage, latency}; }),
           py::arg("max_storage_growth_ratio"), py::arg("max_neighbor_latency_ratio"))
      .def_readwrite("max_storage_growth_ratio", &velographx::ConsolidationPolicy::max_storage_growth_ratio)
      .def_readwrite("max_neighbor_latency_ratio", &velographx::ConsolidationPolicy::max_neighbor_latency_ratio);
  m.def("scale_aware_consolidation_policy", &velographx::scale_aware_consolidation_policy,
        py::arg("directed_edges"), py::arg("latency_ratio") = 1.25);
  py::class_<velographx::ConsolidationSignal>(m, "ConsolidationSignal")
      .def(py::init<>())
      .def(py::init([](double storage, double latency, bool storage_ex, bool latency_ex, bool should) {
             return velographx::ConsolidationSignal{storage, latency, storage_ex, latency_ex, should};
           }),
           py::arg("storage_growth_ratio"), py::arg("neighbor_latency_ratio"), py::arg("storage_limit_exceeded"),
           py::arg("latency_limit_exceeded"), py::arg("should_consolidate"))
      .def_readwrite("storage_growth_ratio", &velographx::ConsolidationSignal::storage_growth_ratio)
      .def_readwrite("neighbor_latency_ratio", &velographx::ConsolidationSignal::neighbor_latency_ratio)
      .def_readwrite("storage_limit_exceeded", &velographx::ConsolidationSignal::storage_limit_exceeded)
      .def_readwrite("latency_limit_exceeded", &velographx::ConsolidationSignal::latency_limit_exceeded)
      .def_readwrite("should_consolidate", &velographx::ConsolidationSignal::should_consolidate);
  m.def("evaluate_consolidation", &velographx::evaluate_consolidation, py::arg("current_storage_bytes"),
        py::arg("canonical_storage_bytes"), py::arg("current_neighbor_latency"),
        py::arg("canonical_neighbor_latency"), py::arg("policy") = velographx::ConsolidationPolicy{});
  py::class_<velographx::ConsolidationControllerConfig>(m, "ConsolidationControllerConfig")
      .def(py::init<>())
      .def(py::init([](std::size_t min_epochs, std::size_t samples, double growth) {
             return velographx::ConsolidationControllerConfig{min_epochs, samples, growth};
           }),
           py::arg("min_epochs_between_consolidations"), py::arg("latency_breach_samples"),
           py::arg("min_storage_growth_for_latency_trigger"))
      .def_readwrite("min_epochs_between_consolidations",
                     &velographx::ConsolidationControllerConfig::min_epochs_between_consolidations)
      .def_readwrite("latency_breach_samples", &velographx::ConsolidationControllerConfig::latency_breach_samples)
      .def_readwrite("min_storage_growth_for_latency_trigger",
                     &velographx::ConsolidationControllerConfig::min_storage_growth_for_latency_trigger);
  py::class_<velographx::ConsolidationController>(m, "ConsolidationController")
      .def(py::init<velographx::ConsolidationControllerConfig>(),
           py::arg("config") = velographx::ConsolidationControllerConfig{})
      .def("observe", &velographx::ConsolidationController::observe, py::arg("signal"), py::arg("epoch"))
      .def("mark_consolidated", &velographx::ConsolidationController::mark_consolidated, py::arg("epoch"))
      .def("latency_breach_streak", &velographx::ConsolidationController::latency_breach_streak)
      .def("last_consolidation_epoch", &velographx::ConsolidationController::last_consolidation_epoch);
  py::class_<velographx::ConsolidationSnapshot>(m, "ConsolidationSnapshot")
      .def_property_readonly(
          "graph", [](velographx::ConsolidationSnapshot& s) -> DynamicGraph& { return s.graph; },
          py::return_value_policy::reference_internal)
      .def_readonly("source_storage_bytes", &velographx::ConsolidationSnapshot::source_storage_bytes)
      .def_readonly("consolidated_storage_bytes", &velographx::ConsolidationSnapshot::consolidated_storage_bytes)
      .def_readonly("directed_edges", &velographx::ConsolidationSnapshot::directed_edges);
  m.def(
      "consolidate_to_csr_snapshot",
      [](const DynamicGraph& g) {
        py::gil_scoped_release nogil;
        return velographx::consolidate_to_csr_snapshot(g);
      },
      py::arg("source"));

  py::class_<PageRankValidation>(m, "PageRankValidation")
      .def_readonly("l1_error", &PageRankValidation::l1_error)
      .def_readonly("linf_error", &PageRankValidation::linf_error)
      .def_readonly("local_residual_l1", &PageRankValidation::local_residual_l1)
      .def_readonly("local_residual_linf", &PageRankValidation::local_residual_linf)
      .def_readonly("reference_residual_l1", &PageRankValidation::reference_residual_l1)
      .def_readonly("reference_residual_linf", &PageRankValidation::reference_residual_linf)
      .def_readonly("reference_iterations", &PageRankValidation::reference_iterations)
      .def_readonly("reference_converged", &PageRankValidation::reference_converged)
      .def_readonly("within_tolerance", &PageRankValidation::within_tolerance)
      .def_readonly("fallback_applied", &PageRankValidation::fallback_applied);

  py::class_<IncrementalBFS>(m, "IncrementalBFS")
      .def(py::init<DynamicGraph&, VertexId, double>(), py::arg("graph"), py::arg("source"),
           py::arg("deletion_fallback_fraction") = 0.35, py::keep_alive<1, 2>())
      .def_property_readonly("distances", [](const IncrementalBFS& x) { return x.distances(); })
      .def("apply", &IncrementalBFS::apply, py::call_guard<py::gil_scoped_release>())
      .def("recompute", &IncrementalBFS::recompute, py::call_guard<py::gil_scoped_release>())
      .def_property_readonly_static("unreachable", [](py::object) { return IncrementalBFS::unreachable; })
      // additions: diagnostics of bfs.hpp:24-27 and a zero-list-copy view
      .def_property_readonly("reachable_count", &IncrementalBFS::reachable_count)
      .def_property_readonly("last_deletion_candidates", &IncrementalBFS::last_deletion_candidates)
      .def_property_readonly("last_affected_vertices", &IncrementalBFS::last_affected_vertices)
      .def_property_readonly("last_used_full_recompute", &IncrementalBFS::last_used_full_recompute)
      .def("distances_numpy", [](const IncrementalBFS& x) { return to_numpy(x.distances()); })
      .def(
          "distances_dlpack",
          [](py::object self, std::uintptr_t stream) {
            const auto& b = self.cast<const IncrementalBFS&>();
            const std::uint32_t* dev = nullptr;
            std::uint64_t n = 0;
            velographx::vgx_check(vgx_bfs_distances_device(b.handle(), &dev, &n));
            vgx_graph* g = nullptr;
            velographx::vgx_check(vgx_bfs_graph(b.handle(), &g));
            velographx::vgx_check(vgx_bfs_hold_view(b.handle(), 1, reinterpret_cast<void*>(stream)));
            return dlpack_view(
                const_cast<std::uint32_t*>(dev), n, device_of(g), DLDataType{1, 32, 1}, self,
                [](void* h) { (void)vgx_bfs_hold_view(static_cast<vgx_bfs*>(h), -1, nullptr); }, b.handle());
          },
          py::arg("stream") = 0,
          "Zero-copy DLPack view (uint32, on the device) of the current distances, read on `stream` "
          "(a hipStream_t; 0: the default stream).  Its contents are those of the current distances "
          "until the object's next apply or recompute, which first waits on the device for the "
          "work queued on `stream` so far; the memory stays valid while the view lives.")
      .def(
          "_apply_device",
          [](IncrementalBFS& b, std::uintptr_t src, std::uintptr_t dst, std::uintptr_t add, std::uint64_t k,
             std::uintptr_t stream) {
            py::gil_scoped_release nogil;
            velographx::vgx_check(vgx_bfs_apply_device(b.handle(), reinterpret_cast<const std::uint32_t*>(src),
                                                       reinterpret_cast<const std::uint32_t*>(dst),
                                                       reinterpret_cast<const std::uint8_t*>(add), k,
                                                       reinterpret_cast<void*>(stream)));
          },
          py::arg("src_ptr"), py::arg("dst_ptr"), py::arg("add_ptr"), py::arg("k"), py::arg("stream") = 0)
      .def("apply_arrays",
           [](IncrementalBFS& b, const U32Array& src, const U32Array& dst,
              const py::array_t<std::uint8_t, py::array::c_style | py::array::forcecast>& add) {
             if (src.size() != dst.size() || src.size() != add.size())
               throw std::invalid_argument("src, dst and add must have equal length");
             const auto* s = src.data();
             const auto* d = dst.data();
             const auto* a = add.data();
             const auto k = static_cast<std::size_t>(src.size());
             py::gil_scoped_release nogil;
             velographx::vgx_check(vgx_bfs_apply(b.handle(), s, d, a, k));
           },
           py::arg("src"), py::arg("dst"), py::arg("add"))
      .def("stats",
           [](const IncrementalBFS& b) {
             auto s = b.stats();
             py::dict d;
             d["reachable_count"] = s.reachable_count;
             d["last_deletion_candidates"] = s.last_deletion_candidates;
             d["last_affected_vertices"] = s.last_affected_vertices;
             d["last_used_full_recompute"] = s.last_used_full_recompute != 0;
             d["levels"] = s.levels;
             d["bottom_up_levels"] = s.bottom_up_levels;
             d["last_relaxed_vertices"] = s.last_relaxed_vertices;
             d["level_launches"] = s.level_launches;
             d["level_kernel_ms"] = s.level_kernel_ms;
             d["steps"] = s.steps;
             d["traversal_ms"] = s.traversal_ms;
             d["traversals"] = s.traversals;
             return d;
           })
      .def("set_profiling",
           [](IncrementalBFS& b, int mode) { velographx::vgx_check(vgx_bfs_set_profiling(b.handle(), mode)); },
           py::arg("mode") = 1)
      .def("set_tuning",
           [](IncrementalBFS& b, double alpha, double beta, std::uint32_t blocks, std::uint32_t levels_per_sync,
              double tail_ratio, double claim_ratio) {
             vgx_bfs_tuning t{alpha, beta, blocks, levels_per_sync, tail_ratio, claim_ratio};
             velographx::vgx_check(vgx_bfs_set_tuning(b.handle(), &t));
           },
           py::arg("alpha") = 15.0, py::arg("beta") = 18.0, py::arg("blocks") = 0, py::arg("levels_per_sync") = 0,
           py::arg("tail_ratio") = 32.0, py::arg("claim_ratio") = 5.0)
      .def(
          "set_source",
          [](IncrementalBFS& b, std::uint32_t source) {
            py::gil_scoped_release nogil;
            velographx::vgx_check(vgx_bfs_set_source(b.handle(), source));
          },
          py::arg("source"))
      .def("get_tuning",
           [](const IncrementalBFS& b) {
             vgx_bfs_tuning t{};
             velographx::vgx_check(vgx_bfs_get_tuning(b.handle(), &t));
             py::dict d;
             d["alpha"] = t.alpha;
             d["beta"] = t.beta;
             d["blocks"] = t.blocks;
             d["levels_per_sync"] = t.levels_per_sync;
             d["tail_ratio"] = t.tail_ratio;
             d["claim_ratio"] = t.claim_ratio;
             return d;
           })
      .def("level_log",
           [](const IncrementalBFS& b) {
             vgx_bfs_level lv[64];
             std::uint32_t cnt = 0;
             velographx::vgx_check(vgx_bfs_level_log(b.handle(), lv, 64, &cnt));
             py::list out;
             for (std::uint32_t i = 0; i < cnt && i < 64; ++i) {
               py::dict d;
               d["bottom_up"] = lv[i].bottom_up == 1;
               d["mode"] = lv[i].bottom_up == 1 ? "bottom-up" : (lv[i].bottom_up == 2 ? "top-down-bitmap" : "top-down");
               d["frontier_vertices"] = lv[i].frontier_vertices;
               d["frontier_edges"] = lv[i].frontier_edges;
               d["pending_vertices"] = lv[i].pending_vertices;
               d["pending_edges"] = lv[i].pending_edges;
               d["kernel_ms"] = lv[i].kernel_ms;
               out.append(d);
             }
             return out;
           })
      .def(
          "autotune",
          [](IncrementalBFS& b, std::uint32_t reps) {
            vgx_bfs_tuning t{};
            double ms = 0.0;
            {
              py::gil_scoped_release nogil;
              velographx::vgx_check(vgx_bfs_autotune(b.handle(), reps, &t, &ms));
            }
            py::dict d;
            d["alpha"] = t.alpha;
            d["beta"] = t.beta;
            d["blocks"] = t.blocks;
            d["levels_per_sync"] = t.levels_per_sync;
            d["tail_ratio"] = t.tail_ratio;
            d["claim_ratio"] = t.claim_ratio;
            d["ms"] = ms;
            return d;
          },
          py::arg("reps") = 3)
      .def_property_readonly("_handle", [](const IncrementalBFS& b) { return reinterpret_cast<std::uintptr_t>(b.handle()); });

  py::class_<IncrementalPageRank>(m, "IncrementalPageRank")
      .def(py::init<DynamicGraph&, double, int>(), py::arg("graph"), py::arg("damping") = 0.85,
           py::arg("precision") = 64, py::keep_alive<1, 2>())
      .def_property_readonly("values", [](const IncrementalPageRank& x) { return x.values(); })
      .def(
          "values_dlpack",
          [](py::object self, std::uintptr_t stream) {
            const auto& pr = self.cast<const IncrementalPageRank&>();
            const void* dev = nullptr;
            std::uint64_t n = 0;
            int precision = 64;
            velographx::vgx_check(vgx_pagerank_values_device(pr.handle(), &dev, &n, &precision));
            vgx_graph* g = nullptr;
            velographx::vgx_check(vgx_pagerank_graph(pr.handle(), &g));
            const DLDataType dt{2, static_cast<std::uint8_t>(precision), 1};
            velographx::vgx_check(vgx_pagerank_hold_view(pr.handle(), 1, reinterpret_cast<void*>(stream)));
            return dlpack_view(
                const_cast<void*>(dev), n, device_of(g), dt, self,
                [](void* h) { (void)vgx_pagerank_hold_view(static_cast<vgx_pagerank*>(h), -1, nullptr); },
                pr.handle());
          },
          py::arg("stream") = 0,
          "Zero-copy DLPack view (float64, or float32 at precision 32, on the device) of the current "
          "ranks, read on `stream` (0: the default stream); as distances_dlpack.")
      .def("apply", &IncrementalPageRank::apply, py::arg("batch"), py::arg("local_iterations") = 24,
           py::arg("tol") = 1e-9, py::arg("full_fallback_fraction") = 0.60, py::call_guard<py::gil_scoped_release>())
      .def("recompute", &IncrementalPageRank::recompute, py::arg("max_iterations") = 200, py::arg("tol") = 1e-12,
           py::call_guard<py::gil_scoped_release>())
      .def("validate_against_full", &IncrementalPageRank::validate_against_full,
           py::arg("reference_max_iterations") = 500, py::arg("reference_tol") = 1e-12, py::arg("l1_tolerance") = 1e-6,
           py::arg("linf_tolerance") = 1e-7, py::call_guard<py::gil_scoped_release>())
      .def("apply_validated", &IncrementalPageRank::apply_validated, py::arg("batch"), py::arg("local_iterations") = 64,
           py::arg("local_tol") = 1e-10, py::arg("full_fallback_fraction") = 0.95,
           py::arg("reference_max_iterations") = 500, py::arg("reference_tol") = 1e-12, py::arg("l1_tolerance") = 1e-6,
           py::arg("linf_tolerance") = 1e-7, py::call_guard<py::gil_scoped_release>())
      .def_property_readonly("last_repaired_vertices", &IncrementalPageRank::last_repaired_vertices)
      .def_property_readonly("last_repair_iterations", &IncrementalPageRank::last_repair_iterations)
      .def_property_readonly("last_residual_l1", &IncrementalPageRank::last_residual_l1)
      .def_property_readonly("last_residual_linf", &IncrementalPageRank::last_residual_linf)
      .def_property_readonly("last_full_recompute_converged", &IncrementalPageRank::last_full_recompute_converged)
      // additions
      .def("values_numpy", [](const IncrementalPageRank& x) { return to_numpy(x.values()); })
      .def("solve_l1", &IncrementalPageRank::solve_l1, py::arg("max_iterations") = 200, py::arg("l1_tol") = 1e-7,
           py::arg("warm_start") = false, py::call_guard<py::gil_scoped_release>())
      .def_property_readonly("last_used_full_recompute", &IncrementalPageRank::last_used_full_recompute)
      .def_property_readonly("precision", &IncrementalPageRank::precision)
      .def(
          "apply_push",
          [](IncrementalPageRank& x, const UpdateBatch& batch, double l1_tol, double threshold,
             double max_push_fraction, std::size_t max_sweeps) {
            vgx_pagerank_push_stats st{};
            {
              py::gil_scoped_release nogil;
              st = x.apply_push(batch, l1_tol, threshold, max_push_fraction, max_sweeps);
            }
            return push_stats_dict(st);
          },
          py::arg("batch"), py::arg("l1_tol") = 1e-7, py::arg("threshold") = 0.0, py::arg("max_push_fraction") = 0.05,
          py::arg("max_sweeps") = 200)
      .def(
          "apply_push_arrays",
          [](IncrementalPageRank& x, const U32Array& src, const U32Array& dst,
             const py::array_t<std::uint8_t, py::array::c_style | py::array::forcecast>& add, double l1_tol,
             double threshold, double max_push_fraction, std::size_t max_sweeps) {
            if (src.size() != dst.size() || src.size() != add.size())
              throw std::invalid_argument("src, dst and add must have equal length");
            const auto* s = src.data();
            const auto* d = dst.data();
            const auto* a = add.data();
            const auto k = static_cast<std::size_t>(src.size());
            vgx_pagerank_push_stats st{};
            {
              py::gil_scoped_release nogil;
              st = x.apply_push(s, d, a, k, l1_tol, threshold, max_push_fraction, max_sweeps);
            }
            return push_stats_dict(st);
          },
          py::arg("src"), py::arg("dst"), py::arg("add"), py::arg("l1_tol") = 1e-7, py::arg("threshold") = 0.0,
          py::arg("max_push_fraction") = 0.05, py::arg("max_sweeps") = 200);

  // addition: adaptive repair-vs-recompute selection
  py::class_<velographx::AdaptivePolicy>(m, "AdaptivePolicy")
      .def(py::init<std::uint64_t, double, double>(), py::arg("vertices"), py::arg("initial_full_us"),
           py::arg("initial_reachable_fraction"))
      .def(
          "decide",
          [](const velographx::AdaptivePolicy& p, double update_fraction, double shallow, bool first_batch) {
            vgx_selector_trace t{};
            t.update_fraction = update_fraction;
            t.shallow_parent_deletion_fraction = shallow;
            t.first_batch = first_batch ? 1 : 0;
            p.decide(t);
            return trace_to_dict(t);
          },
          py::arg("update_fraction"), py::arg("shallow_parent_deletion_fraction") = 0.0,
          py::arg("first_batch") = false)
      .def(
          "observe", [](velographx::AdaptivePolicy& p, const py::dict& trace) { p.observe(trace_from_dict(trace)); },
          py::arg("trace"));
  py::class_<velographx::AdaptiveSelector>(m, "AdaptiveSelector")
      .def(py::init([](IncrementalBFS& b) { return new velographx::AdaptiveSelector(b); }), py::arg("bfs"),
           py::keep_alive<1, 2>())
      .def("apply",
           [](velographx::AdaptiveSelector& s, const UpdateBatch& batch) {
             vgx_selector_trace t;
             {
               py::gil_scoped_release nogil;
               t = s.apply(batch);
             }
             return trace_to_dict(t);
           })
      .def(
          "apply_arrays",
          [](velographx::AdaptiveSelector& s, const U32Array& src, const U32Array& dst,
             const py::array_t<std::uint8_t, py::array::c_style | py::array::forcecast>& add) {
            if (src.size() != dst.size() || src.size() != add.size())
              throw std::invalid_argument("src, dst and add must have equal length");
            const auto* sp = src.data();
            const auto* dp = dst.data();
            const auto* ap = add.data();
            const auto k = static_cast<std::size_t>(src.size());
            vgx_selector_trace t;
            {
              py::gil_scoped_release nogil;
              t = s.apply(sp, dp, ap, k);
            }
            return trace_to_dict(t);
          },
          py::arg("src"), py::arg("dst"), py::arg("add"))
      .def_property_readonly("last_trace", [](const velographx::AdaptiveSelector& s) { return trace_to_dict(s.last_trace()); });
  m.def("policy_reason_name", &velographx::policy_reason_name);

  // addition: batched multi-root BFS
  py::class_<velographx::MultiSourceBFS>(m, "MultiSourceBFS")
      .def(py::init([](DynamicGraph& g, const std::vector<VertexId>& roots) {
             return new velographx::MultiSourceBFS(g, roots);
           }),
           py::arg("graph"), py::arg("roots"), py::keep_alive<1, 2>())
      .def("run",
           [](velographx::MultiSourceBFS& m, const DynamicGraph& g) {
             const auto n = g.vertex_count();
             py::gil_scoped_release nogil;
             m.run(n);
           },
           py::arg("graph"))
      .def("distances", [](const velographx::MultiSourceBFS& m, std::uint32_t r) { return to_numpy(m.distances(r)); })
      .def("levels_u8",
           [](const velographx::MultiSourceBFS& m) {
             auto v = m.levels_u8();
             const auto k = static_cast<py::ssize_t>(m.root_count());
             auto arr = to_numpy(std::move(v));
             return arr.reshape({k, arr.size() / std::max<py::ssize_t>(k, 1)});
           })
      .def("reached", [](const velographx::MultiSourceBFS& m) { return to_numpy(m.reached()); })
      .def_property_readonly("level_count", &velographx::MultiSourceBFS::level_count)
      .def(
          "set_graph",
          [](velographx::MultiSourceBFS& m, bool on) { velographx::vgx_check(vgx_msbfs_set_graph(m.handle(), on)); },
          py::arg("on") = true)
      .def("stats",
           [](const velographx::MultiSourceBFS& m) {
             std::uint32_t pl = 0;
             std::uint64_t rep = 0;
             velographx::vgx_check(vgx_msbfs_stats(m.handle(), &pl, &rep));
             py::dict d;
             d["push_levels"] = pl;
             d["graph_replays"] = rep;
             return d;
           })
      .def(
          "set_tuning",
          [](velographx::MultiSourceBFS& m, std::uint32_t blocks, double alpha) {
            velographx::vgx_check(vgx_msbfs_set_tuning(m.handle(), blocks, alpha));
          },
          py::arg("blocks") = 0, py::arg("alpha") = 4.0)
      .def(
          "autotune",
          [](velographx::MultiSourceBFS& m, std::uint32_t reps) {
            std::uint32_t blocks = 0;
            double alpha = 0.0, ms = 0.0;
            {
              py::gil_scoped_release nogil;
              velographx::vgx_check(vgx_msbfs_autotune(m.handle(), reps, &blocks, &alpha, &ms));
            }
            vgx_msbfs_tuning t{};
            velographx::vgx_check(vgx_msbfs_get_params(m.handle(), &t));
            py::dict d;
            d["blocks"] = blocks;
            d["alpha"] = alpha;
            d["beta"] = t.beta;
            d["short_row"] = t.short_row;
            d["chunk"] = t.chunk;
            d["ms"] = ms;
            return d;
          },
          py::arg("reps") = 3)
      .def(
          "set_params",
          [](velographx::MultiSourceBFS& m, std::uint32_t blocks, double alpha, double beta, std::uint32_t short_row,
             std::uint32_t chunk) {
            const vgx_msbfs_tuning t{blocks, alpha, beta, short_row, chunk};
            velographx::vgx_check(vgx_msbfs_set_params(m.handle(), &t));
          },
          py::arg("blocks") = 0, py::arg("alpha") = 4.0, py::arg("beta") = 0.0, py::arg("short_row") = 32,
          py::arg("chunk") = 1024)
      .def("params",
           [](const velographx::MultiSourceBFS& m) {
             vgx_msbfs_tuning t{};
             velographx::vgx_check(vgx_msbfs_get_params(m.handle(), &t));
             py::dict d;
             d["blocks"] = t.blocks;
             d["alpha"] = t.alpha;
             d["beta"] = t.beta;
             d["short_row"] = t.short_row;
             d["chunk"] = t.chunk;
             return d;
           })
      .def("apply",
           [](velographx::MultiSourceBFS& m, const UpdateBatch& batch) {
             py::gil_scoped_release nogil;
             m.apply(batch);
           },
           py::arg("batch"))
      .def("apply_arrays",
           [](velographx::MultiSourceBFS& m, const U32Array& src, const U32Array& dst,
              const py::array_t<std::uint8_t, py::array::c_style | py::array::forcecast>& add) {
             if (src.size() != dst.size() || src.size() != add.size())
               throw std::invalid_argument("src, dst and add must have equal length");
             const auto* s = src.data();
             const auto* d = dst.data();
             const auto* a = add.data();
             const auto k = static_cast<std::size_t>(src.size());
             py::gil_scoped_release nogil;
             m.apply_arrays(s, d, a, k);
           },
           py::arg("src"), py::arg("dst"), py::arg("add"))
      .def(
          "set_repair",
          [](velographx::MultiSourceBFS& m, int mode, double fallback_fraction) {
            velographx::vgx_check(vgx_msbfs_set_repair(m.handle(), mode, fallback_fraction));
          },
          py::arg("mode") = 0, py::arg("fallback_fraction") = 0.1)
      .def("repair_stats", [](const velographx::MultiSourceBFS& m) {
        vgx_msbfs_repair_info r{};
        velographx::vgx_check(vgx_msbfs_repair_stats(m.handle(), &r));
        py::dict d;
        d["batches"] = r.batches;
        d["repaired"] = r.repaired;
        d["full_passes"] = r.full_passes;
        d["fallbacks"] = r.fallbacks;
        d["last_repaired"] = r.last_repaired != 0;
        d["last_fallback"] = r.last_fallback != 0;
        d["last_candidates"] = r.last_candidates;
        d["last_affected_pairs"] = r.last_affected_pairs;
        d["last_lowered"] = r.last_lowered;
        d["pass_ms"] = r.pass_ms;
        d["repair_us_per_op"] = r.repair_us_per_op;
        return d;
      });

  // addition: weighted personalized PageRank kernels on device pointers
  m.def(
      "ppr_forward",
      [](const DynamicGraph& g, std::uintptr_t w, std::uintptr_t s, std::uintptr_t x, std::uint32_t K, double damping,
         std::uint32_t iterations, std::uintptr_t stream) {
        py::gil_scoped_release nogil;
        velographx::vgx_check(vgx_ppr_forward(g.handle(), reinterpret_cast<const float*>(w),
                                              reinterpret_cast<const float*>(s), reinterpret_cast<float*>(x), K,
                                              damping, iterations, reinterpret_cast<void*>(stream)));
      },
      py::arg("graph"), py::arg("w_ptr"), py::arg("s_ptr"), py::arg("x_ptr"), py::arg("K"), py::arg("damping"),
      py::arg("iterations"), py::arg("stream") = 0);
  m.def(
      "ppr_backward",
      [](const DynamicGraph& g, std::uintptr_t w, std::uintptr_t x, std::uintptr_t c, std::uintptr_t lam,
         std::uintptr_t gs, std::uintptr_t gw, std::uint32_t K, double damping, std::uint32_t iterations,
         std::uintptr_t stream) {
        py::gil_scoped_release nogil;
        velographx::vgx_check(vgx_ppr_backward(g.handle(), reinterpret_cast<const float*>(w),
                                               reinterpret_cast<const float*>(x), reinterpret_cast<const float*>(c),
                                               reinterpret_cast<float*>(lam), reinterpret_cast<float*>(gs),
                                               reinterpret_cast<float*>(gw), K, damping, iterations,
                                               reinterpret_cast<void*>(stream)));
      },
      py::arg("graph"), py::arg("w_ptr"), py::arg("x_ptr"), py::arg("c_ptr"), py::arg("lambda_ptr"),
      py::arg("grad_s_ptr"), py::arg("grad_w_ptr"), py::arg("K"), py::arg("damping"), py::arg("iterations"),
      py::arg("stream") = 0);

  m.def("from_numpy_edges", &graph_from_numpy_edges, py::arg("edges"), py::arg("directed") = false);
  m.def("from_scipy_csr", &graph_from_scipy_csr, py::arg("csr"), py::arg("directed") = false);
  m.def("from_arrow_columns", &graph_from_arrow_columns, py::arg("src"), py::arg("dst"), py::arg("directed") = false);
  m.def("from_arrow_table", &graph_from_arrow_table, py::arg("table"), py::arg("src") = "src", py::arg("dst") = "dst",
        py::arg("directed") = false);

  // additions: workload generator and device info
  m.def(
      "rmat_edges",
      [](std::uint32_t scale, std::uint32_t edgefactor, std::uint64_t seed) {
        const std::size_t mm = static_cast<std::size_t>(edgefactor) << scale;
        py::array_t<std::uint32_t> s(static_cast<py::ssize_t>(mm)), d(static_cast<py::ssize_t>(mm));
        auto* sp = s.mutable_data();
        auto* dp = d.mutable_data();
        {
          py::gil_scoped_release nogil;
          velographx::vgx_check(vgx_rmat_generate(scale, edgefactor, seed, sp, dp));
        }
        return py::make_tuple(s, d);
      },
      py::arg("scale"), py::arg("edgefactor") = 16, py::arg("seed") = 1);
  m.def("device_count", &vgx_device_count);
  m.def("version", &vgx_version);
}
