// Storage-independent graph access -- the interface of
// include/velographx/graph_access.hpp:61-184: the ReadableGraph and
// MutableGraph concepts and the free functions the incremental algorithms
// reach a graph through.  Each function prefers a member of the graph, then
// a `vx_*` function found by argument-dependent lookup, then a fallback
// built from what the graph does offer (the reference's order, so a graph
// type adapted for the reference is adapted here too).
//
// On this engine the algorithms themselves run on a device copy of a
// foreign graph (graph_binding.hpp); these functions are how that copy is
// read from, and kept in step with, the caller's graph.
#pragma once

#include <concepts>
#include <cstddef>
#include <cstdint>
#include <utility>

#include "velographx_amd/types.hpp"

namespace velographx {

namespace access_detail {

// A neighbour entry is an id or an (id, weight) pair.
template <class T>
constexpr VertexId target_of(const T& entry) noexcept {
  if constexpr (requires { entry.first; }) return static_cast<VertexId>(entry.first);
  else return static_cast<VertexId>(entry);
}

template <class G>
concept HasCountMember = requires(const G& g) {
  { g.vertex_count() } -> std::convertible_to<std::size_t>;
};
template <class G>
concept HasCountAdl = requires(const G& g) {
  { vx_vertex_count(g) } -> std::convertible_to<std::size_t>;
};
template <class G>
concept HasDirectedMember = requires(const G& g) {
  { g.directed() } -> std::convertible_to<bool>;
};
template <class G>
concept HasDirectedAdl = requires(const G& g) {
  { vx_is_directed(g) } -> std::convertible_to<bool>;
};
template <class G>
concept HasVersionMember = requires(const G& g) {
  { g.version() } -> std::convertible_to<std::uint64_t>;
};
template <class G>
concept HasVersionAdl = requires(const G& g) {
  { vx_version(g) } -> std::convertible_to<std::uint64_t>;
};

}  // namespace access_detail

// graph_access.hpp:61-95
template <class G>
concept ReadableGraph = (access_detail::HasCountMember<G> || access_detail::HasCountAdl<G>) &&
                        (access_detail::HasDirectedMember<G> || access_detail::HasDirectedAdl<G>);

template <class G>
concept MutableGraph =
    ReadableGraph<G> && (access_detail::HasVersionMember<G> || access_detail::HasVersionAdl<G>);

template <ReadableGraph G>
[[nodiscard]] constexpr std::size_t vertex_count(const G& g) {
  if constexpr (access_detail::HasCountMember<G>) return static_cast<std::size_t>(g.vertex_count());
  else return static_cast<std::size_t>(vx_vertex_count(g));
}

template <ReadableGraph G>
[[nodiscard]] constexpr bool is_directed(const G& g) {
  if constexpr (access_detail::HasDirectedMember<G>) return static_cast<bool>(g.directed());
  else return static_cast<bool>(vx_is_directed(g));
}

template <MutableGraph G>
[[nodiscard]] constexpr std::uint64_t graph_version(const G& g) {
  if constexpr (access_detail::HasVersionMember<G>) return static_cast<std::uint64_t>(g.version());
  else return static_cast<std::uint64_t>(vx_version(g));
}

// graph_access.hpp:97-108: fn(v) for every out-neighbour v of u.
template <ReadableGraph G, class Fn>
void for_each_neighbor(const G& g, VertexId u, Fn&& fn) {
  if constexpr (requires { g.for_each_neighbor(u, fn); }) {
    g.for_each_neighbor(u, std::forward<Fn>(fn));
  } else if constexpr (requires { vx_for_each_neighbor(g, u, fn); }) {
    vx_for_each_neighbor(g, u, std::forward<Fn>(fn));
  } else {
    for (const auto& entry : g.neighbors(u)) fn(access_detail::target_of(entry));
  }
}

// graph_access.hpp:123-140: fn(u) for every in-neighbour u of v; a graph
// with neither an in-row accessor nor an adaptation is scanned whole.
template <ReadableGraph G, class Fn>
void for_each_in_neighbor(const G& g, VertexId v, Fn&& fn) {
  if constexpr (requires { g.for_each_in_neighbor(v, fn); }) {
    g.for_each_in_neighbor(v, std::forward<Fn>(fn));
  } else if constexpr (requires { vx_for_each_in_neighbor(g, v, fn); }) {
    vx_for_each_in_neighbor(g, v, std::forward<Fn>(fn));
  } else if constexpr (requires { g.in_neighbors(v); }) {
    for (const auto& entry : g.in_neighbors(v)) fn(access_detail::target_of(entry));
  } else {
    const std::size_t n = vertex_count(g);
    for (std::size_t u = 0; u < n; ++u) {
      const auto src = static_cast<VertexId>(u);
      for_each_neighbor(g, src, [&](VertexId dst) {
        if (dst == v) fn(src);
      });
    }
  }
}

// graph_access.hpp:142-153
template <ReadableGraph G>
[[nodiscard]] std::size_t neighbor_count(const G& g, VertexId u) {
  if constexpr (requires { g.degree(u); }) {
    return static_cast<std::size_t>(g.degree(u));
  } else if constexpr (requires { vx_neighbor_count(g, u); }) {
    return static_cast<std::size_t>(vx_neighbor_count(g, u));
  } else {
    std::size_t k = 0;
    for_each_neighbor(g, u, [&](VertexId) { ++k; });
    return k;
  }
}

// graph_access.hpp:155-166
template <ReadableGraph G>
[[nodiscard]] bool has_edge(const G& g, VertexId u, VertexId v) {
  if constexpr (requires { g.has_edge(u, v); }) {
    return static_cast<bool>(g.has_edge(u, v));
  } else if constexpr (requires { vx_has_edge(g, u, v); }) {
    return static_cast<bool>(vx_has_edge(g, u, v));
  } else {
    bool hit = false;
    for_each_neighbor(g, u, [&](VertexId w) { hit = hit || w == v; });
    return hit;
  }
}

// graph_access.hpp:177-184
template <class G, class Batch>
void apply_updates(G& g, const Batch& batch) {
  if constexpr (requires { g.apply(batch); }) g.apply(batch);
  else vx_apply_updates(g, batch);
}

}  // namespace velographx
