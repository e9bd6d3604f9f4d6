/*
 * frontier/frontier_settings.hpp -- frontier views and types (reference
 * frontier/frontier_settings.hpp:13-27).
 */
#pragma once

namespace sygraph {
namespace frontier {

enum class frontier_view {
  graph,  /* every vertex of the graph */
  vertex, /* the vertices in the frontier */
  edge,   /* the edges in the frontier */
  none,   /* no frontier */
};

enum class frontier_type {
  bitmap, /* flat bitmap */
  mlb,    /* two-level bitmap (default) */
  none,   /* placeholder */
};

} // namespace frontier
} // namespace sygraph
