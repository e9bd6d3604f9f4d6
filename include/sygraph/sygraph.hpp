/*
 * sygraph.hpp -- umbrella include (reference include/sygraph/sygraph.hpp:7-47).
 * Translation units that include it must be compiled with hipcc for gfx950
 * and linked with libsygraph_hip.
 */
#pragma once

#include <sygraph/formats/coo.hpp>
#include <sygraph/formats/csr.hpp>

#include <sygraph/frontier/frontier.hpp>
#include <sygraph/frontier/frontier_settings.hpp>

#include <sygraph/graph/build.hpp>
#include <sygraph/graph/graph.hpp>
#include <sygraph/graph/properties.hpp>

#include <sygraph/io/read_coo.hpp>
#include <sygraph/io/read_csr.hpp>

#include <sygraph/sycl/event.hpp>
#include <sygraph/utils/memory.hpp>

#include <sygraph/operators/advance/advance.hpp>
#include <sygraph/operators/config.hpp>
#include <sygraph/operators/filter/filter.hpp>
#include <sygraph/operators/for/for.hpp>
#include <sygraph/operators/intersection/intersection.hpp>

#include <sygraph/algorithms/bc.hpp>
#include <sygraph/algorithms/bfs.hpp>
#include <sygraph/algorithms/cc.hpp>
#include <sygraph/algorithms/sssp.hpp>
#include <sygraph/algorithms/tc.hpp>

#include <sygraph/sync/atomics.hpp>

#ifdef ENABLE_PROFILING
#include <sygraph/utils/profiler.hpp>
#endif
