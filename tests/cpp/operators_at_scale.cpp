/*
 * operators_at_scale.cpp -- the drop-in operator API driven the way a user of
 * the reference drives it, on graphs large enough to reach every work
 * mapping: BFS written with advance::frontier<push | pull | pull_all, each
 * load balancer> and a user lambda (the functors of algorithms/bfs.hpp:180-209),
 * SSSP with the push advance and filter::external (algorithms/sssp.hpp:152-210),
 * filter::inplace, and compute::reduce with each of plus, multiplies, minimum
 * and maximum (operators/for/for.hpp:25-92).
 *
 *   operators_at_scale GRAPH SOURCE OUT
 *
 * GRAPH: little-endian int64 n, int64 m, int64 offsets[n + 1], int32 cols[m],
 * uint32 weights[m] (symmetric adjacency).  OUT receives, in order: the BFS
 * depths of the 12 (direction, load balancer) runs (uint32[n] each, unreached
 * n + 1, in the order of kRuns below), the SSSP distances of the 4 push runs
 * (uint32[n] each), the frontier left by filter::inplace (uint8[n] membership)
 * and the five reductions (uint64 each): sum of depths, largest depth,
 * smallest (depth + id % 7), product of (1 + id % 2) over the vertices of
 * depth 1 with id below 64, and the plus-reduction of the vertices of depth
 * 1, whose initial accumulator is 5.
 */
#include <sycl/sycl.hpp>
#include <sygraph/sygraph.hpp>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

using sygraph::frontier::frontier_type;
using sygraph::frontier::frontier_view;
using sygraph::operators::direction;
using sygraph::operators::load_balancer;
using space = sygraph::memory::space;

template<typename T>
std::vector<T> read_array(std::FILE* f, size_t count) {
  std::vector<T> v(count);
  if (count && std::fread(v.data(), sizeof(T), count, f) != count) throw std::runtime_error("short graph file");
  return v;
}

template<typename GraphT>
std::vector<uint32_t> copy_out(sycl::queue& q, const uint32_t* dev, size_t n) {
  std::vector<uint32_t> h(n);
  q.copy(dev, h.data(), n).wait();
  return h;
}

/* BFS of algorithms/bfs.hpp:153-238 with one direction and load balancer for
 * every level */
template<direction Dir, load_balancer Lb, typename GraphT>
std::vector<uint32_t> bfs(sycl::queue& q, GraphT& G, uint32_t source) {
  const size_t n = G.getVertexCount();
  uint32_t* depth = sygraph::memory::detail::memoryAlloc<uint32_t, space::device>(n, q);
  q.fill(depth, static_cast<uint32_t>(n + 1), n).wait();
  q.fill(depth + source, 0u, 1).wait();
  auto in = sygraph::frontier::makeFrontier<frontier_view::vertex, frontier_type::mlb>(q, G);
  auto out = sygraph::frontier::makeFrontier<frontier_view::vertex, frontier_type::mlb>(q, G);
  in.insert(source);
  const uint32_t unreached = static_cast<uint32_t>(n + 1);
  uint32_t iter = 0;
  while (!in.empty()) {
    if constexpr (Dir == direction::push) {
      sygraph::operators::advance::frontier<direction::push, Lb, frontier_view::vertex, frontier_view::vertex>(
          G, in, out,
          [=](auto src, auto dst, auto edge, auto weight) -> bool {
            if (depth[dst] == unreached) {
              depth[dst] = iter + 1;
              return true;
            }
            return false;
          })
          .waitAndThrow();
    } else {
      sygraph::operators::advance::frontier<Dir, Lb, frontier_view::vertex, frontier_view::vertex>(
          G, in, out,
          [=](auto src, auto dst, auto edge, auto weight) -> bool {
            if (depth[src] == unreached && depth[dst] == iter) {
              depth[src] = iter + 1;
              return true;
            }
            return false;
          })
          .waitAndThrow();
    }
    sygraph::frontier::swap(in, out);
    out.clear();
    ++iter;
  }
  auto h = copy_out<GraphT>(q, depth, n);
  sygraph::memory::detail::releaseUSM(depth, q);
  return h;
}

/* SSSP of algorithms/sssp.hpp:152-210: push relaxation with an atomic
 * minimum, then filter::external of the improved vertices not yet taken in
 * this round */
template<load_balancer Lb, typename GraphT>
std::vector<uint32_t> sssp(sycl::queue& q, GraphT& G, uint32_t source) {
  const size_t n = G.getVertexCount();
  uint32_t* dist = sygraph::memory::detail::memoryAlloc<uint32_t, space::device>(n, q);
  int32_t* visited = sygraph::memory::detail::memoryAlloc<int32_t, space::device>(n, q);
  q.fill(dist, static_cast<uint32_t>(n + 1), n).wait();
  q.fill(dist + source, 0u, 1).wait();
  q.fill(visited, -1, n).wait();
  auto in = sygraph::frontier::makeFrontier<frontier_view::vertex, frontier_type::mlb>(q, G);
  auto out = sygraph::frontier::makeFrontier<frontier_view::vertex, frontier_type::mlb>(q, G);
  in.insert(source);
  int32_t iter = 0;
  while (!in.empty()) {
    sygraph::operators::advance::frontier<Lb, frontier_view::vertex, frontier_view::vertex>(
        G, in, out,
        [=](auto src, auto dst, auto edge, auto weight) -> bool {
          const uint32_t ds = sygraph::sync::load(&dist[src]);
          uint32_t cand = ds + weight;
          const uint32_t before = sygraph::sync::min(&dist[dst], &cand);
          return cand < before;
        })
        .waitAndThrow();
    sygraph::operators::filter::external(G, out, in, [=](auto v) -> bool {
      if (visited[v] == iter) return false;
      visited[v] = iter;
      return true;
    }).waitAndThrow();
    out.clear();
    ++iter;
  }
  auto h = copy_out<GraphT>(q, dist, n);
  sygraph::memory::detail::releaseUSM(dist, q);
  sygraph::memory::detail::releaseUSM(visited, q);
  return h;
}

void put(std::FILE* f, const void* p, size_t bytes) {
  if (bytes && std::fwrite(p, 1, bytes, f) != bytes) throw std::runtime_error("short write");
}

} // namespace

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s GRAPH SOURCE OUT\n", argv[0]);
    return 2;
  }
  std::FILE* f = std::fopen(argv[1], "rb");
  if (!f) throw std::runtime_error("cannot open the graph file");
  const auto head = read_array<int64_t>(f, 2);
  const int64_t n = head[0], m = head[1];
  const auto off64 = read_array<int64_t>(f, (size_t)n + 1);
  const auto cols32 = read_array<int32_t>(f, (size_t)m);
  const auto w = read_array<uint32_t>(f, (size_t)m);
  std::fclose(f);
  const uint32_t source = (uint32_t)std::strtoul(argv[2], nullptr, 10);

  std::vector<size_t> off(off64.begin(), off64.end());
  std::vector<uint32_t> cols(cols32.begin(), cols32.end());
  sygraph::formats::CSR<uint32_t, uint32_t, size_t> csr(off, cols, w);
  sycl::queue q{sycl::gpu_selector_v};
  auto G = sygraph::graph::build::fromCSR<space::device>(q, csr);

  std::FILE* o = std::fopen(argv[3], "wb");
  if (!o) throw std::runtime_error("cannot open the output file");
  std::vector<std::vector<uint32_t>> runs;
  runs.push_back(bfs<direction::push, load_balancer::workitem_mapped>(q, G, source));
  runs.push_back(bfs<direction::push, load_balancer::subgroup_mapped>(q, G, source));
  runs.push_back(bfs<direction::push, load_balancer::workgroup_mapped>(q, G, source));
  runs.push_back(bfs<direction::push, load_balancer::bucketing>(q, G, source));
  runs.push_back(bfs<direction::pull, load_balancer::workitem_mapped>(q, G, source));
  runs.push_back(bfs<direction::pull, load_balancer::subgroup_mapped>(q, G, source));
  runs.push_back(bfs<direction::pull, load_balancer::workgroup_mapped>(q, G, source));
  runs.push_back(bfs<direction::pull, load_balancer::bucketing>(q, G, source));
  runs.push_back(bfs<direction::pull_all, load_balancer::workitem_mapped>(q, G, source));
  runs.push_back(bfs<direction::pull_all, load_balancer::subgroup_mapped>(q, G, source));
  runs.push_back(bfs<direction::pull_all, load_balancer::workgroup_mapped>(q, G, source));
  runs.push_back(bfs<direction::pull_all, load_balancer::bucketing>(q, G, source));
  runs.push_back(sssp<load_balancer::workitem_mapped>(q, G, source));
  runs.push_back(sssp<load_balancer::subgroup_mapped>(q, G, source));
  runs.push_back(sssp<load_balancer::workgroup_mapped>(q, G, source));
  runs.push_back(sssp<load_balancer::bucketing>(q, G, source));
  for (const auto& r : runs) put(o, r.data(), 4 * r.size());

  /* the reached vertices as a frontier; filter::inplace drops those of odd depth */
  const std::vector<uint32_t>& depth = runs[3];
  uint32_t* ddev = sygraph::memory::detail::memoryAlloc<uint32_t, space::device>((size_t)n, q);
  q.copy(depth.data(), ddev, (size_t)n).wait();
  auto reached = sygraph::frontier::makeFrontier<frontier_view::vertex, frontier_type::mlb>(q, G);
  auto level1 = sygraph::frontier::makeFrontier<frontier_view::vertex, frontier_type::mlb>(q, G);
  for (int64_t v = 0; v < n; ++v) {
    if (depth[v] != (uint32_t)(n + 1)) reached.insert((uint32_t)v);
    if (depth[v] == 1) level1.insert((uint32_t)v);
  }
  uint64_t sum = 0, mx = 0, mn = ~0ull, prod = 1, plus5 = 5;
  sygraph::operators::compute::reduce<frontier_view::vertex, sycl::plus<uint64_t>>(
      G, reached, sum, [=](auto v, auto& r) { r += (uint64_t)ddev[v]; })
      .waitAndThrow();
  sygraph::operators::compute::reduce<frontier_view::vertex, sycl::maximum<uint64_t>>(
      G, reached, mx, [=](auto v, auto& r) { r.combine((uint64_t)ddev[v]); })
      .waitAndThrow();
  sygraph::operators::compute::reduce<frontier_view::vertex, sycl::minimum<uint64_t>>(
      G, reached, mn, [=](auto v, auto& r) { r += (uint64_t)ddev[v] + (uint64_t)(v % 7); })
      .waitAndThrow();
  sygraph::operators::compute::reduce<frontier_view::vertex, sycl::multiplies<uint64_t>>(
      G, level1, prod, [=](auto v, auto& r) {
        if (v < 64) r.combine(1 + (uint64_t)(v % 2));
      })
      .waitAndThrow();
  sygraph::operators::compute::reduce<frontier_view::vertex, sycl::plus<uint64_t>>(
      G, level1, plus5, [=](auto v, auto& r) { r += 1; })
      .waitAndThrow();
  sygraph::operators::filter::inplace(G, reached, [=](auto v) -> bool { return ddev[v] % 2 == 1; }).waitAndThrow();
  std::vector<uint8_t> member((size_t)n, 0);
  for (int64_t v = 0; v < n; ++v) member[(size_t)v] = reached.check((uint32_t)v) ? 1 : 0;
  put(o, member.data(), member.size());
  const uint64_t red[5] = {sum, mx, mn, prod, plus5};
  put(o, red, sizeof(red));
  std::fclose(o);
  sygraph::memory::detail::releaseUSM(ddev, q);
  std::printf("operators_at_scale: n=%lld m=%lld source=%u done\n", (long long)n, (long long)m, source);
  return 0;
}
