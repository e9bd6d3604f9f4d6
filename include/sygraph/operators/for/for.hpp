/*
 * operators/for/for.hpp -- per-vertex compute and reduction over a frontier
 * (reference operators/for/for.hpp:25-92, impl_for.hpp:25-111):
 *   compute::execute<vertex>(G, F, f):          f(v) for every v in F
 *   compute::reduce<vertex, Op>(G, F, acc, f):  acc = acc (+) f-contributions
 * In reduce the functor receives a reducer `r` and adds its contribution with
 * `r += x` or `r.combine(x)`; both apply Op (plus, multiplies, minimum or
 * maximum).  The reference always sums whatever Op says (impl_for.hpp:99).
 * Each workgroup reduces its lanes; the per-workgroup partials are combined
 * on the host in workgroup order, so the result is deterministic.
 */
#pragma once

#include <concepts>
#include <vector>

#include <sygraph/operators/advance/common.hpp>

namespace sygraph {
namespace operators {
namespace compute {

template<typename T, typename R>
concept ReducerT = std::same_as<T, sycl::plus<R>> || std::same_as<T, sycl::multiplies<R>> || std::same_as<T, sycl::minimum<R>>
                   || std::same_as<T, sycl::maximum<R>>;

namespace detail {

template<typename R, typename Op>
struct Reducer {
  R value;
  __host__ __device__ Reducer& combine(const R& x) {
    value = Op{}(value, x);
    return *this;
  }
  __host__ __device__ Reducer& operator+=(const R& x) { return combine(x); }
};

template<typename InDev, typename F>
__global__ void __launch_bounds__(256) execute_kernel(InDev in, F f) {
  const uint32_t count = *in.getOffsetsSize();
  const int* list = in.getOffsets();
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < count; i += gridDim.x * 256) f((uint32_t)list[i]);
}

template<typename R, typename Op, typename InDev, typename F>
__global__ void __launch_bounds__(256) reduce_kernel(InDev in, F f, R* partial) {
  __shared__ R red[256];
  const uint32_t count = *in.getOffsetsSize();
  const int* list = in.getOffsets();
  Reducer<R, Op> r{Op::identity()};
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < count; i += gridDim.x * 256) f((uint32_t)list[i], r);
  red[threadIdx.x] = r.value;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = Op{}(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

} // namespace detail

template<frontier::frontier_view FW, typename GraphT, typename T, typename LambdaT, frontier::frontier_type FT>
sygraph::Event execute(GraphT& graph,
                       const sygraph::frontier::Frontier<T, FT>& frontier,
                       LambdaT&& functor,
                       frontier::size::frontier_size_t expected_size = frontier::size::fetch_from_memory) {
  static_assert(FW == frontier::frontier_view::vertex, "Invalid frontier view for compute operation.");
  sycl::queue& q = graph.getQueue();
  auto ev = sycl::event::begin(q.handle(), q.stream(), q.origin());
  const uint32_t* dcount = operators::advance::detail::compact_sources(frontier);
  const unsigned grid = operators::advance::detail::push_grid(q, expected_size, dcount);
  std::decay_t<LambdaT> f = functor;
  detail::execute_kernel<<<grid, 256, 0, q.stream()>>>(frontier.getDeviceFrontier(), f);
  operators::advance::detail::check_launch();
  ev.finish();
  return ev;
}

template<frontier::frontier_view FW, typename ReductionOperator, typename GraphT, typename T, typename R,
         frontier::frontier_type FT, typename LambdaT>
  requires ReducerT<ReductionOperator, R>
sygraph::Event reduce(GraphT& graph,
                      const sygraph::frontier::Frontier<T, FT>& frontier,
                      R& accumulator,
                      LambdaT&& function,
                      frontier::size::frontier_size_t expected_size = frontier::size::fetch_from_memory) {
  static_assert(FW == frontier::frontier_view::vertex, "Invalid frontier view for compute operation.");
  sycl::queue& q = graph.getQueue();
  auto ev = sycl::event::begin(q.handle(), q.stream(), q.origin());
  const uint32_t* dcount = operators::advance::detail::compact_sources(frontier);
  const unsigned grid = operators::advance::detail::push_grid(q, expected_size, dcount);
  R* partial = sycl::malloc_device<R>(grid, q);
  std::decay_t<LambdaT> f = function;
  detail::reduce_kernel<R, ReductionOperator><<<grid, 256, 0, q.stream()>>>(frontier.getDeviceFrontier(), f, partial);
  operators::advance::detail::check_launch();
  std::vector<R> host(grid);
  q.copy(partial, host.data(), grid);
  sycl::free(partial, q);
  for (const R& x : host) accumulator = ReductionOperator{}(accumulator, x);
  ev.finish();
  return ev;
}

} // namespace compute
} // namespace operators
} // namespace sygraph
