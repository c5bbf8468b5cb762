// Kernel launches that go either straight to a stream or into a graph under
// construction, and the MINRES scalar block the solver kernels share; used by
// the Krylov solvers and by the preconditioners they enqueue.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <tuple>
#include <type_traits>
#include <vector>

#include "feec_internal.hpp"

namespace feec {

// Enqueues kernels on a stream, or appends them as a linear chain of kernel
// nodes to a graph under construction. Node arguments are kept alive in
// `keep` for as long as the graph exists, so nothing that inspects a node's
// parameters after its creation can see freed memory.
template <class... P>
struct KernelArgs {
  std::tuple<P...> values;
  void* ptrs[sizeof...(P) > 0 ? sizeof...(P) : 1];
  template <class... A>
  explicit KernelArgs(A... a) : values(static_cast<P>(a)...) {
    std::apply(
        [&](auto&... x) {
          int i = 0;
          ((ptrs[i++] = static_cast<void*>(&x)), ...);
        },
        values);
  }
};

struct Enq {
  hipStream_t st = nullptr;
  hipGraph_t graph = nullptr;
  std::vector<std::shared_ptr<void>>* keep = nullptr;
  hipGraphNode_t last = nullptr;

  template <class... P, class... A>
  void operator()(void (*kernel)(P...), dim3 grid, dim3 block, A... a) {
    static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
    if (!graph) {
      KernelArgs<std::decay_t<P>...> args(a...);
      FEEC_HIP(hipLaunchKernel(reinterpret_cast<const void*>(kernel), grid, block, args.ptrs, 0, st));
      return;
    }
    auto args = std::make_shared<KernelArgs<std::decay_t<P>...>>(a...);
    keep->push_back(args);
    hipKernelNodeParams p{};
    p.func = reinterpret_cast<void*>(kernel);
    p.gridDim = grid;
    p.blockDim = block;
    p.sharedMemBytes = 0;
    p.kernelParams = args->ptrs;
    p.extra = nullptr;
    hipGraphNode_t node = nullptr;
    FEEC_HIP(hipGraphAddKernelNode(&node, graph, last ? &last : nullptr, last ? 1 : 0, &p));
    last = node;
  }
};

struct MinresS {
  double beta, oldb, beta1, dbar, epsln, phibar, cs, sn;
  double alfa, residual;
  double invb_prev, oldeps, delta, invg, phi;
  double rtol;
  double extra_alpha;
  long long iter, max_iters;
  int done;
  int pad;
};

}  // namespace feec
