// Grouped launches: one kernel launch runs the same device body for several
// problems of one shape.  blockIdx.z selects the problem; its argument pack
// sits in a device table built on the host from the arguments the
// single-problem launch would have taken.  A problem whose stop flag is set
// (a converged solve) exits at once: every kernel of a converged solve's
// refinement step leaves its iterate and residual unchanged, so skipping
// them is the same as running them.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <utility>

namespace plesmg {

// Plain aggregate holding a kernel's arguments in order (trivially copyable:
// the host writes it byte for byte into the device table).
template <class... A>
struct Pack {};
template <class T, class... R>
struct Pack<T, R...> {
    T v;
    Pack<R...> r;
};

template <std::size_t I>
struct PGet {
    template <class T, class... R>
    __host__ __device__ static const auto& get(const Pack<T, R...>& p) {
        return PGet<I - 1>::get(p.r);
    }
};
template <>
struct PGet<0> {
    template <class T, class... R>
    __host__ __device__ static const T& get(const Pack<T, R...>& p) {
        return p.v;
    }
};

inline Pack<> make_pack() { return {}; }
template <class T, class... R>
Pack<T, R...> make_pack(T v, R... r) {
    return Pack<T, R...>{v, make_pack(r...)};
}

// one problem's entry of a grouped launch: its stop flag (nullptr: always
// run) and its arguments
template <class... A>
struct GEntry {
    const double* done;
    Pack<A...> a;
};

template <auto F, class... A, std::size_t... I>
__device__ __forceinline__ void pcall(const Pack<A...>& p, std::index_sequence<I...>) {
    F(PGet<I>::get(p)...);
}

template <auto F, int LB, int WPE, class... A>
__global__ void __launch_bounds__(LB) __attribute__((amdgpu_waves_per_eu(WPE)))
k_group(const GEntry<A...>* __restrict__ tab) {
    const GEntry<A...>& e = tab[blockIdx.z];
    if (e.done != nullptr && *e.done != 0.0) return;  // uniform over the block
    pcall<F>(e.a, std::index_sequence_for<A...>{});
}

// host side: launch k_group over n entries of the table tab (device memory)
template <auto F, int LB, int WPE, class... A>
hipError_t launch_group(const void* tab, int n, dim3 g, dim3 b, size_t shm, hipStream_t s) {
    hipLaunchKernelGGL((k_group<F, LB, WPE, A...>), dim3(g.x, g.y, (unsigned)n), b, shm, s,
                       (const GEntry<A...>*)tab);
    return hipGetLastError();
}

}  // namespace plesmg
