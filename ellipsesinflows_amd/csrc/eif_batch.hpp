// eif_batch.hpp -- kernels that serve one case per launch or many.
//
// Every kernel of a time step is written as a body, B::run(vb, args...),
// that takes its workgroup's place in the grid (vb) from its caller:
//  - the named kernel (k_predictor, k_exchange, ...) passes the hardware's
//    ids and runs one case, as launched by the single-case path;
//  - k_batch<B, K...> runs the same body for many independent cases in one
//    launch: the workgroups of case 0 come first, then those of case 1, and
//    so on; each workgroup finds its case in the launch's BatchMap and reads
//    that case's arguments (a Pack of the kernel's parameter types) from a
//    per-case descriptor array in device memory.
//
// The host launches through go<B>(kernel, grid, block, lds, stream, args...).
// Normally that is the named kernel's launch.  While a Recorder is set on
// the calling thread, go<B> records the launch instead (the batched kernel,
// the case's grid, block, LDS bytes and its argument Pack), so a step of
// several cases can be recorded case by case and replayed as batched
// launches (eif_gpu_ensemble_run_steps).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstring>
#include <utility>
#include <vector>

namespace eif {

// a workgroup's ids and the grid's dimensions, as its case sees them
struct VB {
  unsigned bx, by, bz;
  unsigned gx, gy, gz;
  __device__ static VB here() {
    return VB{blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, gridDim.y, gridDim.z};
  }
};

// a kernel's arguments as one trivially copyable record, laid out alike on
// the host and the device
template <class... K>
struct Pack;
template <>
struct Pack<> {};
template <class H, class... T>
struct Pack<H, T...> {
  H h;
  Pack<T...> t;
};

template <std::size_t I, class H, class... T>
__host__ __device__ inline const auto &pack_get(const Pack<H, T...> &p) {
  if constexpr (I == 0)
    return p.h;
  else
    return pack_get<I - 1>(p.t);
}

template <class H, class... T, class A, class... As>
inline void pack_set(Pack<H, T...> &p, A &&a, As &&...as) {
  p.h = static_cast<H>(a);
  if constexpr (sizeof...(T) > 0) pack_set(p.t, std::forward<As>(as)...);
}

// cases per batched launch
constexpr int BATCH_MAX = 64;

// the cases of one batched launch: case c owns workgroups
// [start[c], start[c+1]) of a one-dimensional grid, which it sees as a grid
// of gx x gy x gz; its argument Pack is at packs + c * stride
struct BatchMap {
  int n;
  unsigned stride;
  unsigned start[BATCH_MAX + 1];
  unsigned gx[BATCH_MAX], gy[BATCH_MAX], gz[BATCH_MAX];
};

__device__ __forceinline__ int batch_case(const BatchMap &m, VB *vb) {
  const unsigned b = blockIdx.x;
  int c = 0;
  while (c + 1 < m.n && b >= m.start[c + 1]) c++;
  const unsigned l = b - m.start[c];
  vb->gx = m.gx[c];
  vb->gy = m.gy[c];
  vb->gz = m.gz[c];
  vb->bx = l % vb->gx;
  const unsigned r = l / vb->gx;
  vb->by = r % vb->gy;
  vb->bz = r / vb->gy;
  return c;
}

template <class B, class... K, std::size_t... I>
__device__ __forceinline__ void batch_apply(const VB &vb, const Pack<K...> &p,
                                            std::index_sequence<I...>) {
  B::run(vb, pack_get<I>(p)...);
}

// B::T is the body's workgroup size bound; B::M, when nonzero, the second
// launch-bounds operand of its named kernel
template <class B, class... K>
__global__ __launch_bounds__(B::T) void k_batch(const char *__restrict__ packs, BatchMap m) {
  VB vb;
  const int c = batch_case(m, &vb);
  batch_apply<B>(vb, *reinterpret_cast<const Pack<K...> *>(packs + (std::size_t)c * m.stride),
                 std::index_sequence_for<K...>{});
}

template <class B, class... K>
__global__ __launch_bounds__(B::T, B::M) void k_batch_m(const char *__restrict__ packs, BatchMap m) {
  VB vb;
  const int c = batch_case(m, &vb);
  batch_apply<B>(vb, *reinterpret_cast<const Pack<K...> *>(packs + (std::size_t)c * m.stride),
                 std::index_sequence_for<K...>{});
}

// one recorded launch of one case
struct LaunchRec {
  const void *fn;           // the batched kernel of its body and argument types
  std::vector<char> pack;   // the argument Pack's bytes
  dim3 grid, block;
  std::size_t shm;
};

struct Recorder {
  std::vector<LaunchRec> recs;
};

// set while a case's step is recorded (eif_api.hip); null otherwise
extern thread_local Recorder *g_rec;

template <class B, class... K, class... A>
inline void go(void (*kern)(K...), dim3 grid, dim3 block, std::size_t shm, hipStream_t st,
               A &&...a) {
  static_assert(sizeof...(K) == sizeof...(A), "argument count");
  if (Recorder *r = g_rec) {
    Pack<K...> p;
    std::memset(&p, 0, sizeof p);
    pack_set(p, std::forward<A>(a)...);
    LaunchRec rec;
    if constexpr (B::M > 0)
      rec.fn = reinterpret_cast<const void *>(&k_batch_m<B, K...>);
    else
      rec.fn = reinterpret_cast<const void *>(&k_batch<B, K...>);
    rec.pack.assign(reinterpret_cast<const char *>(&p), reinterpret_cast<const char *>(&p) + sizeof p);
    rec.grid = grid;
    rec.block = block;
    rec.shm = shm;
    r->recs.push_back(std::move(rec));
    return;
  }
  hipLaunchKernelGGL(kern, grid, block, shm, st, static_cast<K>(a)...);
}

}  // namespace eif
