// Hand-written device templates shared by every generated OM machine (gfx950).
//
// The generator instantiates these for each kernel of an Orthotope Machine:
//   * om_wrap        cyclic index wrap for halo reads
//                    (the reference's `(x + c + n) % n`, Generator/PlanTrans.hs:459-461,481-483),
//   * om_clamp       keeps halo reads of Open (non-periodic) axes inside the grid,
//   * om_max/om_min  exactly std::max/std::min (the reference's CPU Max/Min,
//                    Generator/PlanTrans.hs:690-691): ties and NaNs resolve as there,
//   * om_resident_workgroups  grid size of a persistent (tile-looping) launch,
//   * om_block_*     wave64 shuffle tree + LDS combine for Reduce: each
//                    workgroup writes one partial, and a one-workgroup finish
//                    kernel folds the partials in a fixed order, replacing the
//                    reference's serial om_reduce_* (Generator/PlanTrans.hs:719)
//                    with a deterministic tree.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Workgroups a persistent launch starts: as many as the device holds at once (per_cu per
// CU, from the launch's LDS and work-items), a multiple of 8 so that workgroup b, which the
// dispatcher places on XCD b % 8, keeps to that XCD's band of tiles; at most ntiles.
static inline unsigned om_resident_workgroups(unsigned ntiles, int per_cu) {
  static int cus = 0;
  if (cus <= 0) {
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d) != hipSuccess ||
        cus <= 0)
      cus = 256;
  }
  unsigned n = ((unsigned)cus * (unsigned)(per_cu > 0 ? per_cu : 1)) & ~7u;
  if (n == 0) n = 8;
  return n < ntiles ? n : ntiles;
}


// Division of a tile index by a tile count d without the division expansion: m is
// floor((2^32 - 1) / d) from the host, so umulhi(x, m) is floor(x / d) or one less for any
// x < 2^31, and one compare corrects it (the operands are uniform: scalar instructions).
static inline unsigned om_udiv_magic(unsigned d) { return d ? 0xFFFFFFFFu / d : 0u; }

__device__ __forceinline__ unsigned om_udiv(unsigned x, unsigned d, unsigned m) {
  const unsigned q = __umulhi(x, m);
  return q + ((q + 1u) * d <= x ? 1u : 0u);
}

// One conditional add and subtract cover -n <= x < 2n.  `wide` (uniform: the host sets it
// when some axis of the grid is shorter than the launch's tile plus its reach) enables the
// remainder for coordinates further out; otherwise the branch around it is scalar.
__device__ __forceinline__ int om_wrap(int x, int n, int wide) {
  x = (x < 0) ? x + n : x;
  x = (x >= n) ? x - n : x;
  if (wide && (unsigned)x >= (unsigned)n) x = ((x % n) + n) % n;
  return x;
}

__device__ __forceinline__ int om_clamp(int x, int n) { return x < 0 ? 0 : (x < n ? x : n - 1); }

template <class T> __device__ __forceinline__ T om_max(T a, T b) { return (a < b) ? b : a; }
template <class T> __device__ __forceinline__ T om_min(T a, T b) { return (b < a) ? b : a; }

struct OmSum {
  template <class T> __device__ __forceinline__ T operator()(T a, T b) const { return a + b; }
};
struct OmSumWrap {   // integer sum with two's-complement wrap (order independent)
  __device__ __forceinline__ int operator()(int a, int b) const {
    return (int)((unsigned)a + (unsigned)b);
  }
};
struct OmMin {
  template <class T> __device__ __forceinline__ T operator()(T a, T b) const { return om_min(a, b); }
};
struct OmMax {
  template <class T> __device__ __forceinline__ T operator()(T a, T b) const { return om_max(a, b); }
};

// Fixed-shape reduction of one value per thread of an NT-thread workgroup; every thread
// gets the result.  The xor-shuffle tree and the wave order are fixed, so results are
// bit-reproducible run to run.
template <int NT, class T, class Op>
__device__ __forceinline__ T om_block_reduce(T v, Op op, T* scratch /* >= NT / 64 */) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  T r = scratch[0];
#pragma unroll
  for (int w = 1; w < NT / 64; ++w) r = op(r, scratch[w]);
  return r;   // every thread gets the block result
}

// Reduce `count` partials in a fixed order.  Thread t keeps eight accumulators: lane
// k folds partials t + (8j + k) NT, so eight independent loads are in flight per thread
// (a one-workgroup reduction of 10^4-10^5 partials is bound by load latency); then the
// eight are folded in order and the workgroup tree finishes.
template <int NT, class T, class Op>
__device__ __forceinline__ T om_partials_reduce(const T* __restrict__ part, int count, T init,
                                                 Op op, T* scratch) {
  T acc[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k] = init;
  int i = threadIdx.x;
  for (; i + 7 * NT < count; i += 8 * NT) {
    T v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = part[i + k * NT];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = op(acc[k], v[k]);
  }
  for (int k = 0; i < count; i += NT, ++k) acc[k] = op(acc[k], part[i]);
  T r = acc[0];
#pragma unroll
  for (int k = 1; k < 8; ++k) r = op(r, acc[k]);
  return om_block_reduce<NT>(r, op, scratch);
}
