// Device helpers shared by the kernel files (internal).
#pragma once

#include <hip/hip_runtime.h>

#include "pmx_kernels.h"

namespace pmx {

__device__ __forceinline__ int nint_piny(double x) {
  // NINT with SIMP_NINT, defines.h:96-97
  return (int)((x) >= 0.0 ? ((x) + 0.5) : ((x)-0.5));
}

// The same rounding (half away from zero) as a double, without the round
// trip through an integer register: trunc(x + copysign(0.5, x)) equals
// (int)(x >= 0 ? x + 0.5 : x - 0.5) for every |x| < 2^31, up to the sign of
// a zero result.
__device__ __forceinline__ double nint_d(double x) { return trunc(x + copysign(0.5, x)); }

__device__ __forceinline__ void atomic_add(double *p, double v) { unsafeAtomicAdd(p, v); }

// Per-bead sums over the threads of one workgroup, written as
// part[(v*P + bead0 + k)*G + blk] for k < nb (the G blocks of one (value,
// bead) slot are contiguous, so the reduction reads them coalesced).  Thread t holds
// bead slot t % nb, so a workgroup covers nb beads.  The order of the sums
// is fixed, so the result does not depend on scheduling.  Every thread of
// the block must call it.
template <int NV>
__device__ __forceinline__ void block_partials(const double (&v)[NV], int nb, int bead0, int P, double *part,
                                               size_t blk, size_t G) {
  __shared__ double red[256];
  const int tid = threadIdx.x, lane = tid & 63, bs = blockDim.x;
  const bool shfl = (64 % nb) == 0 && (bs & 63) == 0;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = v[k];
    if (shfl) {
      for (int off = 32; off >= nb; off >>= 1) s += __shfl_xor(s, off, 64);
      if (lane < nb) red[(tid >> 6) * nb + lane] = s;
    } else {
      red[tid] = s;
    }
    __syncthreads();
    if (tid < nb && bead0 + tid < P) {
      double t = 0.0;
      if (shfl) {
        for (int w = 0; w < (bs >> 6); w++) t += red[w * nb + tid];
      } else {
        for (int u = tid; u < bs; u += nb) t += red[u];
      }
      part[((size_t)k * P + bead0 + tid) * G + blk] = t;
    }
    __syncthreads();
  }
}

// Workgroups are dispatched to the 8 XCDs of an MI355X in turn (workgroup
// id modulo 8), each XCD with its own L2.  Kernels whose work is (unit,
// bead) with units sharing a bead's data (PME tiles, FFT planes) launch
// 8 * units * ceil(P/8) workgroups in one dimension and map them so that
// every unit of one bead runs on one XCD.  Returns false for the padding
// workgroups of the last bead group.
__device__ __forceinline__ bool xcd_unit_bead(int units, int P, int &unit, int &bead) {
  const int L = blockIdx.x, x = L & 7, r = L >> 3;
  bead = x + 8 * (r / units);
  unit = r % units;
  return bead < P;
}

inline unsigned xcd_grid(int units, int P) { return 8u * (unsigned)units * (unsigned)((P + 7) / 8); }

}  // namespace pmx
