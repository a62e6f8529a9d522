// Particle-mesh Ewald on gfx950: charge binning, tiled spread and gather,
// and the 3-D FFT with the influence-function convolution.
//
// Arithmetic restates ewald3d_pme.c (PINY): fractional coordinates and
// B-spline weights (:241-337), charge stuffing (:342-356), the k-space sum
// over the half-space sphere list (:367-430) and the force gather (:500-560).
// The MI355X layout differs from the reference's loop order:
//  - every (charge, bead) is binned by the grid tile holding the first point
//    of its support (k_pme_bin / k_pme_scan / k_pme_scatter, a counting sort);
//  - one workgroup per (tile, bead) spreads the charges of its bin, one lane
//    per charge, into an LDS halo box of (T + n - 1)^3 points with LDS
//    atomics and writes the box out once (no global atomics, no grid clear);
//  - the forward FFT sums, as it loads each grid plane, every point's own
//    box and the overhang of the boxes below it (the fold into the periodic
//    grid never goes through memory);
//  - the gather loads the tile's halo box of (T + n - 1)^3 potential values
//    into LDS and each lane reduces one charge's n^3 products in registers.
#include <hip/hip_runtime.h>

#include "pmx_device.h"

namespace pmx {

namespace {

constexpr int PME_BLOCK = 128;  // lanes per (tile, bead) workgroup: about one per charge
// lanes per charge in the spread (measured: with FP32 weights 90, 80 and 74
// us for 1, 2 and 4 lanes; with FP64 weights, whose registers allow fewer
// wavefronts, more lanes did not help)
template <typename T>
constexpr int spread_lanes() {
  return sizeof(T) == 4 ? 4 : 1;
}

// Complex type of a grid scalar T (double2 for FP64, float2 for the mixed
// precision's FP32 grids).
template <typename T>
struct Cx;
template <>
struct Cx<double> {
  using C = double2;
  __device__ static C make(double x, double y) { return make_double2(x, y); }
};
template <>
struct Cx<float> {
  using C = float2;
  __device__ static C make(float x, float y) { return make_float2(x, y); }
};

// B-spline weights M_n(frac + j - 1), j = 1..NI, by the recursion of
// ewald3d_pme.c:296-337; dmn (order NI-1 differences, :516-530) if DERIV.
template <int NI, bool DERIV, typename T>
__device__ __forceinline__ void bspline(T frac, T (&mn)[NI + 1], T (&dmn)[NI + 1]) {
  T u[NI + 1];
#pragma unroll
  for (int j = 1; j <= NI; j++) u[j] = frac + (T)(j - 1);
  mn[1] = (T)1 - fabs(u[1] - (T)1);
  mn[2] = (T)1 - fabs(u[2] - (T)1);
#pragma unroll
  for (int j = 3; j <= NI; j++) mn[j] = (T)0;
#pragma unroll
  for (int o = 3; o <= NI; o++) {
    const T rn = (T)o, rn1 = (T)1 / ((T)(o - 1));
#pragma unroll
    for (int j = o; j >= 2; j--) mn[j] = (u[j] * mn[j] + (rn - u[j]) * mn[j - 1]) * rn1;
    mn[1] = u[1] * mn[1] * rn1;
    if (DERIV && o == NI - 1) {
      dmn[1] = mn[1];
#pragma unroll
      for (int j = 2; j <= NI; j++) dmn[j] = mn[j] - mn[j - 1];
    }
  }
}

__device__ __forceinline__ void tile_origin(const PmeArgs &a, int tl, int &o0, int &o1, int &o2) {
  const int ta = tl % a.nt[0], tb = (tl / a.nt[0]) % a.nt[1], tc = tl / (a.nt[0] * a.nt[1]);
  o0 = ta * a.T[0];
  o1 = tb * a.T[1];
  o2 = tc * a.T[2];
}

}  // namespace

// --- binning -----------------------------------------------------------------

// one thread per item it = bead*nchrg + charge (grid: charge chunks x beads):
// fractional grid coordinates (ewald3d_pme.c:241-262, 270-276), first support
// point per axis, and the item's slot in its bin.  Slots are counted in an
// LDS histogram per workgroup first, so each (block, tile) pair makes one
// global atomic instead of one per item.
__global__ void __launch_bounds__(256) k_pme_bin(PmeArgs a) {
  extern __shared__ int hist[];  // [ntile]
  const int b = blockIdx.y, ic = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = ic < a.nchrg;
  for (int t = threadIdx.x; t < a.ntile; t += blockDim.x) hist[t] = 0;
  __syncthreads();
  const int it = b * a.nchrg + ic;
  int tile = 0, lpos = 0;
  if (act) {
    const size_t ib = (size_t)a.ichrg[ic] * a.P + b;
    const double x = a.x[ib], y = a.y[ib], z = a.z[ib];
    const double *hi = a.cell.hi;
    double s[3];
    s[0] = x * hi[0] + y * hi[3] + z * hi[6];
    s[1] = x * hi[1] + y * hi[4] + z * hi[7];
    s[2] = x * hi[2] + y * hi[5] + z * hi[8];
    int tcoord[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const int K = a.K[d];
      const double t = s[d] - nint_d((s[d] - 0.5));
      const double sc = t * (double)K;
      int ig = (int)(sc);
      const double frac = sc - (double)ig;
      // support points (0-based) ig - n + 1 .. ig, periodic; ig = K (a
      // coordinate rounding onto the upper face) is the same set as ig = 0
      ig = ((ig % K) + K) % K;
      int s0 = ig - a.n + 1;
      s0 = ((s0 % K) + K) % K;
      a.item_f[3 * (size_t)it + d] = frac;
      a.item_s[3 * (size_t)it + d] = s0;
      tcoord[d] = s0 / a.T[d];
    }
    tile = (tcoord[2] * a.nt[1] + tcoord[1]) * a.nt[0] + tcoord[0];
    lpos = atomicAdd(hist + tile, 1);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < a.ntile; t += blockDim.x) {
    const int n = hist[t];
    hist[t] = n ? atomicAdd(a.bin_cnt + (size_t)b * a.ntile + t, n) : 0;
  }
  __syncthreads();
  if (act) {
    a.item_tile[it] = tile;
    a.item_pos[it] = hist[tile] + lpos;
  }
}

// exclusive scan of the bin counts of one bead (one workgroup per bead)
__global__ void __launch_bounds__(256) k_pme_scan(PmeArgs a) {
  __shared__ int sums[256];
  const int b = blockIdx.x, tid = threadIdx.x, nt = a.ntile;
  const int per = (nt + 255) / 256;
  const int lo = min(nt, tid * per), hi = min(nt, lo + per);
  const int *cnt = a.bin_cnt + (size_t)b * nt;
  int s = 0;
  for (int t = lo; t < hi; t++) s += cnt[t];
  sums[tid] = s;
  __syncthreads();
  if (tid == 0) {
    int run = 0;
    for (int k = 0; k < 256; k++) {
      const int v = sums[k];
      sums[k] = run;
      run += v;
    }
  }
  __syncthreads();
  int run = b * a.nchrg + sums[tid];
  for (int t = lo; t < hi; t++) {
    a.bin_off[(size_t)b * nt + t] = run;
    run += cnt[t];
  }
}

__global__ void __launch_bounds__(256) k_pme_scatter(PmeArgs a) {
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= a.P * a.nchrg) return;
  const int b = it / a.nchrg;
  a.sorted[a.bin_off[(size_t)b * a.ntile + a.item_tile[it]] + a.item_pos[it]] = it;
}

// --- spread --------------------------------------------------------------------

// Weights of one item along the three axes: w[d][m] is the weight of support
// point s_d + m (M_n index j = n - m); l[d] is s_d relative to the tile origin.
template <int NI, bool DERIV, typename T>
__device__ __forceinline__ void item_weights(const PmeArgs &a, int it, const int (&org)[3], T (&w)[3][NI],
                                             T (&dw)[3][NI], int (&l)[3]) {
#pragma unroll
  for (int d = 0; d < 3; d++) {
    T mn[NI + 1], dmn[NI + 1];
    bspline<NI, DERIV, T>((T)a.item_f[3 * (size_t)it + d], mn, dmn);
    l[d] = a.item_s[3 * (size_t)it + d] - org[d];
#pragma unroll
    for (int m = 0; m < NI; m++) {
      w[d][m] = mn[NI - m];
      if (DERIV) dw[d][m] = dmn[NI - m];
    }
  }
}

// one workgroup per (tile, bead), one lane per charge of the bin; the
// charge's n^3 products go into the LDS halo box with LDS atomics.  The box
// accumulates in FP64 for both grid types: the weights are formed in T, and
// the FP32 path rounds the box once when it is stored (FP32 LDS atomics
// measured 5x slower than FP64 ones on MI355X for this access pattern).
template <int NI, typename T>
__global__ void __launch_bounds__(spread_lanes<T>() * PME_BLOCK) k_pme_spread(PmeArgs a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double *box = (double *)lds_raw;
  const int hvol = (int)a.hvol;
  int tl, b;
  if (!xcd_unit_bead(a.ntile, a.P, tl, b)) return;  // uniform over the workgroup
  const int tid = threadIdx.x;
  for (int k = tid; k < hvol; k += blockDim.x) box[k] = 0.0;
  const size_t bin = (size_t)b * a.ntile + tl;
  const int lo = a.bin_off[bin], cnt = a.bin_cnt[bin];
  int org[3];
  tile_origin(a, tl, org[0], org[1], org[2]);
  const int H0 = a.H[0], H01 = a.H[0] * a.H[1];
  __syncthreads();
  // L lanes per charge, each adding every L-th row (mc, mb) of the support
  // (more wavefronts in flight for the LDS atomics)
  constexpr int L = spread_lanes<T>();
  const int part = tid % L;
  for (int c = tid / L; c < cnt; c += blockDim.x / L) {
    const int it = a.sorted[lo + c];
    const T q = (T)a.q[a.ichrg[it % a.nchrg]];
    T w[3][NI], dw[3][NI];
    int l[3];
    item_weights<NI, false, T>(a, it, org, w, dw, l);
#pragma unroll
    for (int m = 0; m < NI; m++) w[0][m] *= q;
    for (int rw = part; rw < NI * NI; rw += L) {
      const int mc = rw / NI, mb = rw - mc * NI;
      const T wbc = w[1][mb] * w[2][mc];
      double *row = box + (l[2] + mc) * H01 + (l[1] + mb) * H0 + l[0];
#pragma unroll
      for (int ma = 0; ma < NI; ma++) atomicAdd(row + ma, (double)(w[0][ma] * wbc));
    }
  }
  __syncthreads();
  T *dst = (T *)a.hbuf + bin * a.hvol;
  for (int k = tid; k < hvol; k += blockDim.x) dst[k] = (T)box[k];
}

// --- gather ---------------------------------------------------------------------

// The potential on the halo box of tile origin org, bead b, into LDS, with
// batches of 8 loads in flight per lane.
template <typename T>
__device__ __forceinline__ void load_halo(const PmeArgs &a, int b, const int (&org)[3], T *box) {
  const int hvol = (int)a.hvol;
  const int Ka = a.K[0], Kb = a.K[1], Kc = a.K[2];
  const int H0 = a.H[0], H1 = a.H[1], H01 = H0 * H1;
  const T *g = (const T *)a.grid + (size_t)b * a.gstride;
  for (int k0 = threadIdx.x; k0 < hvol; k0 += 8 * blockDim.x) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int k = k0 + u * blockDim.x;
      v[u] = (T)0;
      if (k < hvol) {
        const int q0 = (int)__umulhi((unsigned)k, a.mH0), lc = (int)__umulhi((unsigned)q0, a.mH1);
        const int la = k - q0 * H0, lb = q0 - lc * H1;
        // org < K and l < H <= K + n - 1, so org + l < 3 K (n - 1 <= K on
        // every grid): two conditional subtractions wrap it into the grid
        int ga = org[0] + la, gb = org[1] + lb, gc = org[2] + lc;
        ga -= ga >= Ka ? Ka : 0;
        gb -= gb >= Kb ? Kb : 0;
        gc -= gc >= Kc ? Kc : 0;
        ga -= ga >= Ka ? Ka : 0;
        gb -= gb >= Kb ? Kb : 0;
        gc -= gc >= Kc ? Kc : 0;
        v[u] = g[((size_t)gc * Kb + gb) * Ka + ga];
      }
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int k = k0 + u * blockDim.x;
      if (k < hvol) box[k] = v[u];
    }
  }
}

// lanes per charge in the gather: with FP32 grids two, each summing half
// of the support planes (twice the wavefronts on the LDS the halo box
// allows); with FP64 one
template <typename T>
constexpr int gather_lanes() {
  return sizeof(T) == 4 ? 2 : 1;
}

// one workgroup per (tile, bead): the potential on the halo box into LDS,
// then L lanes per charge (ewald3d_pme.c:534-560); lane part sums the planes
// mc = part * NI / 2 ... (part + 1) * NI / 2 - 1 and the pair adds its sums
template <int NI, typename T>
__global__ void __launch_bounds__(gather_lanes<T>() * PME_BLOCK) k_pme_gather(PmeArgs a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *box = (T *)lds_raw;
  int tl, b;
  if (!xcd_unit_bead(a.ntile, a.P, tl, b)) return;
  const int tid = threadIdx.x;
  const size_t bin = (size_t)b * a.ntile + tl;
  const int lo = a.bin_off[bin], cnt = a.bin_cnt[bin];
  if (cnt == 0) return;  // uniform over the workgroup
  int org[3];
  tile_origin(a, tl, org[0], org[1], org[2]);
  const int Ka = a.K[0], Kb = a.K[1], Kc = a.K[2];
  const int H0 = a.H[0], H01 = H0 * a.H[1];
  load_halo<T>(a, b, org, box);
  __syncthreads();
  const double *hi = a.cell.hi;
  constexpr int L = gather_lanes<T>();
  constexpr int HP = L == 1 ? 0 : NI / 2;  // planes of lane 0 when split
  constexpr int KP = NI - HP;              // planes per lane (lane 0 pads odd orders with zero weights)
  const int part = L == 1 ? 0 : (tid & 1);
  for (int c = tid / L; c < cnt; c += blockDim.x / L) {
    const int it = a.sorted[lo + c];
    const int ic = it % a.nchrg, bead = it / a.nchrg;
    const int i = a.ichrg[ic];
    T w[3][NI], dw[3][NI];
    int l[3];
    item_weights<NI, true, T>(a, it, org, w, dw, l);
    // this lane's planes: weights picked once per charge
    T w2[KP], dw2[KP];
#pragma unroll
    for (int k = 0; k < KP; k++) {
      if (L == 1) {
        w2[k] = w[2][k];
        dw2[k] = dw[2][k];
      } else {
        const T a0 = k < HP ? w[2][k] : (T)0, d0 = k < HP ? dw[2][k] : (T)0;
        w2[k] = part ? w[2][HP + k] : a0;
        dw2[k] = part ? dw[2][HP + k] : d0;
      }
    }
    const int mc0 = part * HP;
    // row sums in T, accumulated over rows in FP64
    double da = 0.0, db = 0.0, dc = 0.0;
#pragma unroll
    for (int k = 0; k < KP; k++) {
      const int mc = min(mc0 + k, NI - 1);
#pragma unroll
      for (int mb = 0; mb < NI; mb++) {
        const T *row = box + (l[2] + mc) * H01 + (l[1] + mb) * H0 + l[0];
        T r0 = (T)0, r1 = (T)0;
#pragma unroll
        for (int ma = 0; ma < NI; ma++) {
          const T phi = row[ma];
          r0 += dw[0][ma] * phi;
          r1 += w[0][ma] * phi;
        }
        da += (double)(r0 * (w[1][mb] * w2[k]));
        db += (double)(r1 * (dw[1][mb] * w2[k]));
        dc += (double)(r1 * (w[1][mb] * dw2[k]));
      }
    }
    if (L == 2) {  // the pair's lanes hold the same charge
      da += __shfl_xor(da, 1, 64);
      db += __shfl_xor(db, 1, 64);
      dc += __shfl_xor(dc, 1, 64);
    }
    if (part == 0) {
      const double q = a.q[i];
      da *= (double)Ka * q;
      db *= (double)Kb * q;
      dc *= (double)Kc * q;
      // bead-major, so the lines of one bead are written from its own XCD
      const size_t ib = (size_t)bead * a.N + i;
      a.rx[ib] = -(da * hi[0] + db * hi[1] + dc * hi[2]);
      a.ry[ib] = -(da * hi[3] + db * hi[4] + dc * hi[5]);
      a.rz[ib] = -(da * hi[6] + db * hi[7] + dc * hi[8]);
    }
  }
}

// The k-space part of dE_b/dq_i: the convolved grid (the derivative of the
// k-space energy with respect to the grid charges) interpolated at the charge
// with the B-spline weights themselves (the gather's weights without the
// derivative).  One workgroup per (tile, bead), one lane per charge.
template <int NI, typename T>
__global__ void __launch_bounds__(PME_BLOCK) k_pme_phi(PmeArgs a) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *box = (T *)lds_raw;
  int tl, b;
  if (!xcd_unit_bead(a.ntile, a.P, tl, b)) return;
  const int tid = threadIdx.x;
  const size_t bin = (size_t)b * a.ntile + tl;
  const int lo = a.bin_off[bin], cnt = a.bin_cnt[bin];
  if (cnt == 0) return;  // uniform over the workgroup
  int org[3];
  tile_origin(a, tl, org[0], org[1], org[2]);
  const int H0 = a.H[0], H01 = H0 * a.H[1];
  load_halo<T>(a, b, org, box);
  __syncthreads();
  for (int c = tid; c < cnt; c += blockDim.x) {
    const int it = a.sorted[lo + c];
    const int ic = it % a.nchrg, bead = it / a.nchrg;
    const int i = a.ichrg[ic];
    T w[3][NI], dw[3][NI];
    int l[3];
    item_weights<NI, false, T>(a, it, org, w, dw, l);
    double v = 0.0;
    for (int mc = 0; mc < NI; mc++) {
#pragma unroll
      for (int mb = 0; mb < NI; mb++) {
        const T *row = box + (l[2] + mc) * H01 + (l[1] + mb) * H0 + l[0];
        T r = (T)0;
#pragma unroll
        for (int ma = 0; ma < NI; ma++) r += w[0][ma] * row[ma];
        v += (double)(r * (w[1][mb] * w[2][mc]));
      }
    }
    a.phi[(size_t)i * a.P + bead] += v;
  }
}

// LDS of the gather (the halo box in the grid type) and of the spread (the
// halo box in FP64)
size_t pme_lds_bytes(const PmeArgs &a, bool gather) {
  return a.hvol * (gather && a.f32 ? sizeof(float) : sizeof(double));
}

#define PMX_PME_CASES(X) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)

template <typename T>
hipError_t launch_spread_t(const PmeArgs &a, hipStream_t st) {
  const dim3 grid(xcd_grid(a.ntile, a.P)), block(spread_lanes<T>() * PME_BLOCK);
  const size_t lds = pme_lds_bytes(a, false);
  switch (a.n) {
#define X(NI) \
  case NI:                                                                                                  \
    if (lds > 65536)                                                                                          \
      (void)hipFuncSetAttribute((const void *)k_pme_spread<NI, T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pme_spread<NI, T>), grid, block, lds, st, a);                                \
    break;
    PMX_PME_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_gather_t(const PmeArgs &a, hipStream_t st) {
  const dim3 grid(xcd_grid(a.ntile, a.P)), block(gather_lanes<T>() * PME_BLOCK);
  const size_t lds = pme_lds_bytes(a, true);
  switch (a.n) {
#define X(NI) \
  case NI:                                                                                                  \
    if (lds > 65536)                                                                                          \
      (void)hipFuncSetAttribute((const void *)k_pme_gather<NI, T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pme_gather<NI, T>), grid, block, lds, st, a);                                \
    break;
    PMX_PME_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// spread into halo boxes (the forward FFT folds them into the grid)
hipError_t launch_pme_spread(const PmeArgs &a, hipStream_t st) {
  return a.f32 ? launch_spread_t<float>(a, st) : launch_spread_t<double>(a, st);
}

template <typename T>
hipError_t launch_phi_t(const PmeArgs &a, hipStream_t st) {
  const dim3 grid(xcd_grid(a.ntile, a.P)), block(PME_BLOCK);
  const size_t lds = pme_lds_bytes(a, true);
  switch (a.n) {
#define X(NI) \
  case NI:                                                                                                  \
    if (lds > 65536)                                                                                          \
      (void)hipFuncSetAttribute((const void *)k_pme_phi<NI, T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pme_phi<NI, T>), grid, block, lds, st, a);                                \
    break;
    PMX_PME_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_pme_phi(const PmeArgs &a, hipStream_t st) {
  return a.f32 ? launch_phi_t<float>(a, st) : launch_phi_t<double>(a, st);
}

// --- standard Ewald ---------------------------------------------------------------

// Structure factors, ewald3d_recip (ewald3d.c:27-373) with direct phases:
// one thread per (k, bead) sums q cos and q sin of 2 pi k.s over LDS tiles
// of the charges' fractional coordinates s = h^-1 r, then the k term's
// energy preg |S|^2 and pressure tensor go into block partial rows (the
// energy also on the diagonal, as the PME convolution writes them).
__global__ void __launch_bounds__(EWALD_BLOCK) k_ewald_sk(EwaldArgs a) {
  __shared__ double4 sq[EWALD_BLOCK];
  const int tid = threadIdx.x, b = blockIdx.y;
  const int k = blockIdx.x * EWALD_BLOCK + tid;
  const bool on = k < a.nk;
  const double4 kv = on ? a.kv[k] : make_double4(0, 0, 0, 0);
  const double *hi = a.cell.hi;
  double sr = 0.0, si = 0.0;
  for (int c0 = 0; c0 < a.nchrg; c0 += EWALD_BLOCK) {
    __syncthreads();
    if (c0 + tid < a.nchrg) {
      const int i = a.ichrg[c0 + tid];
      const size_t ib = (size_t)i * a.P + b;
      const double x = a.x[ib], y = a.y[ib], z = a.z[ib];
      sq[tid] = make_double4(x * hi[0] + y * hi[3] + z * hi[6], x * hi[1] + y * hi[4] + z * hi[7],
                             x * hi[2] + y * hi[5] + z * hi[8], a.q[i]);
    }
    __syncthreads();
    const int n = min(EWALD_BLOCK, a.nchrg - c0);
    for (int m = 0; m < n; m++) {
      const double4 c = sq[m];
      double sn, cs;
      sincos(kv.x * c.x + kv.y * c.y + kv.z * c.z, &sn, &cs);
      sr += cs * c.w;
      si += sn * c.w;
    }
  }
  double vals[FFT_NV] = {0, 0, 0, 0, 0, 0, 0};
  if (on) {
    const double4 kg = a.kg[k];
    const double smag = sr * sr + si * si;
    const double e = smag * kv.w, pr = kg.w * smag;
    vals[0] = e;
    vals[1] = pr * kg.x * kg.x + e;
    vals[2] = pr * kg.y * kg.y + e;
    vals[3] = pr * kg.z * kg.z + e;
    vals[4] = pr * kg.x * kg.y;
    vals[5] = pr * kg.x * kg.z;
    vals[6] = pr * kg.y * kg.z;
    a.sk[(size_t)b * a.nk + k] = make_double2(2.0 * kv.w * sr, 2.0 * kv.w * si);
  }
  block_partials<FFT_NV>(vals, 1, b, a.P, a.part, blockIdx.x, gridDim.x);
}

// Forces: one thread per (charge, bead) sums, over LDS tiles of the k list
// in ascending order, k (sin(k.r) 2 preg Re S - cos(k.r) 2 preg Im S), and
// adds q times it to the force (each (charge, bead) is one thread's).
__global__ void __launch_bounds__(EWALD_BLOCK) k_ewald_force(EwaldArgs a) {
  __shared__ double4 skv[EWALD_BLOCK], skg[EWALD_BLOCK];
  __shared__ double2 ssk[EWALD_BLOCK];
  const int tid = threadIdx.x, b = blockIdx.y;
  const int c = blockIdx.x * EWALD_BLOCK + tid;
  const bool on = c < a.nchrg;
  const double *hi = a.cell.hi;
  int i = 0;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  if (on) {
    i = a.ichrg[c];
    const size_t ib = (size_t)i * a.P + b;
    const double x = a.x[ib], y = a.y[ib], z = a.z[ib];
    s0 = x * hi[0] + y * hi[3] + z * hi[6];
    s1 = x * hi[1] + y * hi[4] + z * hi[7];
    s2 = x * hi[2] + y * hi[5] + z * hi[8];
  }
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int k0 = 0; k0 < a.nk; k0 += EWALD_BLOCK) {
    __syncthreads();
    if (k0 + tid < a.nk) {
      skv[tid] = a.kv[k0 + tid];
      skg[tid] = a.kg[k0 + tid];
      ssk[tid] = a.sk[(size_t)b * a.nk + k0 + tid];
    }
    __syncthreads();
    const int n = min(EWALD_BLOCK, a.nk - k0);
    for (int m = 0; m < n; m++) {
      double sn, cs;
      sincos(skv[m].x * s0 + skv[m].y * s1 + skv[m].z * s2, &sn, &cs);
      const double fd = sn * ssk[m].x - cs * ssk[m].y;
      fx += skg[m].x * fd;
      fy += skg[m].y * fd;
      fz += skg[m].z * fd;
    }
  }
  if (on) {  // bead-major, added to the other forces after the pair kernel
    const size_t ib = (size_t)b * a.N + i;
    const double q = a.q[i];
    a.rx[ib] = fx * q;
    a.ry[ib] = fy * q;
    a.rz[ib] = fz * q;
  }
}

hipError_t launch_ewald(const EwaldArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(k_ewald_sk, dim3((a.nk + EWALD_BLOCK - 1) / EWALD_BLOCK, a.P), dim3(EWALD_BLOCK), 0, st, a);
  hipLaunchKernelGGL(k_ewald_force, dim3((a.nchrg + EWALD_BLOCK - 1) / EWALD_BLOCK, a.P), dim3(EWALD_BLOCK), 0, st,
                     a);
  return hipGetLastError();
}

hipError_t launch_pme_gather(const PmeArgs &a, hipStream_t st) {
  return a.f32 ? launch_gather_t<float>(a, st) : launch_gather_t<double>(a, st);
}

// --- FFT --------------------------------------------------------------------------

// Batched 1D complex FFTs held in LDS, cooperative over the workgroup:
// mixed-radix Stockham passes (radix 2, 3, 4, 5 — the radices of FFTPACK,
// fft_generic.f:14-84 — and 8), unnormalised, sign -1 for exp(-i) (DCFFTF) and +1 for
// exp(+i) (DCFFTB).  Line l element e lives at buf[l*dist + e*stride].  A
// pass reads every butterfly's inputs into registers, waits, then writes the
// outputs in place, so no permutation pass is needed.  Lanes run along the
// contiguous direction (lines when stride > 1), which keeps LDS accesses
// free of bank conflicts.  nlines * K <= MAXP and blockDim.x == NT (the
// plane classes of launch_fft_t).
template <int R, typename T>
__device__ __forceinline__ void dft_small(typename Cx<T>::C (&v)[R], int sign) {
  using C = typename Cx<T>::C;
  if constexpr (R == 2) {
    const C a = v[0], b = v[1];
    v[0] = Cx<T>::make(a.x + b.x, a.y + b.y);
    v[1] = Cx<T>::make(a.x - b.x, a.y - b.y);
  } else if constexpr (R == 4) {
    const C a0 = Cx<T>::make(v[0].x + v[2].x, v[0].y + v[2].y);
    const C a1 = Cx<T>::make(v[0].x - v[2].x, v[0].y - v[2].y);
    const C b0 = Cx<T>::make(v[1].x + v[3].x, v[1].y + v[3].y);
    const C b1 = Cx<T>::make(v[1].x - v[3].x, v[1].y - v[3].y);
    // b1 * (sign * i)
    const C c1 = sign < 0 ? Cx<T>::make(b1.y, -b1.x) : Cx<T>::make(-b1.y, b1.x);
    v[0] = Cx<T>::make(a0.x + b0.x, a0.y + b0.y);
    v[2] = Cx<T>::make(a0.x - b0.x, a0.y - b0.y);
    v[1] = Cx<T>::make(a1.x + c1.x, a1.y + c1.y);
    v[3] = Cx<T>::make(a1.x - c1.x, a1.y - c1.y);
  } else if constexpr (R == 8) {
    // two 4-point DFTs of the even and odd inputs, combined with the
    // eighth roots w^k = exp(sign 2 pi i k / 8): X[k] = E[k] + w^k O[k],
    // X[k + 4] = E[k] - w^k O[k]
    C e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft_small<4, T>(e, sign);
    dft_small<4, T>(o, sign);
    const T h = (T)0.70710678118654752440, sg = (T)sign;
    const C w1 = Cx<T>::make(h * (o[1].x - sg * o[1].y), h * (o[1].y + sg * o[1].x));    // (1 + sign i) h
    const C w2 = Cx<T>::make(-sg * o[2].y, sg * o[2].x);                                // sign i
    const C w3 = Cx<T>::make(-h * (o[3].x + sg * o[3].y), h * (sg * o[3].x - o[3].y));  // (-1 + sign i) h
    v[0] = Cx<T>::make(e[0].x + o[0].x, e[0].y + o[0].y);
    v[4] = Cx<T>::make(e[0].x - o[0].x, e[0].y - o[0].y);
    v[1] = Cx<T>::make(e[1].x + w1.x, e[1].y + w1.y);
    v[5] = Cx<T>::make(e[1].x - w1.x, e[1].y - w1.y);
    v[2] = Cx<T>::make(e[2].x + w2.x, e[2].y + w2.y);
    v[6] = Cx<T>::make(e[2].x - w2.x, e[2].y - w2.y);
    v[3] = Cx<T>::make(e[3].x + w3.x, e[3].y + w3.y);
    v[7] = Cx<T>::make(e[3].x - w3.x, e[3].y - w3.y);
  } else {
    // direct DFT: out[k] = sum_m v[m] exp(sign 2 pi i k m / R)
    constexpr double c3[3] = {1.0, -0.5, -0.5};
    constexpr double s3[3] = {0.0, 0.86602540378443864676, -0.86602540378443864676};
    constexpr double c5[5] = {1.0, 0.30901699437494742410, -0.80901699437494742410, -0.80901699437494742410,
                              0.30901699437494742410};
    constexpr double s5[5] = {0.0, 0.95105651629515357212, 0.58778525229247312917, -0.58778525229247312917,
                              -0.95105651629515357212};
    C out[R];
#pragma unroll
    for (int k = 0; k < R; k++) {
      T re = (T)0, im = (T)0;
#pragma unroll
      for (int m = 0; m < R; m++) {
        const int km = (k * m) % R;
        const T cr = (T)((R == 3) ? c3[km] : c5[km]);
        const T ci = (T)(((R == 3) ? s3[km] : s5[km]) * (double)sign);
        re += v[m].x * cr - v[m].y * ci;
        im += v[m].x * ci + v[m].y * cr;
      }
      out[k] = Cx<T>::make(re, im);
    }
#pragma unroll
    for (int k = 0; k < R; k++) v[k] = out[k];
  }
}

// k / d for 0 <= k, k * d < 2^32 as a multiply-high by m = floor(2^32 / d) + 1,
// m formed once per divisor (2^32 / d is at least 1/d from an integer unless
// exact, so its double quotient floors exactly); d = 1, whose m does not fit
// 32 bits, is taken apart.
struct FastDiv {
  unsigned m;
  int d;
  __device__ explicit FastDiv(int d_)
      : m(d_ > 1 ? (unsigned)(unsigned long long)(4294967296.0 / (double)d_) + 1u : 0u), d(d_) {}
  __device__ int q(int k) const { return d > 1 ? (int)__umulhi((unsigned)k, m) : k; }
};

template <int R, typename T, int NT, int MAXP>
__device__ void stockham_pass(typename Cx<T>::C *buf, int K, int nlines, int stride, int dist, int Ns, int sign,
                              const typename Cx<T>::C *__restrict__ tw) {
  using C = typename Cx<T>::C;
  constexpr int MAXB = (MAXP / R + NT - 1) / NT;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int KR = K / R, nb = nlines * KR;
  const bool lines_fast = stride != 1;
  // butterfly bi -> (line l, index j): lines run fastest when stride > 1
  const FastDiv dl(lines_fast ? nlines : KR), dn(Ns);
  auto line_index = [&](int bi, int &l, int &j) {
    const int qt = dl.q(bi), rm = bi - qt * dl.d;
    l = lines_fast ? rm : qt;
    j = lines_fast ? qt : rm;
  };
  // element offsets: the k-th input of a butterfly sits k * KR * stride
  // after its first, the k-th output k * Ns * stride (uniform steps, so the
  // per-lane index work is one multiply-add per butterfly and side)
  const int kin = KR * stride, kout = Ns * stride;
  C v[MAXB][R];
  int lq[MAXB], jqv[MAXB];
#pragma unroll
  for (int q = 0; q < MAXB; q++) {
    const int bi = tid + q * nt;
    if (bi < nb) {
      int l, j;
      line_index(bi, l, j);
      lq[q] = l;
      jqv[q] = j;
      const C *p = buf + (l * dist + j * stride);
#pragma unroll
      for (int k = 0; k < R; k++) v[q][k] = p[k * kin];
    }
  }
  __syncthreads();
  const int tstep = K / (Ns * R);
#pragma unroll
  for (int q = 0; q < MAXB; q++) {
    const int bi = tid + q * nt;
    if (bi < nb) {
      const int l = lq[q], j = jqv[q];
      const int jq = dn.q(j), jm = j - jq * Ns;
      const int jt = jm * tstep;
      int ti = 0;
#pragma unroll
      for (int k = 1; k < R; k++) {  // twiddle exp(sign 2 pi i jm k / (Ns R))
        ti += jt;
        const C t = tw[ti];  // (cos, -sin) of 2 pi m / K, m = jm k tstep
        const T wr = t.x, wi = sign < 0 ? t.y : -t.y;
        const C a = v[q][k];
        v[q][k] = Cx<T>::make(a.x * wr - a.y * wi, a.x * wi + a.y * wr);
      }
      dft_small<R, T>(v[q], sign);
      C *p = buf + (l * dist + (jq * Ns * R + jm) * stride);
#pragma unroll
      for (int k = 0; k < R; k++) p[k * kout] = v[q][k];
    }
  }
  __syncthreads();
}

template <typename T, int NT, int MAXP>
__device__ void lds_fft(typename Cx<T>::C *buf, int K, int nlines, int stride, int dist, int sign,
                        const typename Cx<T>::C *tw, int npass, const int *radix) {
  int Ns = 1;
  for (int s = 0; s < npass; s++) {
    switch (radix[s]) {
      case 8:  // FP32 only (the host plans radix 8 for FP32 grids)
        if constexpr (sizeof(T) == 4) stockham_pass<8, T, NT, MAXP>(buf, K, nlines, stride, dist, Ns, sign, tw);
        break;
      case 2: stockham_pass<2, T, NT, MAXP>(buf, K, nlines, stride, dist, Ns, sign, tw); break;
      case 3: stockham_pass<3, T, NT, MAXP>(buf, K, nlines, stride, dist, Ns, sign, tw); break;
      case 4: stockham_pass<4, T, NT, MAXP>(buf, K, nlines, stride, dist, Ns, sign, tw); break;
      default: stockham_pass<5, T, NT, MAXP>(buf, K, nlines, stride, dist, Ns, sign, tw); break;
    }
    Ns *= radix[s];
  }
}

// stockham_pass with every size a compile-time constant (transform length
// K, NL lines, element stride ST, line distance DS, sub-transform length NS,
// NT lanes): the same loads, twiddles, butterflies and stores in the same
// order, with the index arithmetic folded to shifts and constants.  NS = 1
// multiplies by no twiddle (all of them are 1 there, also in stockham_pass,
// exactly).
template <int R, typename T, int NT, int K, int NL, int ST, int DS, int NS>
__device__ __forceinline__ void stockham_pass_c(typename Cx<T>::C *buf, int sign,
                                                const typename Cx<T>::C *__restrict__ tw) {
  using C = typename Cx<T>::C;
  constexpr int KR = K / R, NB = NL * KR, MAXB = (NB + NT - 1) / NT;
  constexpr bool LINES_FAST = ST != 1;
  constexpr int KIN = KR * ST, KOUT = NS * ST, TSTEP = K / (NS * R);
  const int tid = threadIdx.x;
  C v[MAXB][R];
#pragma unroll
  for (int q = 0; q < MAXB; q++) {
    const int bi = tid + q * NT;
    if (NB % NT == 0 || bi < NB) {
      const int l = LINES_FAST ? bi % NL : bi / KR, j = LINES_FAST ? bi / NL : bi % KR;
      const C *p = buf + (l * DS + j * ST);
#pragma unroll
      for (int k = 0; k < R; k++) v[q][k] = p[k * KIN];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXB; q++) {
    const int bi = tid + q * NT;
    if (NB % NT == 0 || bi < NB) {
      const int l = LINES_FAST ? bi % NL : bi / KR, j = LINES_FAST ? bi / NL : bi % KR;
      const int jq = j / NS, jm = j % NS;
      if (NS > 1) {
        const int jt = jm * TSTEP;
#pragma unroll
        for (int k = 1; k < R; k++) {  // twiddle exp(sign 2 pi i jm k / (NS R))
          const C t = tw[jt * k];
          const T wr = t.x, wi = sign < 0 ? t.y : -t.y;
          const C a = v[q][k];
          v[q][k] = Cx<T>::make(a.x * wr - a.y * wi, a.x * wi + a.y * wr);
        }
      }
      dft_small<R, T>(v[q], sign);
      C *p = buf + (l * DS + (jq * NS * R + jm) * ST);
#pragma unroll
      for (int k = 0; k < R; k++) p[k * KOUT] = v[q][k];
    }
  }
  __syncthreads();
}

// A 64-point transform of NL lines as two radix-8 passes with constant
// sizes: what lds_fft does for a plan of radices (8, 8).
template <typename T, int NT, int NL, int ST, int DS>
__device__ __forceinline__ void lds_fft64(typename Cx<T>::C *buf, int sign, const typename Cx<T>::C *tw) {
  stockham_pass_c<8, T, NT, 64, NL, ST, DS, 1>(buf, sign, tw);
  stockham_pass_c<8, T, NT, 64, NL, ST, DS, 8>(buf, sign, tw);
}

// the plan of axis d is radix 8 twice (a length of 64; FP32 grids)
__device__ __forceinline__ bool plan_88(const FftArgs &a, int d) {
  return a.npass[d] == 2 && a.radix[d][0] == 8 && a.radix[d][1] == 8;
}

template <typename T>
__device__ __forceinline__ const typename Cx<T>::C *twiddles(const FftArgs &a, int d) {
  if constexpr (sizeof(T) == 8)
    return a.tw[d];
  else
    return a.twf[d];
}

// The K[d] twiddles of axis d copied into LDS at dst: the passes read them at
// lane-varying indices, which LDS serves without a vector-memory round trip.
template <typename T>
__device__ __forceinline__ void stage_twiddles(const FftArgs &a, int d, typename Cx<T>::C *dst) {
  const typename Cx<T>::C *src = twiddles<T>(a, d);
  for (int k = threadIdx.x; k < a.K[d]; k += blockDim.x) dst[k] = src[k];
}

// Plane c of bead b's spread charge into LDS as complex values, folded from
// the halo boxes: along each axis, point l of tile t is held by the box of
// tile t - o (periodically) at l + o*T for every o with l + o*T < H.  Tiles
// divide the grid, so the offset is exact also across the periodic boundary.
// With at most two boxes per axis (T >= n - 1) every point makes the same 8
// loads, the absent overlaps re-reading its own element and adding zero, so
// the loads of 4 points per lane are in flight together; smaller tiles take
// the loop over all overlaps.  Sums run in the order (c, b, a) of the boxes.
template <typename T>
__device__ __forceinline__ void fold_plane(const FftArgs &a, int c, int b, typename Cx<T>::C *sm) {
  const int Ka = a.K[0], npt = Ka * a.K[1];
  const int T0 = a.T[0], T1 = a.T[1], T2 = a.T[2];
  const int H0 = a.H[0], H1 = a.H[1], H2 = a.H[2], H01 = H0 * H1;
  const int nt0 = a.nt[0], nt1 = a.nt[1], nt2 = a.nt[2];
  const int hv = (int)a.hvol;
  const T *hb = (const T *)a.hbuf + (size_t)b * a.ntile * a.hvol;
  const int tc = c / T2, lc = c - tc * T2;
  if (a.nbr[0] <= 2 && a.nbr[1] <= 2 && a.nbr[2] <= 2) {
    const bool ec = lc + T2 < H2;
    const int rc0 = tc * nt1, rc1 = ec ? (tc ? tc - 1 : nt2 - 1) * nt1 : rc0;
    const int oc0 = lc * H01, oc1 = ec ? (lc + T2) * H01 : oc0;
    for (int w0 = threadIdx.x; w0 < npt; w0 += 4 * blockDim.x) {
      T v[4][8];
      unsigned m[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int w = min(w0 + u * (int)blockDim.x, npt - 1);
        const int gb = (int)__umulhi((unsigned)w, a.mKa), ga = w - gb * Ka;
        const int ta = (int)__umulhi((unsigned)ga, a.mT0), la = ga - ta * T0;
        const int tb = (int)__umulhi((unsigned)gb, a.mT1), lb = gb - tb * T1;
        const bool ea = la + T0 < H0, eb = lb + T1 < H1;
        const int sa1 = ea ? (ta ? ta - 1 : nt0 - 1) : ta, sb1 = eb ? (tb ? tb - 1 : nt1 - 1) : tb;
        const int xa = ea ? T0 : 0, xb = eb ? T1 * H0 : 0, e = lb * H0 + la;
        m[u] = (ea ? 1u : 0u) | (eb ? 2u : 0u);
        v[u][0] = hb[((rc0 + tb) * nt0 + ta) * hv + oc0 + e];
        v[u][1] = hb[((rc0 + tb) * nt0 + sa1) * hv + oc0 + e + xa];
        v[u][2] = hb[((rc0 + sb1) * nt0 + ta) * hv + oc0 + e + xb];
        v[u][3] = hb[((rc0 + sb1) * nt0 + sa1) * hv + oc0 + e + xa + xb];
        v[u][4] = hb[((rc1 + tb) * nt0 + ta) * hv + oc1 + e];
        v[u][5] = hb[((rc1 + tb) * nt0 + sa1) * hv + oc1 + e + xa];
        v[u][6] = hb[((rc1 + sb1) * nt0 + ta) * hv + oc1 + e + xb];
        v[u][7] = hb[((rc1 + sb1) * nt0 + sa1) * hv + oc1 + e + xa + xb];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int w = w0 + u * (int)blockDim.x;
        const bool ea = m[u] & 1u, eb = m[u] & 2u;
        T s = v[u][0];
        s += ea ? v[u][1] : (T)0;
        s += eb ? v[u][2] : (T)0;
        s += (ea && eb) ? v[u][3] : (T)0;
        s += ec ? v[u][4] : (T)0;
        s += (ec && ea) ? v[u][5] : (T)0;
        s += (ec && eb) ? v[u][6] : (T)0;
        s += (ec && ea && eb) ? v[u][7] : (T)0;
        if (w < npt) sm[w] = Cx<T>::make(s, (T)0);
      }
    }
    return;
  }
  for (int w = threadIdx.x; w < npt; w += blockDim.x) {
    const int gb = w / Ka, ga = w - gb * Ka;
    const int ta = ga / T0, la = ga - ta * T0, tb = gb / T1, lb = gb - tb * T1;
    T s = (T)0;
    for (int oc = 0; oc < a.nbr[2] && lc + oc * T2 < H2; oc++) {
      const int sc = (tc - oc % nt2 + nt2) % nt2;
      for (int ob = 0; ob < a.nbr[1] && lb + ob * T1 < H1; ob++) {
        const int sb = (tb - ob % nt1 + nt1) % nt1;
        for (int oa = 0; oa < a.nbr[0] && la + oa * T0 < H0; oa++) {
          const int sa = (ta - oa % nt0 + nt0) % nt0;
          s += hb[((sc * nt1 + sb) * nt0 + sa) * hv + (lc + oc * T2) * H01 + (lb + ob * T1) * H0 + la + oa * T0];
        }
      }
    }
    sm[w] = Cx<T>::make(s, (T)0);
  }
}

// forward transform of one c-plane of one bead: the folded charge plane,
// real along a, then b.  LDS holds the plane as Kb lines of Ka complex
// values; only a' = 0..Ka/2 is kept (Hermitian half).  spec layout:
// [bead][c][b][a'] complex.
template <typename T, int NT, int MAXP>
__global__ void __launch_bounds__(NT) k_fft_fwd_ab(FftArgs a) {
  using C = typename Cx<T>::C;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  C *sm = (C *)lds_raw;
  int c, b;
  if (!xcd_unit_bead(a.K[2], a.P, c, b)) return;
  const int Ka = a.K[0], Kb = a.K[1], Ha = Ka / 2 + 1;
  C *twa = sm + Ka * Kb, *twb = twa + Ka;  // twiddles after the plane
  stage_twiddles<T>(a, 0, twa);
  stage_twiddles<T>(a, 1, twb);
  fold_plane<T>(a, c, b, sm);
  __syncthreads();
  if (sizeof(T) == 4 && NT == 256 && plan_88(a, 0) && plan_88(a, 1)) {  // 64 x 64 planes
    lds_fft64<T, NT, 64, 1, 64>(sm, -1, twa);
    lds_fft64<T, NT, 33, 64, 1>(sm, -1, twb);
  } else {
    lds_fft<T, NT, MAXP>(sm, Ka, Kb, 1, Ka, -1, twa, a.npass[0], a.radix[0]);
    lds_fft<T, NT, MAXP>(sm, Kb, Ha, Ka, 1, -1, twb, a.npass[1], a.radix[1]);
  }
  C *sp = (C *)a.spec + (size_t)b * a.sstride + (size_t)c * Kb * Ha;
  const FastDiv dh(Ha);
  for (int w = threadIdx.x; w < Kb * Ha; w += blockDim.x) {
    const int bb = dh.q(w), aa = w - bb * Ha;
    sp[w] = sm[bb * Ka + aa];
  }
}

// along c for one (bead, b) row: forward FFT, influence function over the
// half-space sphere list (ewald3d_pme.c:367-430): energy, virial, and the
// product with preg*bweight, then the inverse FFT.  Energy and virial are
// summed in FP64 into one partial row per (b row, bead).
template <typename T, int NT, int MAXP>
__global__ void __launch_bounds__(NT) k_fft_c_conv(FftArgs a) {
  using C = typename Cx<T>::C;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  C *sm = (C *)lds_raw;
  int bb, b;
  if (!xcd_unit_bead(a.K[1], a.P, bb, b)) return;
  const int Ka = a.K[0], Kb = a.K[1], Kc = a.K[2], Ha = Ka / 2 + 1;
  C *twc = sm + Kc * Ha;  // twiddles after the lines
  stage_twiddles<T>(a, 2, twc);
  C *sp = (C *)a.spec + (size_t)b * a.sstride;
  const FastDiv dh(Ha);
  for (int w0 = threadIdx.x; w0 < Kc * Ha; w0 += 8 * blockDim.x) {
    C v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int w = w0 + u * blockDim.x;
      const int cc = dh.q(w);
      // a select, not a conditional store: the latter put v on the stack in FP64
      v[u] = w < Kc * Ha ? sp[((size_t)cc * Kb + bb) * Ha + (w - cc * Ha)] : Cx<T>::make((T)0, (T)0);
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int w = w0 + u * blockDim.x;
      if (w < Kc * Ha) sm[w] = v[u];
    }
  }
  __syncthreads();
  const bool c64 = sizeof(T) == 4 && NT == 256 && plan_88(a, 0) && plan_88(a, 2);  // Ha = 33, Kc = 64
  if (c64) lds_fft64<T, NT, 33, 33, 1>(sm, -1, twc);
  else lds_fft<T, NT, MAXP>(sm, Kc, Ha, Ha, 1, -1, twc, a.npass[2], a.radix[2]);
  const double *hi = a.cell.hi;
  const double tpi = 2.0 * M_PI, falp2 = a.falp2;
  const int kb = bb <= Kb / 2 ? bb : bb - Kb;
  double e = 0.0, v11 = 0.0, v22 = 0.0, v33 = 0.0, v12 = 0.0, v13 = 0.0, v23 = 0.0;
  for (int w = threadIdx.x; w < Kc * Ha; w += blockDim.x) {
    const int cc = dh.q(w), aa = w - cc * Ha;
    const double W = a.infl[((size_t)cc * Kb + bb) * Ha + aa];
    const C s = sm[w];
    if (W != 0.0) {
      const int kc = cc <= Kc / 2 ? cc : cc - Kc;
      const double ga = tpi * (double)aa, gb = tpi * (double)kb, gc = tpi * (double)kc;
      const double xk = (ga * hi[0] + gb * hi[1] + gc * hi[2]);
      const double yk = (ga * hi[3] + gb * hi[4] + gc * hi[5]);
      const double zk = (ga * hi[6] + gb * hi[7] + gc * hi[8]);
      const double g2 = xk * xk + yk * yk + zk * zk;
      const double wsum = (aa == 0) ? 0.5 : 1.0;  // a' = 0 holds both +k and -k
      const double sx = (double)s.x, sy = (double)s.y;
      const double sm2 = (sx * sx + sy * sy) * wsum;
      e += W * sm2;
      const double prep = -2.0 * W * ((g2 / falp2) + 1.0) / g2 * sm2;
      v11 += prep * xk * xk;
      v22 += prep * yk * yk;
      v33 += prep * zk * zk;
      v12 += prep * xk * yk;
      v13 += prep * xk * zk;
      v23 += prep * yk * zk;
    }
    sm[w] = Cx<T>::make((T)(s.x * W), (T)(s.y * W));
  }
  {
    // the k-space pressure also carries the energy on its diagonal
    // (ewald3d_pme.c:412-414); reduced before the inverse transform so the
    // sums are not held in registers through it
    const double vals[FFT_NV] = {e, v11 + e, v22 + e, v33 + e, v12, v13, v23};
    block_partials<FFT_NV>(vals, 1, b, a.P, a.part, bb, a.K[1]);
  }
  __syncthreads();
  if (c64) lds_fft64<T, NT, 33, 33, 1>(sm, +1, twc);
  else lds_fft<T, NT, MAXP>(sm, Kc, Ha, Ha, 1, +1, twc, a.npass[2], a.radix[2]);
  for (int w = threadIdx.x; w < Kc * Ha; w += blockDim.x) {
    const int cc = dh.q(w), aa = w - cc * Ha;
    sp[((size_t)cc * Kb + bb) * Ha + aa] = sm[w];
  }
}

// inverse transform of one c-plane: along b, then complex-to-real along a.
template <typename T, int NT, int MAXP>
__global__ void __launch_bounds__(NT) k_fft_inv_ba(FftArgs a) {
  using C = typename Cx<T>::C;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  C *sm = (C *)lds_raw;
  int c, b;
  if (!xcd_unit_bead(a.K[2], a.P, c, b)) return;
  const int Ka = a.K[0], Kb = a.K[1], Ha = Ka / 2 + 1;
  C *twa = sm + Ka * Kb, *twb = twa + Ka;  // twiddles after the plane
  stage_twiddles<T>(a, 0, twa);
  stage_twiddles<T>(a, 1, twb);
  const C *sp = (const C *)a.spec + (size_t)b * a.sstride + (size_t)c * Kb * Ha;
  const FastDiv dka(Ka);
  for (int w0 = threadIdx.x; w0 < Kb * Ka; w0 += 8 * blockDim.x) {
    C v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int w = w0 + u * blockDim.x;
      const int bb = dka.q(w), aa = w - bb * Ka;
      v[u] = (w < Kb * Ka && aa < Ha) ? sp[bb * Ha + aa] : Cx<T>::make((T)0, (T)0);
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int w = w0 + u * blockDim.x;
      if (w < Kb * Ka) sm[w] = v[u];
    }
  }
  __syncthreads();
  const bool ab64 = sizeof(T) == 4 && NT == 256 && plan_88(a, 0) && plan_88(a, 1);  // 64 x 64 planes
  if (ab64) lds_fft64<T, NT, 33, 64, 1>(sm, +1, twb);
  else lds_fft<T, NT, MAXP>(sm, Kb, Ha, Ka, 1, +1, twb, a.npass[1], a.radix[1]);
  // Hermitian completion along a: X[Ka - a'] = conj(X[a'])
  for (int w = threadIdx.x; w < Kb * Ka; w += blockDim.x) {
    const int bb = dka.q(w), aa = w - bb * Ka;
    if (aa >= Ha) {
      const C v = sm[bb * Ka + (Ka - aa)];
      sm[w] = Cx<T>::make(v.x, -v.y);
    }
  }
  __syncthreads();
  if (ab64) lds_fft64<T, NT, 64, 1, 64>(sm, +1, twa);
  else lds_fft<T, NT, MAXP>(sm, Ka, Kb, 1, Ka, +1, twa, a.npass[0], a.radix[0]);
  T *g = (T *)a.grid + (size_t)b * a.gstride + (size_t)c * Ka * Kb;
  for (int w = threadIdx.x; w < Ka * Kb; w += blockDim.x) g[w] = sm[w].x;
}

template <typename T, int NT, int MAXP>
hipError_t launch_fft_class(const FftArgs &f, hipStream_t st) {
  using C = typename Cx<T>::C;
  const int Ka = f.K[0], Kb = f.K[1], Kc = f.K[2], Ha = Ka / 2 + 1;
  // the plane (or the lines) and the twiddles of the axes it transforms
  const size_t lab = ((size_t)Ka * Kb + Ka + Kb) * sizeof(C), lch = ((size_t)Kc * Ha + Kc) * sizeof(C);
  if (lab > FFT_LDS_DYN || lch > FFT_LDS_DYN) return hipErrorInvalidValue;  // refused at setup already
  if (lab > 65536 || lch > 65536) {  // beyond 64 KB the kernels ask for the larger LDS
    hipError_t e;
    if ((e = hipFuncSetAttribute((const void *)k_fft_fwd_ab<T, NT, MAXP>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lab)) != hipSuccess)
      return e;
    if ((e = hipFuncSetAttribute((const void *)k_fft_c_conv<T, NT, MAXP>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lch)) != hipSuccess)
      return e;
    if ((e = hipFuncSetAttribute((const void *)k_fft_inv_ba<T, NT, MAXP>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lab)) != hipSuccess)
      return e;
  }
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fft_fwd_ab<T, NT, MAXP>), dim3(xcd_grid(Kc, f.P)), dim3(NT), lab, st, f);
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fft_c_conv<T, NT, MAXP>), dim3(xcd_grid(Kb, f.P)), dim3(NT), lch, st, f);
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fft_inv_ba<T, NT, MAXP>), dim3(xcd_grid(Kc, f.P)), dim3(NT), lab, st, f);
  return hipGetLastError();
}

// Two plane classes: planes of up to FFT_MAX_PLANE points with 256 lanes,
// larger ones (up to FFT_LDS_MAX bytes, the LDS a workgroup may hold on
// MI355X) with 1024 lanes, so the registers that stage a Stockham pass stay
// few in both.
template <typename T>
hipError_t launch_fft_t(const FftArgs &f, hipStream_t st) {
  using C = typename Cx<T>::C;
  const int Ka = f.K[0], Kb = f.K[1], Kc = f.K[2], Ha = Ka / 2 + 1;
  if ((size_t)Ka * Kb <= FFT_MAX_PLANE && (size_t)Kc * Ha <= FFT_MAX_PLANE)
    return launch_fft_class<T, 256, FFT_MAX_PLANE>(f, st);
  return launch_fft_class<T, 1024, (int)(FFT_LDS_MAX / sizeof(C))>(f, st);
}

// forward 3-D FFT, convolution with energy and virial, inverse 3-D FFT
hipError_t launch_pme_fft(const FftArgs &f, hipStream_t st) {
  return f.f32 ? launch_fft_t<float>(f, st) : launch_fft_t<double>(f, st);
}

}  // namespace pmx
