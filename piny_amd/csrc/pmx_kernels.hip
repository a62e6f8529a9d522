// HIP kernels of the force evaluation, written for gfx950 (wave64).
//
// Device layout: positions and forces are atom-major with the bead index
// fastest, x[i*P + b], so that the lanes of a wavefront that handle the P
// beads of one pair read and write contiguous memory.  Reference citations
// name the PINY routine whose arithmetic each kernel restates.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "pmx_device.h"

namespace pmx {

// ---------------------------------------------------------------------------
// helpers

__global__ void k_fill_int(int *p, int v, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

// a fill kernel instead of hipMemsetAsync, which must not be recorded into
// graphs
__global__ void k_fill_double(double *p, double v, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

// The per-bead sums from the partial rows: one workgroup per (bead, slot)
// output (and one more for the scalar *sacc = the sum of spart[0 .. sG)),
// its 256 lanes over the rows' entries, then a fixed-order tree in LDS, so
// the result does not depend on timing.
__global__ void __launch_bounds__(256) k_reduce_partials(RedArgs a) {
  __shared__ double red[256];
  const int o = blockIdx.x, tid = threadIdx.x;
  const bool scalar = o == a.P * a.nslot;
  if (scalar && !a.sacc) return;  // uniform over the workgroup
  double s0 = 0.0, s1 = 0.0;
  if (scalar) {
    for (int g = tid; g < a.sG; g += 256) s0 += a.spart[g];
  } else {
    const int b = o / a.nslot, slot = o % a.nslot;
    for (int si = 0; si < a.nstage; si++) {
      const RedStage &r = a.st[si];
      for (int v = 0; v < r.nv; v++) {
        if (r.slot[v] != slot) continue;
        const double *row = a.part + r.off + ((size_t)v * a.P + b) * r.G;
        // a weight of 1 leaves every sum as it is (the products are exact)
        const double w = v == 0 ? 1.0 : r.wv;
        int g = tid;
        for (; g + 256 < r.G; g += 512) {
          s0 += w * row[g];
          s1 += w * row[g + 256];
        }
        for (; g < r.G; g += 256) s0 += w * row[g];
      }
    }
  }
  red[tid] = s0 + s1;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) red[tid] += red[tid + h];
    __syncthreads();
  }
  if (tid == 0) {
    if (scalar) *a.sacc = red[0];
    else a.acc[(size_t)(o / a.nslot) * a.nslot + (o % a.nslot)] = red[0];
  }
}

// ---------------------------------------------------------------------------
// layout conversion

__global__ void k_bead_to_atom(int N, int P, const double *__restrict__ src, double *__restrict__ dst) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)N * P) return;
  const int i = (int)(t / P), b = (int)(t % P);
  dst[t] = src[(size_t)b * N + i];
}

__global__ void k_atom_to_bead(int N, int P, const double *__restrict__ src, double *__restrict__ dst) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)N * P) return;
  const int b = (int)(t / N), i = (int)(t % N);
  dst[t] = src[(size_t)i * P + b];
}

// get_ghost_pos, ghost_control.c:28-64 (same summation order, no FMA)
__global__ void k_ghost_pos(int ng, int P, const int *__restrict__ gatom, const int *__restrict__ gncomp,
                            const int *__restrict__ gcomp, const double *__restrict__ gcoef, double *x,
                            double *y, double *z) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ng * P) return;
  const int g = t / P, b = t % P;
  const size_t ig = (size_t)gatom[g] * P + b;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int k = 0; k < gncomp[g]; k++) {
    const size_t ia = (size_t)gcomp[4 * g + k] * P + b;
    const double c = gcoef[4 * g + k];
    sx += c * x[ia];
    sy += c * y[ia];
    sz += c * z[ia];
  }
  x[ig] = sx;
  y[ig] = sy;
  z[ig] = sz;
}

// (xmod, q) of every site in one 32-byte record for the force kernels
__global__ void k_pack_sites(int N, int pimd, const double *__restrict__ xm, const double *__restrict__ ym,
                             const double *__restrict__ zm, const double *__restrict__ q, double4 *site) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  site[i] = pimd ? make_double4(xm[i], ym[i], zm[i], q[i]) : make_double4(0.0, 0.0, 0.0, q[i]);
}

// Bead displacements from the centroid reference xmod in FP32, for the
// mixed-precision pair kernel: the separation of a pair is then
// (u_i - u_j) + (xmod_i - xmod_j imaged), the same quantity the FP64 kernel
// forms as x_i - x_j plus the xmod image shift.
__global__ void k_pack_disp(int N, int P, const double *__restrict__ x, const double *__restrict__ y,
                            const double *__restrict__ z, const double *__restrict__ xm,
                            const double *__restrict__ ym, const double *__restrict__ zm, float4 *__restrict__ u) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)N * P) return;
  const int i = (int)(t / P);
  u[t] = make_float4((float)(x[t] - xm[i]), (float)(y[t] - ym[i]), (float)(z[t] - zm[i]), 0.0f);
}

// k_pack_disp and k_site_umax in one pass for P <= 64: a group of P lanes
// holds one site's beads (64 / P sites per wavefront); each lane stores its
// bead's displacement and the group's largest |u|^2 meets by shuffles.
__global__ void __launch_bounds__(256) k_pack_disp_umax(int N, int P, const double *__restrict__ x,
                                                        const double *__restrict__ y, const double *__restrict__ z,
                                                        const double *__restrict__ xm, const double *__restrict__ ym,
                                                        const double *__restrict__ zm, float4 *__restrict__ u,
                                                        double *__restrict__ umax) {
  const int lane = threadIdx.x & 63;
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int S = 64 / P, g = lane / P, k = lane - g * P, lead = g * P;
  const int i = wave * S + g;
  const bool act = g < S && i < N;
  float m = 0.0f;
  if (act) {
    const size_t t = (size_t)i * P + k;
    const float4 v = make_float4((float)(x[t] - xm[i]), (float)(y[t] - ym[i]), (float)(z[t] - zm[i]), 0.0f);
    u[t] = v;
    m = v.x * v.x + v.y * v.y + v.z * v.z;
  }
  float mx = m;
  if ((P & (P - 1)) == 0) {  // groups of 2^k lanes: butterfly
    for (int off = 1; off < P; off <<= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  } else {
    for (int o = 1; o < P; o++) mx = fmaxf(mx, __shfl(m, lead + (k + o) % P, 64));
  }
  if (act && k == 0) umax[i] = sqrt((double)mx);
}

// The largest |u| of every site over its beads (one thread per site).
__global__ void k_site_umax(int N, int P, const float4 *__restrict__ u, double *__restrict__ umax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  float m = 0.0f;
  for (int b = 0; b < P; b++) {
    const float4 v = u[(size_t)i * P + b];
    m = fmaxf(m, v.x * v.x + v.y * v.y + v.z * v.z);
  }
  umax[i] = sqrt((double)m);
}

// Bead-1 copy, largest displacement since the last list build
// (check_ver_list, nbr_list_control.c:182-239), bead spread
// (get_pimd_spread, pimd_estimators.c:121-222) and, when asked, xmod, the
// mode-1 coordinate of the bead transform (control_pimd_trans.c:175-181
// takes xmod from mode 1): the bead centroid for normal modes (centroid = 1)
// or bead 1 for staging (centroid = 2).  Each workgroup writes its maxima
// of the squares; a group of P lanes holds one site's beads (64 / P sites
// per wavefront; with P > 64 one site per wavefront in chunks of 64 beads),
// so the loads are coalesced; the group's lead lane forms the centroid sum
// over beads 0, 1, ... in order from the wavefront's LDS rows.
__device__ __forceinline__ double shfl_d(double v, int src) { return __shfl(v, src, 64); }

__global__ void __launch_bounds__(256) k_list_check(int N, int P, const double *__restrict__ x,
                                                    const double *__restrict__ y, const double *__restrict__ z,
                                                    const double *__restrict__ x0, const double *__restrict__ y0,
                                                    const double *__restrict__ z0, double *b0x, double *b0y,
                                                    double *b0z, double *maxes, int centroid,
                                                    double *xm, double *ym, double *zm, CheckPack pk) {
#pragma clang fp contract(off)
  __shared__ double red[2][4];
  __shared__ double stage[4][3 * 64];  // one chunk of bead positions per wavefront
  const int lane = threadIdx.x & 63;
  const int G = P <= 64 ? P : 64;           // lanes per site
  const int S = 64 / G;                     // sites per wavefront
  const int g = lane / G, k = lane - g * G;  // site group, bead within the chunk
  const int lead = g * G;                   // lane of the site's bead 0 (of each chunk)
  const int nunit = (N + S - 1) / S;
  double dmax = 0.0, smax = 0.0;
  // grid-stride over wavefront units: a bounded number of workgroups, each
  // writing one pair of maxima that the host reduces
  for (int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6); wave < nunit;
       wave += (int)((gridDim.x * blockDim.x) >> 6)) {
    const int i = wave * S + g;
    const bool act = g < S && i < N;
    const size_t base = (size_t)(act ? i : 0) * P;
    double sx = 0.0, sy = 0.0, sz = 0.0, xi = 0.0, yi = 0.0, zi = 0.0;
    double xk = 0.0, yk = 0.0, zk = 0.0;  // this lane's bead of the first chunk
    for (int c0 = 0; c0 < P; c0 += G) {
      const bool on = act && c0 + k < P;
      const double xb = on ? x[base + c0 + k] : 0.0, yb = on ? y[base + c0 + k] : 0.0,
                   zb = on ? z[base + c0 + k] : 0.0;
      if (c0 == 0) {
        xi = shfl_d(xb, lead);
        yi = shfl_d(yb, lead);
        zi = shfl_d(zb, lead);
        xk = xb;
        yk = yb;
        zk = zb;
      }
      if (on) {
        const double ex = xb - xi, ey = yb - yi, ez = zb - zi;
        smax = fmax(smax, ex * ex + ey * ey + ez * ez);
      }
      // the group's lead lane forms the ordered sum over the chunk from LDS
      // (the wavefront's own rows; lanes of one wavefront need no barrier
      // beyond the ordering of their LDS accesses)
      const int n = min(G, P - c0);
      double *sw = stage[threadIdx.x >> 6];
      sw[lane] = xb;
      sw[64 + lane] = yb;
      sw[128 + lane] = zb;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (k == 0) {
#pragma unroll 8
        for (int m = 0; m < n; m++) {
          const double vx = sw[lead + m], vy = sw[64 + lead + m], vz = sw[128 + lead + m];
          if (c0 + m == 0) {
            sx = vx;
            sy = vy;
            sz = vz;
          } else {
            sx += vx;
            sy += vy;
            sz += vz;
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");  // reads done before the next chunk's stores
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (act && k == 0) {
      b0x[i] = xi;
      b0y[i] = yi;
      b0z[i] = zi;
      const double dx = xi - x0[i], dy = yi - y0[i], dz = zi - z0[i];
      dmax = fmax(dmax, dx * dx + dy * dy + dz * dz);
      if (centroid == 1) {
        xm[i] = sx / P;
        ym[i] = sy / P;
        zm[i] = sz / P;
      } else if (centroid == 2) {
        xm[i] = xi;
        ym[i] = yi;
        zm[i] = zi;
      }
    }
    if (pk.disp) {  // uniform; P divides 64, so the site's beads are this group's lanes
      // the site's xmod as the lead lane has it (formed above, or given)
      double rx = 0.0, ry = 0.0, rz = 0.0;
      if (act && k == 0) {
        rx = centroid == 1 ? sx / P : centroid == 2 ? xi : xm[i];
        ry = centroid == 1 ? sy / P : centroid == 2 ? yi : ym[i];
        rz = centroid == 1 ? sz / P : centroid == 2 ? zi : zm[i];
      }
      rx = shfl_d(rx, lead);
      ry = shfl_d(ry, lead);
      rz = shfl_d(rz, lead);
      // k_pack_disp_umax and k_pack_sites
      float m = 0.0f;
      if (act) {
        const float4 v = make_float4((float)(xk - rx), (float)(yk - ry), (float)(zk - rz), 0.0f);
        pk.disp[base + k] = v;
        m = v.x * v.x + v.y * v.y + v.z * v.z;
      }
      for (int off = 1; off < G; off <<= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
      if (act && k == 0) {
        pk.umax[i] = sqrt((double)m);
        pk.site[i] = make_double4(rx, ry, rz, pk.q[i]);
      }
    }
  }
  for (int off = 32; off >= 1; off >>= 1) {
    dmax = fmax(dmax, __shfl_xor(dmax, off, 64));
    smax = fmax(smax, __shfl_xor(smax, off, 64));
  }
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0) {
    red[0][w] = dmax;
    red[1][w] = smax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {  // the host takes the maxima over the workgroups
    for (int v = 1; v < nw; v++) {
      dmax = fmax(dmax, red[0][v]);
      smax = fmax(smax, red[1][v]);
    }
    maxes[blockIdx.x] = dmax;
    maxes[gridDim.x + blockIdx.x] = smax;
  }
}

// ---------------------------------------------------------------------------
// Verlet list: nolst_ver_gen + verlist_create (verlist_control.c:134-441,
// verlist_create.c:223-503).  One wavefront per site i scans j in ascending
// chunks of 64; a ballot keeps the reference's order, so each per-site
// segment comes out sorted.  Two lists are written in ELL form (row i at
// i*cap): the reference's half list (j < i, r^2 < cutskin^2) and the list the
// force kernel walks, which keeps only pairs that can interact (LJ table
// non-null or q_i q_j != 0) and, when full is set, holds both directions
// (every j != i).  The minimum image is odd in the separation, so the full
// list is exactly the half list and its transpose.  No FMA contraction: the
// r^2 test must see the reference's bits.
__device__ __forceinline__ double min_image_sep(const Cell &c, int pimd, double xi, double yi, double zi,
                                                double xj, double yj, double zj, double xmi, double ymi,
                                                double zmi, double xmj, double ymj, double zmj, double &dx,
                                                double &dy, double &dz) {
#pragma clang fp contract(off)
  dx = xi - xj;
  dy = yi - yj;
  dz = zi - zj;
  if (!pimd) {
    if (c.cubic) {
      const double asx = dx * c.hi[0], asy = dy * c.hi[4], asz = dz * c.hi[8];
      const double sx = asx - nint_d(asx), sy = asy - nint_d(asy), sz = asz - nint_d(asz);
      dx = sx * c.h[0];
      dy = sy * c.h[4];
      dz = sz * c.h[8];
    } else {
      const double asx = dx * c.hi[0] + dy * c.hi[3] + dz * c.hi[6];
      const double asy = dx * c.hi[1] + dy * c.hi[4] + dz * c.hi[7];
      const double asz = dx * c.hi[2] + dy * c.hi[5] + dz * c.hi[8];
      const double sx = asx - nint_d(asx), sy = asy - nint_d(asy), sz = asz - nint_d(asz);
      dx = sx * c.h[0] + sy * c.h[3] + sz * c.h[6];
      dy = sx * c.h[1] + sy * c.h[4] + sz * c.h[7];
      dz = sx * c.h[2] + sy * c.h[5] + sz * c.h[8];
    }
  } else {
    const double dxs = xmi - xmj, dys = ymi - ymj, dzs = zmi - zmj;
    if (c.cubic) {
      const double sx = -nint_d(dxs * c.hi[0]), sy = -nint_d(dys * c.hi[4]), sz = -nint_d(dzs * c.hi[8]);
      dx += (sx * c.h[0]);
      dy += (sy * c.h[4]);
      dz += (sz * c.h[8]);
    } else {
      const double asx = dxs * c.hi[0] + dys * c.hi[3] + dzs * c.hi[6];
      const double asy = dxs * c.hi[1] + dys * c.hi[4] + dzs * c.hi[7];
      const double asz = dxs * c.hi[2] + dys * c.hi[5] + dzs * c.hi[8];
      const double sx = -nint_d(asx), sy = -nint_d(asy), sz = -nint_d(asz);
      dx += (sx * c.h[0] + sy * c.h[3] + sz * c.h[6]);
      dy += (sx * c.h[1] + sy * c.h[4] + sz * c.h[7]);
      dz += (sx * c.h[2] + sy * c.h[5] + sz * c.h[8]);
    }
  }
  return dx * dx + dy * dy + dz * dz;
}

__global__ void __launch_bounds__(256) k_list_build(ListArgs a) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (i >= a.N) return;
  const double xi = a.bx[i], yi = a.by[i], zi = a.bz[i];
  double xmi = 0.0, ymi = 0.0, zmi = 0.0;
  if (a.pimd) {
    xmi = a.xm[i];
    ymi = a.ym[i];
    zmi = a.zm[i];
  }
  const int ti = a.type[i];
  const double qi = a.q[i];
  const int eo = a.ex_off[i], en = a.ex_num[i];
  int nfull = 0, nforce = 0;
  int *rowf = a.jfull + (size_t)i * a.cap;
  int *rowi = a.jforce + (size_t)i * a.cap;
  // candidates: every j (ascending, so rows come out sorted) or the sites of
  // the 27 link cells around i's cell (rows sorted afterwards by k_sort_rows)
  const int jend = a.full ? a.N : i;
  const int ncand_cells = a.use_cells ? 27 : 1;
  int ci[3] = {0, 0, 0};
  if (a.use_cells) {
    const int c0 = a.site_cell[i];
    ci[0] = c0 % a.nc[0];
    ci[1] = (c0 / a.nc[0]) % a.nc[1];
    ci[2] = c0 / (a.nc[0] * a.nc[1]);
  }
  for (int cc = 0; cc < ncand_cells; cc++) {
    int base = 0, count = jend;
    const int *sites = nullptr;
    if (a.use_cells) {
      int n3[3];
      const int off[3] = {cc % 3 - 1, (cc / 3) % 3 - 1, cc / 9 - 1};
#pragma unroll
      for (int d = 0; d < 3; d++) n3[d] = (ci[d] + off[d] + a.nc[d]) % a.nc[d];
      const int cell = (n3[2] * a.nc[1] + n3[1]) * a.nc[0] + n3[0];
      base = a.cell_start[cell];
      count = a.cell_cnt[cell];
      sites = a.cell_sites;
    }
    for (int j0 = 0; j0 < count; j0 += 64) {
      const int jj = j0 + lane;
      const int j = (jj < count) ? (sites ? sites[base + jj] : jj) : a.N;
      bool keep = false, inter = false;
      int t = 0;
      if (j < a.N && (a.full || j < i) && j != i) {
        bool excl = false;
        for (int k = 0; k < en; k++) excl |= (a.ex_j[eo + k] == j);
        if (!excl) {
          double dx, dy, dz;
          double xmj = 0.0, ymj = 0.0, zmj = 0.0;
          if (a.pimd) {
            xmj = a.xm[j];
            ymj = a.ym[j];
            zmj = a.zm[j];
          }
          const double r2 = min_image_sep(a.cell, a.pimd, xi, yi, zi, a.bx[j], a.by[j], a.bz[j], xmi, ymi, zmi, xmj,
                                          ymj, zmj, dx, dy, dz);
          t = pair_index(a.pt.ntype, ti, a.type[j]);
          const bool in = r2 < a.pt.cutskin2[t];
          keep = in && j < i;
          inter = in && (a.pt.lj[t] || (qi * a.q[j] != 0.0));
        }
      }
      if (a.ref_only) inter = false;
      const unsigned long long mk = __ballot(keep), mi = __ballot(inter);
      const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
      if (keep) {
        const int p = nfull + __popcll(mk & below);
        if (p < a.cap) rowf[p] = j;
      }
      if (inter) {
        const int p = nforce + __popcll(mi & below);
        if (p < a.cap) rowi[p] = j | (t << ENTRY_SHIFT);
      }
      nfull += __popcll(mk);
      nforce += __popcll(mi);
    }
  }
  if (lane == 0) {
    a.nfull[i] = nfull;
    if (!a.ref_only) a.nforce[i] = nforce;
    const int mx = nfull > nforce ? nfull : nforce;
    if (mx > a.cap) atomicMax(a.overflow, mx);
    atomicAdd(a.overflow + 1, nfull);  // pair count of the reference list
  }
}

// Link cells (make_lnk_lst.c:37-355 bins atoms into cells of fractional
// space): the bead-1 position's cell, a counting sort of the sites by cell.
__global__ void __launch_bounds__(256) k_cell_count(ListArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  const Cell &c = a.cell;
  const double x = a.bx[i], y = a.by[i], z = a.bz[i];
  int idx[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double f = x * c.hi[d] + y * c.hi[d + 3] + z * c.hi[d + 6];
    f -= floor(f);
    int k = (int)(f * (double)a.nc[d]);
    idx[d] = k < 0 ? 0 : (k >= a.nc[d] ? a.nc[d] - 1 : k);
  }
  const int cell = (idx[2] * a.nc[1] + idx[1]) * a.nc[0] + idx[0];
  a.site_cell[i] = cell;
  a.site_pos[i] = atomicAdd(a.cell_cnt + cell, 1);
}

// exclusive scan of the cell counts (one workgroup)
__global__ void __launch_bounds__(256) k_cell_scan(int ncell, const int *cnt, int *start) {
  __shared__ int sums[256];
  const int tid = threadIdx.x;
  const int per = (ncell + 255) / 256;
  const int lo = min(ncell, tid * per), hi = min(ncell, lo + per);
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
  int run = sums[tid];
  for (int t = lo; t < hi; t++) {
    start[t] = run;
    run += cnt[t];
  }
}

__global__ void __launch_bounds__(256) k_cell_scatter(ListArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  a.cell_sites[a.cell_start[a.site_cell[i]] + a.site_pos[i]] = i;
}

// Puts each cell's sites in ascending index order (one workgroup per cell;
// the counting sort's atomics leave them in arbitrary order), so that the
// cell order, and the force-list order that follows it, is the same on
// every build; then writes the sites' records in that order.  A cell of
// more than CELL_SORT_MAX sites sets overflow[2] and the host builds with
// the plain scan instead.
__global__ void __launch_bounds__(256) k_cell_order(ListArgs a) {
  __shared__ int ids[CELL_SORT_MAX];
  const int c = blockIdx.x, n = a.cell_cnt[c], o = a.cell_start[c];
  if (n > CELL_SORT_MAX) {  // uniform over the workgroup
    if (threadIdx.x == 0) a.overflow[2] = 1;
    return;
  }
  for (int k = threadIdx.x; k < n; k += blockDim.x) ids[k] = a.cell_sites[o + k];
  __syncthreads();
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const int v = ids[k];
    int rank = 0;
    for (int m = 0; m < n; m++) rank += ids[m] < v;
    const int d = o + rank;
    a.cell_sites[d] = v;
    a.cell_i[d] = v;
    a.cell_p[d] = make_double4(a.bx[v], a.by[v], a.bz[v], a.q[v]);
    a.cell_m[d] = a.pimd ? make_double4(a.xm[v], a.ym[v], a.zm[v], (double)a.type[v])
                         : make_double4(0.0, 0.0, 0.0, (double)a.type[v]);
  }
}

// Verlet list by link cells (make_lnk_lst.c:37-355, lnk_ver_gen_pad in
// verlist_control.c:1738-2254): cells of fractional space at least
// r_list / H wide, at least 2H + 1 per axis, so that the (2H + 1)^3 cells
// around a site's cell hold every site within the list radius, each once.
// One workgroup per cell.  For each row of 2H + 1 cells along a of the
// stencil (contiguous in cell order, up to the wrap) the workgroup stages
// the sites' records in LDS (CELL_TILE at a time); each wavefront takes
// the cell's own sites in turn and tests 64 staged candidates per step:
// the reference's image separation and cut+skin test (min_image_sep, the
// same bits as the plain scan), then, for pairs inside, the exclusion
// list.  A ballot appends the interacting pairs to the site's force-list
// row (every j != i, or j < i for the half list), in stencil order; the
// reference's half-list pairs (j < i, every type pair) are only counted
// (nver).  Rows of the cell's sites are counted in LDS across the tiles.
template <int H, int NW>
__global__ void __launch_bounds__(64 * NW) k_list_cells(ListArgs a) {
#pragma clang fp contract(off)
  __shared__ double4 tp[CELL_TILE], tm[CELL_TILE];
  __shared__ int ti[CELL_TILE];
  __shared__ int rowf[CELL_ROWS_MAX], rowr[CELL_ROWS_MAX];
  __shared__ int red[NW];
  constexpr int W = 2 * H + 1;
  const int c = blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int ni = a.cell_cnt[c], oi = a.cell_start[c];
  if (ni > CELL_ROWS_MAX) {  // uniform over the workgroup
    if (threadIdx.x == 0) a.overflow[2] = 1;
    return;
  }
  for (int k = threadIdx.x; k < ni; k += blockDim.x) {
    rowf[k] = 0;
    rowr[k] = 0;
  }
  const int ca = c % a.nc[0], cb = (c / a.nc[0]) % a.nc[1], cc = c / (a.nc[0] * a.nc[1]);
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int st = 0; st < W * W; st++) {
    const int nb = (cb + st % W - H + a.nc[1]) % a.nc[1], nz = (cc + st / W - H + a.nc[2]) % a.nc[2];
    // the row's cells: their starts and the running count before each
    int rs[W], pre[W + 1];
    pre[0] = 0;
#pragma unroll
    for (int d = 0; d < W; d++) {
      const int na = (ca + d - H + a.nc[0]) % a.nc[0];
      const int cell = (nz * a.nc[1] + nb) * a.nc[0] + na;
      rs[d] = a.cell_start[cell];
      pre[d + 1] = pre[d] + a.cell_cnt[cell];
    }
    const int tot = pre[W];
    for (int t0 = 0; t0 < tot; t0 += CELL_TILE) {
      const int nt = min(CELL_TILE, tot - t0);
      __syncthreads();  // the previous tile's readers are done (and the row counters are cleared)
      for (int e = threadIdx.x; e < nt; e += blockDim.x) {
        const int g = t0 + e;
        int src = 0;
#pragma unroll
        for (int d = 0; d < W; d++)
          if (g >= pre[d] && g < pre[d + 1]) src = rs[d] + (g - pre[d]);
        tp[e] = a.cell_p[src];
        tm[e] = a.cell_m[src];
        ti[e] = a.cell_i[src];
      }
      __syncthreads();
      for (int s = w; s < ni; s += NW) {
        const double4 pi = a.cell_p[oi + s], mi = a.cell_m[oi + s];
        const int i = a.cell_i[oi + s], typi = (int)mi.w;
        const int eo = a.ex_off[i], en = a.ex_num[i];
        int nf = rowf[s], nr = rowr[s];
        int *row = a.jforce + (size_t)i * a.cap;
        for (int b0 = 0; b0 < nt; b0 += 64) {
          const int e = b0 + lane;
          bool keep = false, inter = false;
          int j = 0, t = 0;
          if (e < nt) {
            j = ti[e];
            if (j != i) {
              const double4 pj = tp[e], mj = tm[e];
              double dx, dy, dz;
              const double r2 =
                  min_image_sep(a.cell, a.pimd, pi.x, pi.y, pi.z, pj.x, pj.y, pj.z, mi.x, mi.y, mi.z, mj.x, mj.y, mj.z,
                                dx, dy, dz);
              t = pair_index(a.pt.ntype, typi, (int)mj.w);
              if (r2 < a.pt.cutskin2[t]) {
                bool excl = false;
                for (int k = 0; k < en; k++) excl |= (a.ex_j[eo + k] == j);
                if (!excl) {
                  keep = j < i;
                  inter = (a.full || j < i) && (a.pt.lj[t] || (pi.w * pj.w != 0.0));
                }
              }
            }
          }
          const uint64_t mi2 = __ballot(inter);
          if (inter) {
            const int p = nf + __popcll(mi2 & below);
            if (p < a.cap) row[p] = j | (t << ENTRY_SHIFT);
          }
          nf += __popcll(mi2);
          nr += __popcll(__ballot(keep));
        }
        if (lane == 0) {
          rowf[s] = nf;
          rowr[s] = nr;
        }
      }
    }
  }
  __syncthreads();
  int sum = 0;
  for (int s = threadIdx.x; s < ni; s += blockDim.x) {
    const int i = a.cell_i[oi + s], nf = rowf[s];
    a.nforce[i] = nf;
    if (nf > a.cap) atomicMax(a.overflow, nf);
    sum += rowr[s];
  }
  for (int off = 32; off >= 1; off >>= 1) sum += __shfl_xor(sum, off, 64);
  if (lane == 0) red[w] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int k = 0; k < NW; k++) t += red[k];
    atomicAdd(a.overflow + 1, t);
  }
}

// nw wavefronts per cell's workgroup (4, 8 or 16): more of them keep more
// loads in flight; each site's row is still written by one wavefront, in
// stencil order.
template <int NW>
void launch_list_cells_w(const ListArgs &a, int halfwidth, int ncell, hipStream_t st) {
  if (halfwidth == 2)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_list_cells<2, NW>), dim3(ncell), dim3(64 * NW), 0, st, a);
  else
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_list_cells<1, NW>), dim3(ncell), dim3(64 * NW), 0, st, a);
}

hipError_t launch_list_cells(const ListArgs &a, int halfwidth, int ncell, int nw, hipStream_t st) {
  if (nw >= 16) launch_list_cells_w<16>(a, halfwidth, ncell, st);
  else if (nw >= 8) launch_list_cells_w<8>(a, halfwidth, ncell, st);
  else launch_list_cells_w<4>(a, halfwidth, ncell, st);
  return hipGetLastError();
}

// Sorts each list row ascending by site index (one workgroup per row,
// bitonic sort in LDS over n2 >= cap keys): the link-cell build then yields
// the same rows, in the same order, as the ascending scan.
__global__ void __launch_bounds__(256) k_sort_rows(int cap, int n2, const int *nfull, const int *nforce, int *jfull,
                                                   int *jforce) {
  extern __shared__ int keys[];
  const int i = blockIdx.x, tid = threadIdx.x;
  for (int which = 0; which < 2; which++) {
    int *row = (which == 0 ? jfull : jforce) + (size_t)i * cap;
    const int cnt = which == 0 ? nfull[i] : nforce[i];
    int m = 64;  // this row's sort width: a power of two >= cnt, <= n2
    while (m < cnt && m < n2) m <<= 1;
    for (int k = tid; k < m; k += blockDim.x) keys[k] = k < cnt ? row[k] : ENTRY_MASK | 0x7f000000;
    __syncthreads();
    for (int len = 2; len <= m; len <<= 1) {
      for (int half = len >> 1; half > 0; half >>= 1) {
        for (int k = tid; k < m; k += blockDim.x) {
          const int p = k ^ half;
          if (p > k) {
            const int a = keys[k], b = keys[p];
            const bool up = (k & len) == 0;
            const bool swap = up ? ((a & ENTRY_MASK) > (b & ENTRY_MASK)) : ((a & ENTRY_MASK) < (b & ENTRY_MASK));
            if (swap) {
              keys[k] = b;
              keys[p] = a;
            }
          }
        }
        __syncthreads();
      }
    }
    for (int k = tid; k < cnt; k += blockDim.x) row[k] = keys[k];
    __syncthreads();
  }
}

// Orders the LDS writes of a wavefront before its other lanes read them
// (and the reads before the next writes).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Splits every force-list row, stably, into the entries whose pair lies
// inside its cutoff at the build positions (the centroid reference xmod with
// beads, bead 1 without), then the rest of the list shell, into out.  The
// pair kernels evaluate a row's entries in order, several per wavefront at a
// time; grouped like this, the pairs of one iteration mostly agree on the
// cutoff test and the lanes of a wavefront seldom idle.  One wavefront per
// row.
__global__ void __launch_bounds__(256) k_split_rows(ListArgs a, int *out) {
  // the inside flags of a row's chunks, kept from the counting pass for the
  // placing pass (rows of up to SPLIT_CHUNKS chunks; longer rows test again)
  __shared__ uint64_t msk[4][SPLIT_CHUNKS];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = (int)(((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (i >= a.N) return;  // uniform over the wavefront
  const int cnt = a.nforce[i];
  const int *row = a.jforce + (size_t)i * a.cap;
  int *dst = out + (size_t)i * a.cap;
  const double xi = a.pimd ? a.xm[i] : a.bx[i], yi = a.pimd ? a.ym[i] : a.by[i], zi = a.pimd ? a.zm[i] : a.bz[i];
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  auto inside = [&](int e) {
    const int j = e & ENTRY_MASK, t = e >> ENTRY_SHIFT;
    const double xj = a.pimd ? a.xm[j] : a.bx[j], yj = a.pimd ? a.ym[j] : a.by[j],
                 zj = a.pimd ? a.zm[j] : a.bz[j];
    double dx, dy, dz;
    const double r2 = min_image_sep(a.cell, 0, xi, yi, zi, xj, yj, zj, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, dx, dy, dz);
    return r2 < a.pt.cut[t] * a.pt.cut[t];
  };
  int ninner = 0;
  for (int k0 = 0, c = 0; k0 < cnt; k0 += 64, c++) {
    const int k = k0 + lane;
    const uint64_t mi = __ballot(k < cnt && inside(row[k]));
    if (c < SPLIT_CHUNKS && lane == 0) msk[w][c] = mi;
    ninner += __popcll(mi);
  }
  wave_sync();
  int pin = 0, pout = ninner;
  for (int k0 = 0, c = 0; k0 < cnt; k0 += 64, c++) {
    const int k = k0 + lane;
    const bool valid = k < cnt;
    const int e = valid ? row[k] : 0;
    const uint64_t mi = c < SPLIT_CHUNKS ? msk[w][c] : __ballot(valid && inside(e));
    const uint64_t mv = __ballot(valid);
    const bool inner = (mi >> lane) & 1ull;
    if (valid) dst[inner ? pin + __popcll(mi & below) : pout + __popcll(mv & ~mi & below)] = e;
    pin += __popcll(mi);
    pout += __popcll(mv & ~mi);
  }
}

// ---------------------------------------------------------------------------
// Real-space pair forces over the interacting list: force_npol
// (force_npol.c:28-328) with shave_skin_opt on and RESPA off.  A wavefront
// owns one site i at a time; its 64 lanes are (slot, bead) with B beads per
// slot, so the B lanes of a slot evaluate the same pair for B beads and
// touch contiguous memory.  The i force is reduced over slots with shuffles.
// With the full list (the default: k_pair_staged_tab/_dir, k_pair_mixed) every pair
// is met from both sites: each wavefront stores its own sites' forces and
// the energies and virial carry a factor 1/2, so there are no atomics.  With
// the half list (k_pair_half, a tuning switch) the j forces are scattered
// with FP64 atomics.
//
// MODE 0 interpolates PINY's spline tables in FP64 (npol_vspl_fetch); MODE 1
// evaluates the tabulated functions directly in FP64; MODE 2 does that in
// FP32 with the separation taken in FP64 and every sum kept in FP64.

__device__ __forceinline__ double spline4(const double2 *__restrict__ tab, int k, double del, bool second) {
  // npol_vspl_fetch, force_npol.c:339-376 (k is PINY's 1-based index)
  const double2 tm2 = tab[k - 3], tm1 = tab[k - 2], t0 = tab[k - 1], tp1 = tab[k];
  const double fm2 = second ? tm2.y : tm2.x, fm1 = second ? tm1.y : tm1.x;
  const double c0 = second ? t0.y : t0.x, fp1 = second ? tp1.y : tp1.x;
  const double oneth = 1.0 / 3.0, onesi = 1.0 / 6.0;
  const double c1 = oneth * fp1 + 0.5 * c0 - fm1 + onesi * fm2;
  const double c2 = 0.5 * fp1 - c0 + 0.5 * fm1;
  const double c3 = onesi * fp1 - 0.5 * c0 + 0.5 * fm1 - onesi * fm2;
  return (c0 + del * (c1 + del * (c2 + del * c3)));
}

template <typename R>
__device__ __forceinline__ R rcp_r(R x);
template <>
__device__ __forceinline__ double rcp_r<double>(double x) {
  return 1.0 / x;
}
template <>
__device__ __forceinline__ float rcp_r<float>(float x) {
  return __builtin_amdgcn_rcpf(x);  // v_rcp_f32, 1 ulp
}
template <typename R>
__device__ __forceinline__ R exp_r(R x);
template <>
__device__ __forceinline__ double exp_r<double>(double x) {
  return exp(x);
}
template <>
__device__ __forceinline__ float exp_r<float>(float x) {
  return __expf(x);
}

// energy (ev) and -(1/r) dV/dr (returned) of one pair at distance r, for
// the LJ term (lj) and the erfc term with charge product q2; pp holds the
// type pair's constants
template <typename R, typename PT = double>
__device__ __forceinline__ R pair_direct(const PairArgs &a, const PT *pp, R q2, R r, R rinv, R &vlj, R &vc) {
  R sw = (R)1, dsw = (R)0;
  if (a.swit) {
    const R rheal = (R)a.rheal, irheal = (R)a.irheal;
    const R rs = r - (R)pp[PP_CUT] + rheal;
    if (rs > (R)0) {  // inside the switching shell (below it sw = 1, dsw = 0)
      R rp = rs * irheal;
      rp = rp < (R)1 ? rp : (R)1;
      sw = (R)1 + rp * rp * ((R)2 * rp - (R)3);
      dsw = (R)(-6) * (rp * irheal) * (rp - (R)1) * rinv;
    }
  }
  const R r2i = rinv * rinv;
  constexpr bool f32 = sizeof(R) == 4;
  R dvt = (R)0;
  vlj = (R)0;
  vc = (R)0;
  if (pp[PP_LJ] != (PT)0) {
    R v, dv;
    if constexpr (f32) {  // r^-6 (C12 r^-6 - C6) and 6 r^-6 (2 C12 r^-6 - C6) / r^2
      const R r6i = r2i * r2i * r2i, c6 = (R)pp[PP_C6], t12 = (R)pp[PP_C12] * r6i;
      v = r6i * (t12 - c6);
      dv = ((R)6 * r6i) * ((R)2 * t12 - c6) * r2i;
    } else {
      const R eps = (R)pp[PP_EPS];
      const R s6 = (R)pp[PP_SIG6] * (r2i * r2i * r2i);
      v = (R)4 * eps * s6 * (s6 - (R)1);
      dv = (R)24 * eps * s6 * (s6 * (R)2 - (R)1) * r2i;
    }
    vlj = v * sw - (R)pp[PP_VSLJ];
    dvt = dv * sw + v * dsw;
  }
  if (q2 != (R)0) {
    const R ralp = r * (R)a.alp;
    const R eee = exp_r<R>(-ralp * ralp);
    const R tt = rcp_r<R>((R)1 + (R)kErfP * ralp);
    R g = (R)kErfE[9], dg;
    if constexpr (f32) {
      // -d erfc(alp r)/dr = 2 alp / sqrt(pi) exp(-alp^2 r^2); the polynomial's
      // own derivative differs from it by at most 1.1e-7 relative (5e-9 in
      // the force term), below FP32 rounding
#pragma unroll
      for (int k = 8; k >= 1; k--) g = g * tt + (R)kErfE[k];
      g = g * tt * eee;
      dg = (R)a.spalp * eee;
    } else {
      dg = (R)(9.0 * kErfE[9]);
#pragma unroll
      for (int k = 8; k >= 1; k--) {
        g = g * tt + (R)kErfE[k];
        dg = dg * tt + (R)(k * kErfE[k]);
      }
      g = g * tt * eee;
      dg = dg * tt * tt * eee * (R)a.p_alpha + (R)a.two_alpha2 * g * r;
    }
    const R v = g * rinv;
    const R dv = (g * r2i + dg * rinv) * rinv;
    vc = q2 * (v * sw - (R)pp[PP_VSC]);
    dvt += q2 * (dv * sw + v * dsw);
  }
  return dvt;
}

// Accumulators of one lane of the pair kernel.
struct PairAcc {
  double vlj, vc, p11, p22, p33, p12, p13, p23, npr, fx, fy, fz;
};

// One pair term for one bead: separation (dx, dy, dz) and squared distance
// r2d (FP64), type pair constants pp; returns the scalar -(1/r) dV/dr and
// adds the energies, the i force and the virial.
// The RESPA weight of a pair force at distance r (npol_spldata_fetch and
// npol_force_calc, force_npol.c:807-816,1029-1060): the switch
// 1 + b^2 (2b - 3), b = (r - cutoff_res + rheal) / rheal clamped to [0, 1],
// is 1 well inside the RESPA cutoff and 0 beyond it.
__device__ __forceinline__ double respa_weight(const PairArgs &a, double cres, double r) {
  if (a.respa == RESPA_FULL) return a.w_full;
  double br = (r - cres + a.rheal) * a.irheal;
  br = fmax(br, 0.0);
  br = fmin(br, 1.0);
  const double sw = br * br * (br * 2.0 - 3.0) + 1.0;
  return a.respa == RESPA_RES ? sw * a.w_res : sw * a.w_res + (1.0 - sw) * a.w_full;
}

template <int MODE>
__device__ __forceinline__ bool pair_term(const PairArgs &a, const double *pp, int t, double qq, double dx,
                                          double dy, double dz, double r2d, PairAcc &s, double &dvd) {
  using R = typename std::conditional<MODE == 2, float, double>::type;
  const double cut = pp[PP_CUT];
  if constexpr (MODE == 2) {
    if (!(r2d < cut * cut)) return false;
    const float r2 = (float)r2d;
    const float rinv = __frsqrt_rn(r2);
    R ev_lj, ev_c;
    dvd = (double)pair_direct<float>(a, pp, (float)qq, r2 * rinv, rinv, ev_lj, ev_c);
    s.vlj += (double)ev_lj;
    s.vc += (double)ev_c;
  } else {
    const double rd = sqrt(r2d);
    if (!(rd < cut)) return false;
    if constexpr (MODE == 0) {
      const double rmin = pp[PP_RMIN], dri = pp[PP_DRI];
      int kk = (int)(((rd - rmin) * dri) + 0.5) + 3;
      kk = min(kk, a.nsplin - 2);
      kk = max(kk, 3);
      const double rsp = ((double)(kk - 3)) * pp[PP_DR] + rmin;
      const double del = (rd - rsp) * dri;
      const size_t toff = (size_t)t * a.nsplin;
      double dvt = 0.0;
      if (pp[PP_LJ] != 0.0) {
        s.vlj += spline4(a.tab_lj + toff, kk, del, false);
        dvt += spline4(a.tab_lj + toff, kk, del, true);
      }
      if (qq != 0.0) {
        s.vc += qq * spline4(a.tab_c + toff, kk, del, false);
        dvt += qq * spline4(a.tab_c + toff, kk, del, true);
      }
      dvd = dvt;
    } else {
      double ev_lj, ev_c;
      dvd = pair_direct<double>(a, pp, qq, rd, 1.0 / rd, ev_lj, ev_c);
      s.vlj += ev_lj;
      s.vc += ev_c;
    }
  }
  if (a.respa) dvd *= respa_weight(a, pp[PP_CUTRES], sqrt(r2d));  // uniform over the launch
  const double fx1 = dx * dvd, fy1 = dy * dvd, fz1 = dz * dvd;
  s.fx += fx1;
  s.fy += fy1;
  s.fz += fz1;
  s.p11 += dx * fx1;
  s.p22 += dy * fy1;
  s.p33 += dz * fz1;
  s.p12 += dx * fy1;
  s.p13 += dx * fz1;
  s.p23 += dy * fz1;
  s.npr += 1.0;
  return true;
}

template <int B, int MODE>
__global__ void __launch_bounds__(256) k_pair_half(PairArgs a) {
  __shared__ double spp[MAX_PAIRS * PP_NUM];
  for (int k = threadIdx.x; k < a.npair * PP_NUM; k += blockDim.x) spp[k] = a.pp[k];
  __syncthreads();
  constexpr int S = 64 / B;
  const int lane = threadIdx.x & 63;
  const int slot = lane / B;
  const int bead = (lane % B) + blockIdx.y * B;
  const bool bact = bead < a.P;
  const int P = a.P;
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwave = (int)((gridDim.x * blockDim.x) >> 6);
  const Cell &c = a.cell;
  PairAcc s{};
  for (int i = wave; i < a.N; i += nwave) {
    const size_t ib = (size_t)i * P + (bact ? bead : 0);
    const double xi = a.x[ib], yi = a.y[ib], zi = a.z[ib];
    const double4 si = a.site[i];
    const int cnt = a.cnt[i];
    const int *row = a.jlist + (size_t)i * a.cap;
    s.fx = s.fy = s.fz = 0.0;
    // two list entries per slot and iteration; the next iteration's entries
    // are loaded while the current pairs are evaluated
    int k0 = slot;
    int en0 = k0 < cnt ? row[k0] : 0, en1 = k0 + S < cnt ? row[k0 + S] : 0;
    for (; k0 < cnt; k0 += 2 * S) {
      const int e[2] = {en0, en1};
      const bool v1 = k0 + S < cnt;
      const int kn = k0 + 2 * S;
      en0 = kn < cnt ? row[kn] : 0;
      en1 = kn + S < cnt ? row[kn + S] : 0;
      double xj[2], yj[2], zj[2];
      double4 sj[2];
#pragma unroll
      for (int u = 0; u < 2; u++) {
        const int j = e[u] & ENTRY_MASK;
        const size_t jb = (size_t)j * P + (bact ? bead : 0);
        xj[u] = a.x[jb];
        yj[u] = a.y[jb];
        zj[u] = a.z[jb];
        sj[u] = a.site[j];
      }
#pragma unroll
      for (int u = 0; u < 2; u++) {
        if (u == 1 && !v1) break;
        const int j = e[u] & ENTRY_MASK, t = e[u] >> ENTRY_SHIFT;
        double dx, dy, dz;
        const double r2d = min_image_sep(c, a.pimd, xi, yi, zi, xj[u], yj[u], zj[u], si.x, si.y, si.z, sj[u].x,
                                         sj[u].y, sj[u].z, dx, dy, dz);
        double dvd;
        if (bact && pair_term<MODE>(a, spp + t * PP_NUM, t, si.w * sj[u].w, dx, dy, dz, r2d, s, dvd)) {
          const size_t jb = (size_t)j * P + bead;
          atomic_add(a.fx + jb, -(dx * dvd));
          atomic_add(a.fy + jb, -(dy * dvd));
          atomic_add(a.fz + jb, -(dz * dvd));
        }
      }
    }
    double fxi = s.fx, fyi = s.fy, fzi = s.fz;
    for (int off = 32; off >= B; off >>= 1) {
      fxi += __shfl_xor(fxi, off, 64);
      fyi += __shfl_xor(fyi, off, 64);
      fzi += __shfl_xor(fzi, off, 64);
    }
    if (slot == 0 && bact && cnt > 0) {
      atomic_add(a.fx + ib, fxi);
      atomic_add(a.fy + ib, fyi);
      atomic_add(a.fz + ib, fzi);
    }
  }
  const double vals[PAIR_NV] = {s.vlj, s.vc, s.p11, s.p22, s.p33, s.p12, s.p13, s.p23, s.npr};
  block_partials<PAIR_NV>(vals, B, blockIdx.y * B, P, a.part, blockIdx.x, gridDim.x);
}

// The FP32 pair term of pair_direct with type pair t's constants from LDS:
// sA[t] = (cut^2, cut, C12, C6), sB[t] = (LJ shift, Coulomb shift, LJ flag,
// RESPA cutoff); lj says whether the pair has an LJ term.  ct holds the
// erfc term's cubic pieces: with CSW switched and shifted already (the
// switch and the constants are then read only for the LJ term), else
// erfc(alp r)/r itself.  The energies are added to vlj and vc; the erfc term
// is formed for every pair (q2 = 0 adds zeros), so only the LJ term takes a
// branch.
template <bool CSW>
__device__ __forceinline__ float pair_direct_f32(const PairArgs &a, const float4 *sA, const float4 *sB, int t,
                                                 bool lj, float2 q, float r, float rinv, float &vlj, float &vc,
                                                 const float4 *ct) {
  // q = (q_i q_j, -q_i q_j / dr of the table grid)
  const float q2 = q.x;
  float sw = 1.f, dsw = 0.f;
  float4 A = make_float4(0.f, 0.f, 0.f, 0.f), B = A;
  auto switch_at = [&]() {
    const float rheal = (float)a.rheal, irheal = (float)a.irheal;
    const float rs = r - A.y + rheal;
    if (rs > 0.f) {  // inside the switching shell (below it sw = 1, dsw = 0)
      float rp = rs * irheal;
      rp = rp < 1.f ? rp : 1.f;
      sw = 1.f + rp * rp * (2.f * rp - 3.f);
      dsw = -6.f * (rp * irheal) * (rp - 1.f) * rinv;
    }
  };
  if (!CSW) {
    A = sA[t];
    B = sB[t];
    if (a.swit) switch_at();
  }
  float dvt;
  {
    // the erfc term and its r derivative from the cubic piece of the table
    // u < 0 (r below the table) converts to k = 0 and extrapolates the first piece
    const float u = r * a.ct_idr - a.ct_r0 * a.ct_idr;
    // r < r_c <= the grid's end: k <= CT_N, the repeated last piece
    const unsigned k = __float2uint_rz(u);
    const float tt = u - (float)k;
    const float4 cc = ct[k];
    // value and t derivative by one Horner scheme
    const float p1 = fmaf(cc.w, tt, cc.z);
    const float d1 = fmaf(cc.w, tt, p1);
    const float p2 = fmaf(p1, tt, cc.y);
    const float d2 = fmaf(d1, tt, p2);
    const float v = fmaf(p2, tt, cc.x);
    if (CSW) {
      vc += q2 * v;
      dvt = (d2 * q.y) * rinv;  // -(q2 / r) dV/dr
    } else {
      const float dv = -(d2 * a.ct_idr) * rinv;  // -(1/r) dV/dr
      vc += q2 * (v * sw - B.y);
      dvt = q2 * (dv * sw + v * dsw);
    }
  }
  if (lj) {  // r^-6 (C12 r^-6 - C6) and 6 r^-6 (2 C12 r^-6 - C6) / r^2
    if (CSW) {
      A = sA[t];
      B = sB[t];
      if (a.swit) switch_at();
    }
    const float r2i = rinv * rinv;
    const float r6i = r2i * r2i * r2i, t12 = A.z * r6i;
    const float v = r6i * (t12 - A.w);
    const float dv = (6.f * r6i) * (2.f * t12 - A.w) * r2i;
    vlj += v * sw - B.x;
    dvt += dv * sw + v * dsw;
  }
  return dvt;
}

// Mixed-precision pair kernel over the full list.  A wavefront owns site i;
// its row is taken in chunks of 64 entries.  First every lane stages one
// entry: the bead-independent part of the separation, the imaged centroid
// separation D_ij (FP64: xmod_i - xmod_j - h nint(hi (xmod_i - xmod_j)),
// period.c:121-206; without PIMD the imaged x_i - x_j itself), rounded to
// FP32, with the type pair's cutoff (not TM = 1) and q_i q_j, and the byte
// offset of site j's displacements, into the wavefront's LDS slice.  Then
// the (slot, bead) lanes evaluate the chunk's pairs in FP32: r = u_i - u_j +
// D_ij with the bead displacements u = x - xmod, the LJ and erfc terms of
// pair_direct (erfc(alp r)/r from an LDS table of cubic pieces of PINY's
// polynomial).
// Energies, virial and forces are summed in FP32 within a chunk and in FP64
// across chunks and lanes.
//
// TM selects how the terms are formed.  TM = 1: the LJ and erfc terms come
// whole from the switched, shifted function tables (build_pair_tables): per
// (pair, bead) one table index and one cubic piece per term, no switch
// arithmetic.  TM = 2 (every type pair with the same erfc function): the erfc
// term from its switched, shifted table, LJ and its switch evaluated
// directly (pair_direct_f32<true>).  TM = 0: LJ, the erfc table and the
// switch combined in pair_direct_f32<false>.
template <int B, bool PIMD, int TM, bool RESPA>
__global__ void __launch_bounds__(256) k_pair_mixed(PairArgs a) {
  constexpr bool TAB = TM == 1;
  constexpr int S = 64 / B;
  // the pair loop loads displacements up to 4 S entries past a chunk's end
  constexpr int SPAD = 64 + 4 * S;
  __shared__ float4 sppA[MAX_PAIRS], sppB[MAX_PAIRS];  // pair_direct_f32's constants per type pair
  __shared__ int4 stt[MAX_PAIRS];  // (cut^2 as bits, TAB: erfc table, LJ table or -1, culling radius as bits)
  extern __shared__ float4 sct[];                       // the tables: TAB nftab * CT_N, else CT_N + 1 pieces
  __shared__ float4 sdq[4][64];
  __shared__ int styp[4][64];  // type pair of the staged entries
  // not TAB: sdq's w holds the entry's cut^2, negated for a type pair with
  // an LJ term, and sq2 its charge product and that times -1/dr of the erfc
  // table (read only inside the cutoff)
  __shared__ float2 sq2[4][64];
  // byte offset of site j's first bead displacement in u per staged entry,
  // 0 past the chunk's end (a valid address whose value goes unused)
  __shared__ unsigned soff[4][PIMD ? SPAD : 1];
  for (int t = threadIdx.x; t < a.npair; t += blockDim.x) {
    const double *pp = a.pp + (size_t)t * PP_NUM;
    sppA[t] = make_float4((float)pp[PP_CUT2], (float)pp[PP_CUT], (float)pp[PP_C12], (float)pp[PP_C6]);
    sppB[t] = make_float4((float)pp[PP_VSLJ], (float)pp[PP_VSC], (float)pp[PP_LJ], (float)pp[PP_CUTRES]);
    stt[t] = make_int4(__float_as_int((float)pp[PP_CUT2]), a.ftoff[t].x, a.ftoff[t].y,
                       __float_as_int((float)pp[PP_CULL]));
  }
  const int ntab = TAB ? a.nftab * CT_N : CT_N + 1;
  const float4 *src = TAB ? a.ftab : (TM == 2 ? a.cswtab : a.ctab);
  for (int k = threadIdx.x; k < ntab; k += blockDim.x) sct[k] = src[k];
  __syncthreads();
  const float r0i = a.ct_r0 * a.ct_idr;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (PIMD)
    for (int k = 64 + lane; k < SPAD; k += 64) soff[w][k] = 0u;
  const int slot = lane / B;
  const int bead = (lane % B) + blockIdx.y * B;
  const bool bact = bead < a.P;
  const int P = a.P;
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwave = (int)((gridDim.x * blockDim.x) >> 6);
  const Cell &c = a.cell;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  double vlj = 0, vc = 0, p11 = 0, p22 = 0, p33 = 0, p12 = 0, p13 = 0, p23 = 0, npr = 0;
  for (int i = wave; i < a.N; i += nwave) {
    const size_t ib = (size_t)i * P + (bact ? bead : 0);
    const double4 si = a.site[i];
    const float4 ui = PIMD ? a.u[ib] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int cnt = a.cnt[i];
    const int *row = a.jlist + (size_t)i * a.cap;
    double fxi = 0.0, fyi = 0.0, fzi = 0.0;
    for (int k0 = 0; k0 < cnt; k0 += 64) {
      const int ks = k0 + lane;
      // entries that no bead can bring inside the cutoff
      // (|D| - max|u_i| - max|u_j| >= r_c, with room for the FP32
      // rounding) are dropped here and the chunk is compacted
      bool keep = false;
      int e = 0;
      double dx = 0.0, dy = 0.0, dz = 0.0, q2 = 0.0;
      if (ks < cnt) {
        e = row[ks];
        const int j = e & ENTRY_MASK;
        const double4 sj = a.site[j];
        if (PIMD) {
          min_image_sep(c, 1, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, si.x, si.y, si.z, sj.x, sj.y, sj.z, dx, dy, dz);
          dx += si.x - sj.x;
          dy += si.y - sj.y;
          dz += si.z - sj.z;
        } else {
          min_image_sep(c, 0, a.x[i], a.y[i], a.z[i], a.x[j], a.y[j], a.z[j], 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, dx,
                        dy, dz);
        }
        const double lim =
            ((double)__int_as_float(stt[e >> ENTRY_SHIFT].w) + (PIMD ? a.umax[i] + a.umax[j] : 0.0)) *
            (1.0 + 1e-5);
        keep = dx * dx + dy * dy + dz * dz < lim * lim;
        q2 = si.w * sj.w;
      }
      const uint64_t mk = __ballot(keep);
      const int nk = __popcll(mk);
      if (keep) {
        const int pos = __popcll(mk & below);
        float dw = (float)q2;
        styp[w][pos] = e >> ENTRY_SHIFT;
        if (!TAB) {
          const int t = e >> ENTRY_SHIFT;
          const float c2 = sppA[t].x;
          sq2[w][pos] = make_float2(dw, -dw * a.ct_idr);
          dw = sppB[t].z != 0.f ? -c2 : c2;
        }
        sdq[w][pos] = make_float4((float)dx, (float)dy, (float)dz, dw);
        if (PIMD) soff[w][pos] = (unsigned)(e & ENTRY_MASK) * (unsigned)P * (unsigned)sizeof(float4);
      }
      if (PIMD && lane >= nk) soff[w][lane] = 0u;
      wave_sync();
      float cvlj = 0.f, cvc = 0.f, c11 = 0.f, c22 = 0.f, c33 = 0.f, c12 = 0.f, c13 = 0.f, c23 = 0.f;
      float cfx = 0.f, cfy = 0.f, cfz = 0.f;
      int cn = 0;
      const unsigned boff = (unsigned)(bact ? bead : 0) * (unsigned)sizeof(float4);
      // one pair of the chunk for this lane's bead, with the displacement uj
      // of site j
      auto pair = [&](int p, const float4 &uj) {
        const float4 d = sdq[w][p];
        const int t = styp[w][p];
        float dx = d.x, dy = d.y, dz = d.z;
        if (PIMD) {
          dx += ui.x - uj.x;
          dy += ui.y - uj.y;
          dz += ui.z - uj.z;
        }
        const float r2 = dx * dx + dy * dy + dz * dz;
        float dvd;
        if constexpr (TAB) {
          const int4 T = stt[t];
          if (!bact || !(r2 < __int_as_float(T.x))) return;
          const float rinv = __frsqrt_rn(r2);
          // the piece of r = r2 rinv; below the grid the first piece extrapolates
          const float u = (r2 * rinv) * a.ct_idr - r0i;
          const unsigned k = min(__float2uint_rz(u), (unsigned)(CT_N - 1));
          const float tt = u - (float)k;
          float dv = 0.f;  // dV/dt of the pair
          if (d.w != 0.f) {
            const float4 cc = sct[T.y + k];
            cvc += d.w * (cc.x + tt * (cc.y + tt * (cc.z + tt * cc.w)));
            dv = d.w * (cc.y + tt * (2.f * cc.z + 3.f * tt * cc.w));
          }
          if (T.z >= 0) {
            const float4 cl = sct[T.z + k];
            cvlj += cl.x + tt * (cl.y + tt * (cl.z + tt * cl.w));
            dv += cl.y + tt * (2.f * cl.z + 3.f * tt * cl.w);
          }
          dvd = -(dv * a.ct_idr) * rinv;  // -(1/r) dV/dr
        } else {
          const float c2 = d.w;
          if (!bact || !(r2 < fabsf(c2))) return;
          const float rinv = __frsqrt_rn(r2);
          dvd = pair_direct_f32<TM == 2>(a, sppA, sppB, t, c2 < 0.f, sq2[w][p], r2 * rinv, rinv, cvlj, cvc, sct);
          if constexpr (RESPA) {  // respa_weight in FP32; sppB[t].w is the RESPA cutoff
            float fm = (float)a.w_full;
            if (a.respa != RESPA_FULL) {
              float br = (r2 * rinv - sppB[t].w + (float)a.rheal) * (float)a.irheal;
              br = fminf(fmaxf(br, 0.f), 1.f);
              const float sw = br * br * (br * 2.f - 3.f) + 1.f;
              fm = a.respa == RESPA_RES ? sw * (float)a.w_res : sw * (float)a.w_res + (1.f - sw) * (float)a.w_full;
            }
            dvd *= fm;
          }
        }
        const float fx1 = dx * dvd, fy1 = dy * dvd, fz1 = dz * dvd;
        cfx += fx1;
        cfy += fy1;
        cfz += fz1;
        c11 += dx * fx1;
        c22 += dy * fy1;
        c33 += dz * fz1;
        c12 += dx * fy1;
        c13 += dx * fz1;
        c23 += dy * fz1;
        cn++;
      };
      // the displacement of pair p into u (p < SPAD; past the chunk's end
      // it reads this bead of site 0, unused)
      // The load is issued on every path: with the load under a branch the
      // compiler waits for every outstanding load before each pair
      // (vmcnt(0)), including the prefetches, and the pair loop runs at the
      // memory latency.  A 32-bit byte offset from the kernel argument lets
      // the load take the scalar base address.
      auto fetch = [&](float4 &u, int p) {
        if (PIMD) u = *(const float4 *)((const char *)a.u + (soff[w][p] + boff));
      };
      // four pairs per iteration, each displacement loaded a whole iteration
      // ahead into a register of its own: no register copies, which would
      // wait for the load (one pair ahead through copies: 0.46 ms; two: 0.42
      // ms, and three raised the kernel above 128 VGPRs, to 3 wavefronts per
      // SIMD: 0.67 ms; four ahead, unrolled: 0.41 ms at 120 VGPRs)
      float4 uA = make_float4(0.f, 0.f, 0.f, 0.f), uB = uA, uC = uA, uD = uA;
      fetch(uA, slot);
      fetch(uB, slot + S);
      fetch(uC, slot + 2 * S);
      fetch(uD, slot + 3 * S);
      int p = slot;
      for (; p + 3 * S < nk; p += 4 * S) {
        pair(p, uA);
        fetch(uA, p + 4 * S);
        pair(p + S, uB);
        fetch(uB, p + 5 * S);
        pair(p + 2 * S, uC);
        fetch(uC, p + 6 * S);
        pair(p + 3 * S, uD);
        fetch(uD, p + 7 * S);
      }
      if (p < nk) pair(p, uA);
      if (p + S < nk) pair(p + S, uB);
      if (p + 2 * S < nk) pair(p + 2 * S, uC);
      vlj += (double)cvlj;
      vc += (double)cvc;
      p11 += (double)c11;
      p22 += (double)c22;
      p33 += (double)c33;
      p12 += (double)c12;
      p13 += (double)c13;
      p23 += (double)c23;
      npr += (double)cn;
      fxi += (double)cfx;
      fyi += (double)cfy;
      fzi += (double)cfz;
      wave_sync();
    }
    for (int off = 32; off >= B; off >>= 1) {
      fxi += __shfl_xor(fxi, off, 64);
      fyi += __shfl_xor(fyi, off, 64);
      fzi += __shfl_xor(fzi, off, 64);
    }
    if (slot == 0 && bact) {  // the first writer of this step's forces
      a.fx[ib] = fxi;
      a.fy[ib] = fyi;
      a.fz[ib] = fzi;
    }
  }
  const double vals[PAIR_NV] = {vlj * 0.5, vc * 0.5, p11 * 0.5, p22 * 0.5, p33 * 0.5,
                                p12 * 0.5, p13 * 0.5, p23 * 0.5, npr * 0.5};
  block_partials<PAIR_NV>(vals, B, blockIdx.y * B, P, a.part, blockIdx.x, gridDim.x);
}

// FP64 pair kernel over the full list, staged like the mixed one: per
// chunk of 64 entries one lane per entry forms, in FP64, the imaged
// centroid separation plus the partner's centroid, F = D_ij + xmod_j, and
// drops the entries no bead can bring inside the cutoff; the (slot, bead)
// lanes then form r = (x_i - xmod_i) + F - x_j (the separation the
// reference forms as x_i - x_j plus the xmod image shift, to rounding) and
// evaluate the term with pair_term<MODE> (tables or direct).  Without beads
// F is the imaged separation itself.
template <int B, int MODE>
__device__ __forceinline__ void pair_staged(PairArgs a) {
  __shared__ double spp[MAX_PAIRS * PP_NUM];
  __shared__ double4 sdf[4][64];
  __shared__ int sent[4][64];
  for (int k = threadIdx.x; k < a.npair * PP_NUM; k += blockDim.x) spp[k] = a.pp[k];
  __syncthreads();
  constexpr int S = 64 / B;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int slot = lane / B;
  const int bead = (lane % B) + blockIdx.y * B;
  const bool bact = bead < a.P;
  const int P = a.P;
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwave = (int)((gridDim.x * blockDim.x) >> 6);
  const Cell &c = a.cell;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  PairAcc s{};
  for (int i = wave; i < a.N; i += nwave) {
    const size_t ib = (size_t)i * P + (bact ? bead : 0);
    const double4 si = a.site[i];
    double uix = 0.0, uiy = 0.0, uiz = 0.0;
    if (a.pimd) {
      uix = a.x[ib] - si.x;
      uiy = a.y[ib] - si.y;
      uiz = a.z[ib] - si.z;
    }
    const int cnt = a.cnt[i];
    const int *row = a.jlist + (size_t)i * a.cap;
    s.fx = s.fy = s.fz = 0.0;
    for (int k0 = 0; k0 < cnt; k0 += 64) {
      const int ks = k0 + lane;
      bool keep = false;
      int e = 0;
      double dx = 0.0, dy = 0.0, dz = 0.0, q2 = 0.0, lim = 0.0;
      if (ks < cnt) {
        e = row[ks];
        const int j = e & ENTRY_MASK;
        const double4 sj = a.site[j];
        if (a.pimd) {
          min_image_sep(c, 1, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, si.x, si.y, si.z, sj.x, sj.y, sj.z, dx, dy, dz);
          dx += si.x - sj.x;
          dy += si.y - sj.y;
          dz += si.z - sj.z;
          lim = (spp[(e >> ENTRY_SHIFT) * PP_NUM + PP_CULL] + a.umax[i] + a.umax[j]) * (1.0 + 1e-9);
          keep = dx * dx + dy * dy + dz * dz < lim * lim;
          dx += sj.x;  // F = D + xmod_j
          dy += sj.y;
          dz += sj.z;
        } else {
          min_image_sep(c, 0, a.x[i], a.y[i], a.z[i], a.x[j], a.y[j], a.z[j], 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, dx,
                        dy, dz);
          lim = spp[(e >> ENTRY_SHIFT) * PP_NUM + PP_CULL] * (1.0 + 1e-9);
          keep = dx * dx + dy * dy + dz * dz < lim * lim;
        }
        q2 = si.w * sj.w;
      }
      const uint64_t mk = __ballot(keep);
      if (keep) {
        const int pos = __popcll(mk & below);
        sdf[w][pos] = make_double4(dx, dy, dz, q2);
        sent[w][pos] = e;
      }
      wave_sync();
      const int nk = __popcll(mk);
      // two pairs per iteration, each bead position of site j loaded a whole
      // iteration ahead into registers of its own (no copies that would
      // wait for the load)
      const int bsel = bact ? bead : 0;
      auto pos = [&](int p) {
        double3 r = make_double3(0.0, 0.0, 0.0);
        if (a.pimd && p < nk) {
          const size_t jb = (size_t)(sent[w][p] & ENTRY_MASK) * P + bsel;
          r = make_double3(a.x[jb], a.y[jb], a.z[jb]);
        }
        return r;
      };
      auto pair = [&](int p, const double3 &xj) {
        const double4 f = sdf[w][p];
        const int t = sent[w][p] >> ENTRY_SHIFT;
        double rx = f.x, ry = f.y, rz = f.z;
        if (a.pimd) {
          rx = uix + (f.x - xj.x);
          ry = uiy + (f.y - xj.y);
          rz = uiz + (f.z - xj.z);
        }
        double dvd;
        if (bact) pair_term<MODE>(a, spp + t * PP_NUM, t, f.w, rx, ry, rz, rx * rx + ry * ry + rz * rz, s, dvd);
      };
      double3 xA = pos(slot), xB = pos(slot + S);
      int p = slot;
      for (; p + S < nk; p += 2 * S) {
        pair(p, xA);
        xA = pos(p + 2 * S);
        pair(p + S, xB);
        xB = pos(p + 3 * S);
      }
      if (p < nk) pair(p, xA);
      wave_sync();
    }
    double fxi = s.fx, fyi = s.fy, fzi = s.fz;
    for (int off = 32; off >= B; off >>= 1) {
      fxi += __shfl_xor(fxi, off, 64);
      fyi += __shfl_xor(fyi, off, 64);
      fzi += __shfl_xor(fzi, off, 64);
    }
    if (slot == 0 && bact) {  // the first writer of this step's forces
      a.fx[ib] = fxi;
      a.fy[ib] = fyi;
      a.fz[ib] = fzi;
    }
  }
  const double vals[PAIR_NV] = {s.vlj * 0.5, s.vc * 0.5, s.p11 * 0.5, s.p22 * 0.5, s.p33 * 0.5,
                                s.p12 * 0.5, s.p13 * 0.5, s.p23 * 0.5, s.npr * 0.5};
  block_partials<PAIR_NV>(vals, B, blockIdx.y * B, P, a.part, blockIdx.x, gridDim.x);
}

// The table form holds 4 wavefronts per SIMD (at most 128 VGPRs, no
// spills: 0.89 instead of 1.08 ms when the compiler chose 129); the direct
// form, whose polynomial needs more registers, would spill under that
// bound and runs at what the compiler picks.
template <int B>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_pair_staged_tab(PairArgs a) {
  pair_staged<B, 0>(a);
}
template <int B>
__global__ void __launch_bounds__(256) k_pair_staged_dir(PairArgs a) {
  pair_staged<B, 1>(a);
}

// The largest |x - xmod| of every site over its beads, in FP64 (one thread
// per site), for the FP64 staged kernel's culling.
__global__ void k_site_umax64(int N, int P, const double *__restrict__ x, const double *__restrict__ y,
                              const double *__restrict__ z, const double *__restrict__ xm,
                              const double *__restrict__ ym, const double *__restrict__ zm,
                              double *__restrict__ umax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double m = 0.0;
  for (int b = 0; b < P; b++) {
    const size_t t = (size_t)i * P + b;
    const double dx = x[t] - xm[i], dy = y[t] - ym[i], dz = z[t] - zm[i];
    m = fmax(m, dx * dx + dy * dy + dz * dz);
  }
  umax[i] = sqrt(m);
}

// ---------------------------------------------------------------------------
// Parameter gradients of the real-space term (the backward of force_npol).
// Per (site, bead): phi_i = sum_j q_j vc(r_ij), the derivative of the bead's
// energy with respect to q_i (vc: the erfc-Coulomb function per unit charge
// product, switched and shifted as tabulated; from the tables when MODE is 0,
// directly otherwise).  Per LJ type pair slot: the LJ energy per unit eps and
// sigma dV/dsigma per unit eps, from the LJ form itself.  Sites without
// charge get phi = 0: the set of charged sites is fixed.  Full list only (each
// pair is seen from both ends, so the pair sums are halved).
template <int B, int MODE>
__global__ void __launch_bounds__(256) k_pair_grad(PairArgs a) {
  __shared__ double spp[MAX_PAIRS * PP_NUM];
  for (int k = threadIdx.x; k < a.npair * PP_NUM; k += blockDim.x) spp[k] = a.pp[k];
  __syncthreads();
  constexpr int S = 64 / B;
  const int lane = threadIdx.x & 63;
  const int slot = lane / B;
  const int bead = (lane % B) + blockIdx.y * B;
  const bool bact = bead < a.P;
  const int P = a.P;
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwave = (int)((gridDim.x * blockDim.x) >> 6);
  const Cell &c = a.cell;
  double ge[MAX_LJ_GRAD] = {0, 0, 0, 0}, gs[MAX_LJ_GRAD] = {0, 0, 0, 0};
  for (int i = wave; i < a.N; i += nwave) {
    const size_t ib = (size_t)i * P + (bact ? bead : 0);
    const double xi = a.x[ib], yi = a.y[ib], zi = a.z[ib];
    const double4 si = a.site[i];
    const int cnt = a.cnt[i];
    const int *row = a.jlist + (size_t)i * a.cap;
    double phi = 0.0;
    for (int k = slot; k < cnt; k += S) {
      const int e = row[k];
      const int j = e & ENTRY_MASK, t = e >> ENTRY_SHIFT;
      const double *pp = spp + t * PP_NUM;
      const size_t jb = (size_t)j * P + (bact ? bead : 0);
      const double4 sj = a.site[j];
      double dx, dy, dz;
      const double r2d = min_image_sep(c, a.pimd, xi, yi, zi, a.x[jb], a.y[jb], a.z[jb], si.x, si.y, si.z, sj.x,
                                       sj.y, sj.z, dx, dy, dz);
      const double rd = sqrt(r2d);
      if (!bact || !(rd < pp[PP_CUT])) continue;
      const double rinv = 1.0 / rd;
      double sw = 1.0;
      if (a.swit) {
        double rp = (rd - pp[PP_CUT] + a.rheal) * a.irheal;
        rp = rp < 1.0 ? rp : 1.0;
        rp = rp > 0.0 ? rp : 0.0;
        sw = 1.0 + rp * rp * (2.0 * rp - 3.0);
      }
      if (sj.w != 0.0) {
        double vc;
        if constexpr (MODE == 0) {
          const double rmin = pp[PP_RMIN], dri = pp[PP_DRI];
          int kk = (int)(((rd - rmin) * dri) + 0.5) + 3;
          kk = min(kk, a.nsplin - 2);
          kk = max(kk, 3);
          const double rsp = ((double)(kk - 3)) * pp[PP_DR] + rmin;
          vc = spline4(a.tab_c + (size_t)t * a.nsplin, kk, (rd - rsp) * dri, false);
        } else {
          double vlj_u, vc_u;
          pair_direct<double>(a, pp, 1.0, rd, rinv, vlj_u, vc_u);
          vc = vc_u;
        }
        phi += sj.w * vc;
      }
      const int g = (int)pp[PP_GSLOT];
      if (g >= 0) {
        const double r2i = rinv * rinv;
        const double s6 = pp[PP_SIG6] * (r2i * r2i * r2i);
        const double u = 4.0 * s6 * (s6 - 1.0) * sw - pp[PP_VSLJ1];
        const double d = 4.0 * (12.0 * s6 * s6 - 6.0 * s6) * sw - pp[PP_DSLJ1];
#pragma unroll
        for (int q = 0; q < MAX_LJ_GRAD; q++)
          if (q == g) {
            ge[q] += u;
            gs[q] += d;
          }
      }
    }
    for (int off = 32; off >= B; off >>= 1) phi += __shfl_xor(phi, off, 64);
    if (slot == 0 && bact) a.phi[ib] = (si.w != 0.0) ? phi : 0.0;
  }
  double vals[PAIRG_NV];
#pragma unroll
  for (int q = 0; q < MAX_LJ_GRAD; q++) {
    vals[q] = 0.5 * ge[q];
    vals[MAX_LJ_GRAD + q] = 0.5 * gs[q];
  }
  block_partials<PAIRG_NV>(vals, B, blockIdx.y * B, P, a.part, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// bonded terms and excluded-pair corrections

__device__ __forceinline__ void intra_sep(const IntraArgs &a, int i, int j, int b, double &dx, double &dy,
                                          double &dz) {
  const size_t ib = (size_t)i * a.P + b, jb = (size_t)j * a.P + b;
  dx = a.x[ib] - a.x[jb];
  dy = a.y[ib] - a.y[jb];
  dz = a.z[ib] - a.z[jb];
  if (a.intra_perds) {
    const Cell &c = a.cell;
    if (!a.pimd) {
      const double asx = dx * c.hi[0] + dy * c.hi[3] + dz * c.hi[6];
      const double asy = dx * c.hi[1] + dy * c.hi[4] + dz * c.hi[7];
      const double asz = dx * c.hi[2] + dy * c.hi[5] + dz * c.hi[8];
      const double sx = asx - nint_d(asx), sy = asy - nint_d(asy), sz = asz - nint_d(asz);
      dx = sx * c.h[0] + sy * c.h[3] + sz * c.h[6];
      dy = sx * c.h[1] + sy * c.h[4] + sz * c.h[7];
      dz = sx * c.h[2] + sy * c.h[5] + sz * c.h[8];
    } else {
      const double dxs = a.xm[i] - a.xm[j], dys = a.ym[i] - a.ym[j], dzs = a.zm[i] - a.zm[j];
      const double sx = -nint_d(dxs * c.hi[0] + dys * c.hi[3] + dzs * c.hi[6]);
      const double sy = -nint_d(dxs * c.hi[1] + dys * c.hi[4] + dzs * c.hi[7]);
      const double sz = -nint_d(dxs * c.hi[2] + dys * c.hi[5] + dzs * c.hi[8]);
      dx += (sx * c.h[0] + sy * c.h[3] + sz * c.h[6]);
      dy += (sx * c.h[1] + sy * c.h[4] + sz * c.h[7]);
      dz += (sx * c.h[2] + sy * c.h[5] + sz * c.h[8]);
    }
  }
}

__device__ __forceinline__ void add_force(const IntraArgs &a, int i, int b, double fx, double fy, double fz) {
  const size_t ib = (size_t)i * a.P + b;
  atomic_add(a.fx + ib, fx);
  atomic_add(a.fy + ib, fy);
  atomic_add(a.fz + ib, fz);
}

// Grid-stride loop over (term, bead) items t = term*P + bead.  The block size
// is a multiple of P and so is the grid stride: each thread keeps one bead,
// accumulates its terms' energy and virial in registers, and the block
// writes one partial row per bead.
#define PMX_TERM_LOOP(nterm)                                                                       \
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < (size_t)(nterm) * a.P;       \
       t += (size_t)gridDim.x * blockDim.x)

__device__ __forceinline__ void add_virial(double (&acc)[INTRA_NV], double e, double dx, double dy, double dz,
                                           double fx, double fy, double fz) {
  acc[0] += e;
  acc[1] += dx * fx;
  acc[2] += dy * fy;
  acc[3] += dz * fz;
  acc[4] += dx * fy;
  acc[5] += dx * fz;
  acc[6] += dy * fz;
}

// One term of each intramolecular kind for bead b: separations from
// sep(i, j, dx, dy, dz) (intra_sep, or positions held in LDS), energy and
// virial into acc, forces through add(site, fx, fy, fz) (global atomics in
// the per-term kernels, per-thread LDS rows in k_intra_mol).

// ecor, ecorr.c:28-265 (no switching on the table)
template <class Sep, class Add>
__device__ __forceinline__ void ecor_term(const IntraArgs &a, int i, int j, double q12, double (&acc)[INTRA_NV],
                                          Sep sep, Add add) {
  double dx, dy, dz;
  sep(i, j, dx, dy, dz);
  const double r12 = sqrt(dx * dx + dy * dy + dz * dz);
  int kk = (int)(((r12 - a.e_rmin) * a.e_dri) + 0.5) + 3;
  kk = min(kk, a.nsplin - 2);
  kk = max(kk, 3);
  const double rsp = ((double)(kk - 3)) * a.e_dr + a.e_rmin;
  const double del = (r12 - rsp) * a.e_dri;
  const double e = spline4(a.ecor_tab, kk, del, false) * q12;
  const double dvt = spline4(a.ecor_tab, kk, del, true) * q12;
  const double fx = dx * dvt, fy = dy * dvt, fz = dz * dvt;
  add(i, fx, fy, fz);
  add(j, -fx, -fy, -fz);
  add_virial(acc, e, dx, dy, dz, fx, fy, fz);
}

// bond, bond.c:29-312
template <class Sep, class Add>
__device__ __forceinline__ void bond_term(const IntraArgs &a, int i, int j, int ty, double (&acc)[INTRA_NV], Sep sep,
                                          Add add) {
  const double *cc = a.bc + 7 * ty;
  double dx, dy, dz;
  sep(i, j, dx, dy, dz);
  const double r12 = sqrt(dx * dx + dy * dy + dz * dz);
  const double rr0 = r12 - a.beq[ty];
  const double e =
      (((((cc[6] * rr0 + cc[5]) * rr0 + cc[4]) * rr0 + cc[3]) * rr0 + cc[2]) * rr0 + cc[1]) * rr0 + cc[0];
  const double dvb = (((((6.0 * cc[6]) * rr0 + 5.0 * cc[5]) * rr0 + 4.0 * cc[4]) * rr0 + 3.0 * cc[3]) * rr0 +
                      2.0 * cc[2]) * rr0 + cc[1];
  const double pre = -dvb / r12;
  const double fx = dx * pre, fy = dy * pre, fz = dz * pre;
  add(i, fx, fy, fz);
  add(j, -fx, -fy, -fz);
  add_virial(acc, e, dx, dy, dz, fx, fy, fz);
}

// bend, bend.c:28-434
template <class Sep, class Add>
__device__ __forceinline__ void bend_term(const IntraArgs &a, int i1, int i2, int i3, int ty,
                                          double (&acc)[INTRA_NV], Sep sep, Add add) {
  const double *cc = a.ac + 7 * ty, *ss = a.as + 7 * ty;
  double dx12, dy12, dz12, dx23, dy23, dz23;
  sep(i1, i2, dx12, dy12, dz12);
  sep(i3, i2, dx23, dy23, dz23);
  const double r122i = 1.0 / (dx12 * dx12 + dy12 * dy12 + dz12 * dz12);
  const double r322i = 1.0 / (dx23 * dx23 + dy23 * dy23 + dz23 * dz23);
  const double rpmag = sqrt(r122i) * sqrt(r322i);
  double cost = (dx12 * dx23 + dy12 * dy23 + dz12 * dz23) * rpmag;
  cost = (cost < 1.0 ? cost : 1.0);
  cost = (cost > -1.0 ? cost : -1.0);
  double sint = sqrt(1.0 - cost * cost);
  sint = (sint > 1.0e-8 ? sint : 1.0e-8);
  const double cos122 = cost * r122i, cos322 = cost * r322i;
  const double vbc = ((((((cc[6] * cost + cc[5]) * cost + cc[4]) * cost + cc[3]) * cost + cc[2]) * cost + cc[1]) *
                          cost + cc[0]);
  const double sisum = (((((ss[6] * sint + ss[5]) * sint + ss[4]) * sint + ss[3]) * sint + ss[2]) * sint);
  const double e = vbc + sisum * cost + ss[1] * sint;
  const double dvbc = (((((6.0 * cc[6]) * cost + 5.0 * cc[5]) * cost + 4.0 * cc[4]) * cost + 3.0 * cc[3]) * cost +
                       2.0 * cc[2]) * cost + cc[1];
  double dvbs = ((((5.0 * ss[6]) * sint + 4.0 * ss[5]) * sint + 3.0 * ss[4]) * sint + 2.0 * ss[3]) * sint + ss[2];
  dvbs = dvbs * cost + ss[1];
  const double pre = -(dvbc + sisum - dvbs * cost / sint);
  const double f1x = (dx23 * rpmag - dx12 * cos122) * pre;
  const double f1y = (dy23 * rpmag - dy12 * cos122) * pre;
  const double f1z = (dz23 * rpmag - dz12 * cos122) * pre;
  const double f3x = (dx12 * rpmag - dx23 * cos322) * pre;
  const double f3y = (dy12 * rpmag - dy23 * cos322) * pre;
  const double f3z = (dz12 * rpmag - dz23 * cos322) * pre;
  add(i1, f1x, f1y, f1z);
  add(i2, -(f1x + f3x), -(f1y + f3y), -(f1z + f3z));
  add(i3, f3x, f3y, f3z);
  acc[0] += e;
  acc[1] += dx12 * f1x + dx23 * f3x;
  acc[2] += dy12 * f1y + dy23 * f3y;
  acc[3] += dz12 * f1z + dz23 * f3z;
  acc[4] += dx12 * f1y + dx23 * f3y;
  acc[5] += dx12 * f1z + dx23 * f3z;
  acc[6] += dy12 * f1z + dy23 * f3z;
}

// Per-term kernels (topologies that do not split into small molecules):
// one thread per (term, bead), forces by global atomics.
// The forces are added times the term's RESPA weight W (the virial rows
// stay unweighted; the reduction weighs them).
#define PMX_TERM_KERNEL(name, nterm, W, CALL)                                                 \
  __global__ void __launch_bounds__(256) name(IntraArgs a) {                                   \
    double acc[INTRA_NV] = {0, 0, 0, 0, 0, 0, 0};                                              \
    const double w = W;                                                                        \
    PMX_TERM_LOOP(nterm) {                                                                     \
      const int k = (int)(t / a.P), b = (int)(t % a.P);                                        \
      auto sep = [&](int i, int j, double &dx, double &dy, double &dz) { intra_sep(a, i, j, b, dx, dy, dz); }; \
      auto add = [&](int i, double fx, double fy, double fz) { add_force(a, i, b, fx * w, fy * w, fz * w); }; \
      CALL;                                                                                    \
    }                                                                                          \
    block_partials<INTRA_NV>(acc, a.P, 0, a.P, a.part, blockIdx.x, gridDim.x);                 \
  }
PMX_TERM_KERNEL(k_ecor, a.necor, a.w_ecor, ecor_term(a, a.e1[k], a.e2[k], a.q[a.e1[k]] * a.q[a.e2[k]], acc, sep, add))
PMX_TERM_KERNEL(k_bond, a.nbond, a.w_intra, bond_term(a, a.b1[k], a.b2[k], a.btyp[k], acc, sep, add))
PMX_TERM_KERNEL(k_bend, a.nbend, a.w_intra, bend_term(a, a.a1[k], a.a2[k], a.a3[k], a.atyp[k], acc, sep, add))
#undef PMX_TERM_KERNEL

// Every intramolecular term of a force call in one pass, for topologies
// whose terms (bonds, bends, excluded pairs, ghosts) split into molecules of
// at most MOL_MAX consecutive sites that follow a few templates (local site
// indices, types, charge products and ghost coefficients; one for a water
// box).  The templates sit in LDS, so a thread's only global reads are its
// molecule's template id, positions and forces, the excluded-pair table
// and the k-space forces.  One thread per (molecule, bead) holds the
// molecule's positions and forces in LDS rows of its own (the forces
// starting from the pair forces; 6 * mol_max rows of blockDim doubles,
// dynamic LDS), adds the excluded-pair, bond and bend terms, then does what
// k_pimd_finish (or k_add_rec) and k_ghost_force do for these sites, in the
// same order, and writes the forces once: no atomics.  Energies and virials
// go to the partial rows of the ecor, bond, bend and finish stages.
__global__ void __launch_bounds__(256) k_intra_mol(IntraArgs a) {
  extern __shared__ double lds_mol[];
  __shared__ MolTpl tpl[TPL_MAX];
  const int tid = threadIdx.x, bs = blockDim.x, R = 3 * a.mol_max;
  {
    const int *src = (const int *)a.tpl;
    int *dst = (int *)tpl;
    const int nw = a.ntpl * (int)(sizeof(MolTpl) / sizeof(int));
    for (int k = tid; k < nw; k += bs) dst[k] = src[k];
  }
  __syncthreads();
  double *fl = lds_mol, *pl = lds_mol + (size_t)R * bs;  // [row][thread]
  double ae[INTRA_NV] = {0, 0, 0, 0, 0, 0, 0}, ab[INTRA_NV] = {0, 0, 0, 0, 0, 0, 0};
  double aa[INTRA_NV] = {0, 0, 0, 0, 0, 0, 0};
  double cf[FIN_NV] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  const size_t NP = (size_t)a.N * a.P;
  const double dP = (double)a.P;
  PMX_TERM_LOOP(a.nmol) {
    const int m = (int)(t / a.P), b = (int)(t % a.P);
    const int2 mt = a.mol[m];  // (first site, template)
    const int lo = mt.x;
    const MolTpl &T = tpl[mt.y];
    const int n = T.n;
    for (int s = 0; s < n; s++) {
      const size_t ib = (size_t)(lo + s) * a.P + b;
      fl[(3 * s) * bs + tid] = a.fx[ib];
      fl[(3 * s + 1) * bs + tid] = a.fy[ib];
      fl[(3 * s + 2) * bs + tid] = a.fz[ib];
      pl[(3 * s) * bs + tid] = a.x[ib];
      pl[(3 * s + 1) * bs + tid] = a.y[ib];
      pl[(3 * s + 2) * bs + tid] = a.z[ib];
    }
    // terms address sites by global index (lo + local), as the per-term kernels
    auto add = [&](int i, double fx, double fy, double fz) {
      const int s = i - lo;
      fl[(3 * s) * bs + tid] += fx;
      fl[(3 * s + 1) * bs + tid] += fy;
      fl[(3 * s + 2) * bs + tid] += fz;
    };
    auto sep = [&](int i, int j, double &dx, double &dy, double &dz) {
      if (a.intra_perds) {
        intra_sep(a, i, j, b, dx, dy, dz);
        return;
      }
      const int si = i - lo, sj = j - lo;
      dx = pl[(3 * si) * bs + tid] - pl[(3 * sj) * bs + tid];
      dy = pl[(3 * si + 1) * bs + tid] - pl[(3 * sj + 1) * bs + tid];
      dz = pl[(3 * si + 2) * bs + tid] - pl[(3 * sj + 2) * bs + tid];
    };
    // each term's forces times its RESPA weight (1 without RESPA: exact)
    auto add_e = [&](int i, double fx, double fy, double fz) { add(i, fx * a.w_ecor, fy * a.w_ecor, fz * a.w_ecor); };
    auto add_i = [&](int i, double fx, double fy, double fz) {
      add(i, fx * a.w_intra, fy * a.w_intra, fz * a.w_intra);
    };
    if (a.necor > 0)
      for (int q = 0; q < T.ne; q++) ecor_term(a, lo + T.e[q][0], lo + T.e[q][1], T.eq[q], ae, sep, add_e);
    for (int q = 0; q < T.nb; q++) bond_term(a, lo + T.bd[q][0], lo + T.bd[q][1], T.bd[q][2], ab, sep, add_i);
    for (int q = 0; q < T.na; q++)
      bend_term(a, lo + T.an[q][0], lo + T.an[q][1], lo + T.an[q][2], T.an[q][3], aa, sep, add_i);
    for (int s = 0; s < n; s++) {
      const int i = lo + s;
      double fx = fl[(3 * s) * bs + tid], fy = fl[(3 * s + 1) * bs + tid], fz = fl[(3 * s + 2) * bs + tid];
      if (a.rec) {
        const size_t r = (size_t)b * a.N + i;
        fx += a.rec[r] * a.w_rec;
        fy += a.rec[NP + r] * a.w_rec;
        fz += a.rec[2 * NP + r] * a.w_rec;
      }
      if (a.pimd) {
        fx /= dP;
        fy /= dP;
        fz /= dP;
        if (a.P > 1) {
          const double dx = pl[(3 * s) * bs + tid] - a.xm[i], dy = pl[(3 * s + 1) * bs + tid] - a.ym[i],
                       dz = pl[(3 * s + 2) * bs + tid] - a.zm[i];
          cf[0] -= dx * fx;
          cf[1] -= dy * fy;
          cf[2] -= dz * fz;
          cf[3] -= dx * fy;
          cf[4] -= dy * fx;
          cf[5] -= dx * fz;
          cf[6] -= dz * fx;
          cf[7] -= dy * fz;
          cf[8] -= dz * fy;
          cf[9] -= dx * fx + dy * fy + dz * fz;
        }
      }
      fl[(3 * s) * bs + tid] = fx;
      fl[(3 * s + 1) * bs + tid] = fy;
      fl[(3 * s + 2) * bs + tid] = fz;
    }
    for (int q = 0; q < T.ng; q++) {  // distrib_ghost_force
      const int sg = T.gh[q][0];
      const double gx = fl[(3 * sg) * bs + tid], gy = fl[(3 * sg + 1) * bs + tid], gz = fl[(3 * sg + 2) * bs + tid];
      for (int k = 0; k < T.gh[q][1]; k++) {
        const double cg = T.gc[q][k];
        add(lo + T.gh[q][2 + k], cg * gx, cg * gy, cg * gz);
      }
      fl[(3 * sg) * bs + tid] = 0.0;
      fl[(3 * sg + 1) * bs + tid] = 0.0;
      fl[(3 * sg + 2) * bs + tid] = 0.0;
    }
    for (int s = 0; s < n; s++) {
      const size_t ib = (size_t)(lo + s) * a.P + b;
      a.fx[ib] = fl[(3 * s) * bs + tid];
      a.fy[ib] = fl[(3 * s + 1) * bs + tid];
      a.fz[ib] = fl[(3 * s + 2) * bs + tid];
    }
  }
  block_partials<INTRA_NV>(ae, a.P, 0, a.P, a.part_ecor, blockIdx.x, gridDim.x);
  block_partials<INTRA_NV>(ab, a.P, 0, a.P, a.part_bond, blockIdx.x, gridDim.x);
  block_partials<INTRA_NV>(aa, a.P, 0, a.P, a.part_bend, blockIdx.x, gridDim.x);
  block_partials<FIN_NV>(cf, a.P, 0, a.P, a.part_fin, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Parameter gradients of the excluded-pair, self and background terms and of
// the bonded terms.

// One thread per (charged site, bead): phi_i += sum over the site's
// Ewald-corrected excluded partners j of q_j ecor(r_ij) (the table and
// interpolation of k_ecor), plus d(vself + vbgr)/dq_i
// (ewald3d_self_bgr.c:55-61).
__global__ void __launch_bounds__(256) k_ecor_phi(IntraArgs a) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)a.nchrg * a.P) return;
  const int i = a.ichrg[t / a.P], b = (int)(t % a.P);
  double phi = a.cself * a.q[i] + a.cbgr;
  const int o = a.ecs_off[i], n = a.ecs_num[i];
  for (int k = 0; k < n; k++) {
    const int j = a.ecs_j[o + k];
    double dx, dy, dz;
    intra_sep(a, i, j, b, dx, dy, dz);
    const double r12 = sqrt(dx * dx + dy * dy + dz * dz);
    int kk = (int)(((r12 - a.e_rmin) * a.e_dri) + 0.5) + 3;
    kk = min(kk, a.nsplin - 2);
    kk = max(kk, 3);
    const double rsp = ((double)(kk - 3)) * a.e_dr + a.e_rmin;
    phi += a.q[j] * spline4(a.ecor_tab, kk, (r12 - rsp) * a.e_dri, false);
  }
  a.phi[(size_t)i * a.P + b] += phi;
}

// Sums of a workgroup's threads in a fixed order (tree in LDS); every
// thread of the block calls it, thread 0 gets the result.
__device__ __forceinline__ double block_sum(double v, double *red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int h = blockDim.x >> 1; h > 0; h >>= 1) {
    if (tid < h) red[tid] += red[tid + h];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// One workgroup of 256 per (bond type, bead): sum over that type's bonds of
// (r - eq)^m, m = 0..6, the derivatives of V with respect to c_m
// (bond.c:150-166).
__global__ void __launch_bounds__(256) k_bond_moments(IntraArgs a) {
  __shared__ double red[256];
  const int ty = blockIdx.x, b = blockIdx.y;
  double m[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int k = threadIdx.x; k < a.nbond; k += blockDim.x) {
    if (a.btyp[k] != ty) continue;
    double dx, dy, dz;
    intra_sep(a, a.b1[k], a.b2[k], b, dx, dy, dz);
    const double rr0 = sqrt(dx * dx + dy * dy + dz * dz) - a.beq[ty];
    double p = 1.0;
#pragma unroll
    for (int q = 0; q < 7; q++) {
      m[q] += p;
      p *= rr0;
    }
  }
  for (int q = 0; q < 7; q++) {
    const double v = block_sum(m[q], red);
    if (threadIdx.x == 0) a.gbond[((size_t)b * a.nbond_type + ty) * 7 + q] = v;
  }
}

// One workgroup of 256 per (bend type, bead): cos^m (m = 0..6), sin and
// cos sin^(m-1) (m = 2..6) summed over the type's bends, with the cosine and
// sine clamped as k_bend clamps them (bend.c:253-270).
__global__ void __launch_bounds__(256) k_bend_moments(IntraArgs a) {
  __shared__ double red[256];
  const int ty = blockIdx.x, b = blockIdx.y;
  double m[BEND_GRAD_NV];
#pragma unroll
  for (int q = 0; q < BEND_GRAD_NV; q++) m[q] = 0.0;
  for (int k = threadIdx.x; k < a.nbend; k += blockDim.x) {
    if (a.atyp[k] != ty) continue;
    double dx12, dy12, dz12, dx23, dy23, dz23;
    intra_sep(a, a.a1[k], a.a2[k], b, dx12, dy12, dz12);
    intra_sep(a, a.a3[k], a.a2[k], b, dx23, dy23, dz23);
    const double r122i = 1.0 / (dx12 * dx12 + dy12 * dy12 + dz12 * dz12);
    const double r322i = 1.0 / (dx23 * dx23 + dy23 * dy23 + dz23 * dz23);
    const double rpmag = sqrt(r122i) * sqrt(r322i);
    double cost = (dx12 * dx23 + dy12 * dy23 + dz12 * dz23) * rpmag;
    cost = (cost < 1.0 ? cost : 1.0);
    cost = (cost > -1.0 ? cost : -1.0);
    double sint = sqrt(1.0 - cost * cost);
    sint = (sint > 1.0e-8 ? sint : 1.0e-8);
    double p = 1.0;
#pragma unroll
    for (int q = 0; q < 7; q++) {
      m[q] += p;
      p *= cost;
    }
    m[7] += sint;
    double sp = sint;
#pragma unroll
    for (int q = 8; q < BEND_GRAD_NV; q++) {
      m[q] += cost * sp;
      sp *= sint;
    }
  }
  for (int q = 0; q < BEND_GRAD_NV; q++) {
    const double v = block_sum(m[q], red);
    if (threadIdx.x == 0) a.gbend[((size_t)b * a.nbend_type + ty) * BEND_GRAD_NV + q] = v;
  }
}

// ---------------------------------------------------------------------------
// PIMD finish: the PME forces added (rec, bead-major), f /= P
// (energy_control_pimd.c:116-126) and the virial correction
// -sum (x - xmod) (x) f (pimd_estimators.c:223-323), grid-stride over (site,
// bead) items.
__global__ void __launch_bounds__(256) k_pimd_finish(FinishArgs a) {
  double c[FIN_NV] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  const double dP = (double)a.P;
  const size_t NP = (size_t)a.N * a.P;
  PMX_TERM_LOOP(a.N) {
    const int i = (int)(t / a.P);
    double fx = a.fx[t], fy = a.fy[t], fz = a.fz[t];
    if (a.rec) {
      const size_t r = (t - (size_t)i * a.P) * a.N + i;
      fx += a.rec[r] * a.w_rec;
      fy += a.rec[NP + r] * a.w_rec;
      fz += a.rec[2 * NP + r] * a.w_rec;
    }
    fx /= dP;
    fy /= dP;
    fz /= dP;
    a.fx[t] = fx;
    a.fy[t] = fy;
    a.fz[t] = fz;
    if (a.P > 1) {
      const double dx = a.x[t] - a.xm[i], dy = a.y[t] - a.ym[i], dz = a.z[t] - a.zm[i];
      c[0] -= dx * fx;
      c[1] -= dy * fy;
      c[2] -= dz * fz;
      c[3] -= dx * fy;
      c[4] -= dy * fx;
      c[5] -= dx * fz;
      c[6] -= dz * fx;
      c[7] -= dy * fz;
      c[8] -= dz * fy;
      c[9] -= dx * fx + dy * fy + dz * fz;
    }
  }
  block_partials<FIN_NV>(c, a.P, 0, a.P, a.part, blockIdx.x, gridDim.x);
}

// The PME forces (bead-major [3][P][N]) added to the site-major forces,
// outside PIMD.
__global__ void k_add_rec(int N, int P, const double *__restrict__ rec, double w, double *fx, double *fy,
                          double *fz) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x, NP = (size_t)N * P;
  if (t >= NP) return;
  const size_t i = t / P, r = (t - i * P) * N + i;
  fx[t] += rec[r] * w;
  fy[t] += rec[NP + r] * w;
  fz[t] += rec[2 * NP + r] * w;
}

// distrib_ghost_force, ghost_control.c:71-125
__global__ void k_ghost_force(int ng, int P, const int *__restrict__ gatom, const int *__restrict__ gncomp,
                              const int *__restrict__ gcomp, const double *__restrict__ gcoef, double *fx,
                              double *fy, double *fz) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ng * P) return;
  const int g = t / P, b = t % P;
  const size_t ig = (size_t)gatom[g] * P + b;
  const double gx = fx[ig], gy = fy[ig], gz = fz[ig];
  for (int k = 0; k < gncomp[g]; k++) {
    const size_t ia = (size_t)gcomp[4 * g + k] * P + b;
    const double c = gcoef[4 * g + k];
    atomic_add(fx + ia, c * gx);
    atomic_add(fy + ia, c * gy);
    atomic_add(fz + ia, c * gz);
  }
  fx[ig] = 0.0;
  fy[ig] = 0.0;
  fz[ig] = 0.0;
}

// Mode forces and springs of energy_control_pimd (energy_control_pimd.c:
// 198-214): the bead forces transformed to mode coordinates plus the spring
// force -2 prekf veig u of mode_energy_control (pimd_estimators.c:71-95),
// with the modes u of the bead positions.  Normal modes (path_int_md_typ
// centroid): convert_pimd_mode_cent / convert_pimd_force_cent
// (transform_cnt.c:240-316, 415-489), a DFT over the beads in PINY's storage
// order (0 centroid, P-1 the alternating mode, 2k-1 / 2k the real /
// imaginary part of frequency k), mode positions divided by P and paired
// mode forces doubled.  Staging: convert_pimd_mode_stag /
// convert_pimd_force_stag (transform_stg.c:31-160), u_1 = x_1,
// u_k = x_k - (rat1_k x_(k+1) + rat2_k x_1), the force of mode 1 the sum of
// the bead forces and the others the recurrence F_k = f_k + rat1_(k-1) F_(k-1).
//
// A group of P lanes holds one site's beads (64 / P sites per wavefront,
// P <= 64); each lane loads its bead, keeps the bead forces in fxt, stages
// (x, y, z, fx, fy, fz) in the wavefront's LDS rows, then forms one mode
// from the site's rows (reads of one address per group: LDS broadcasts).
// The spring energy of a workgroup goes to part[blockIdx.x], summed in a
// fixed order.  PT > 0 is the bead count at compile time (powers of two up
// to 64): the transform's sums unroll and their LDS reads issue together;
// PT = 0 takes P at run time.
template <int PT>
__global__ void __launch_bounds__(256) k_mode_forces(ModeArgs a) {
  __shared__ double3 spos[4][64], sfrc[4][64];
  __shared__ double2 stw[64];
  __shared__ double red[4];
  const int P = PT > 0 ? PT : a.P;
  for (int m = threadIdx.x; m < P; m += blockDim.x) stw[m] = make_double2(a.cs[2 * m], a.cs[2 * m + 1]);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int S = 64 / P, g = lane / P, m = lane - g * P, lead = g * P;
  const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * S + g;
  const bool act = g < S && i < a.N;
  const size_t t = (size_t)(act ? i : 0) * P + m;
  double3 xb = make_double3(0.0, 0.0, 0.0), fb = xb;
  if (act) {
    xb = make_double3(a.x[t], a.y[t], a.z[t]);
    fb = make_double3(a.fx[t], a.fy[t], a.fz[t]);
    if (a.fxt) {
      a.fxt[t] = fb.x;
      a.fyt[t] = fb.y;
      a.fzt[t] = fb.z;
    }
  }
  spos[w][lane] = xb;
  sfrc[w][lane] = fb;
  __syncthreads();  // the twiddles and every wavefront's rows
  double ake = 0.0;
  if (act) {
    const double3 *px = spos[w] + lead, *pf = sfrc[w] + lead;
    double3 u, F;
    if (a.stag) {
      const double3 x1 = px[0];
      if (m == 0) {
        u = x1;
        F = pf[0];
        for (int b = 1; b < P; b++) {
          F.x += pf[b].x;
          F.y += pf[b].y;
          F.z += pf[b].z;
        }
      } else {
        // 1-based bead ip = m + 1: rat1 = (ip-1)/ip, rat2 = 1/ip; the last
        // bead pairs with bead 1
        const double r1 = ((double)m) / ((double)(m + 1)), r2 = 1.0 / ((double)(m + 1));
        const double3 nx = (m + 1 < P) ? px[m + 1] : x1;
        u = make_double3(xb.x - (r1 * nx.x + r2 * x1.x), xb.y - (r1 * nx.y + r2 * x1.y),
                         xb.z - (r1 * nx.z + r2 * x1.z));
        F = pf[1];
        for (int j = 2; j <= m; j++) {
          const double r = ((double)(j - 1)) / ((double)j);
          F.x = pf[j].x + (r * F.x);
          F.y = pf[j].y + (r * F.y);
          F.z = pf[j].z + (r * F.z);
        }
      }
    } else if (PT >= 2) {
      // lanes work in pairs over half the beads each: (0, P-1) sum with the
      // weights 1 and cos(pi b) of frequencies 0 and P/2, (2k-1, 2k) with the
      // cosine and sine of frequency k; each lane hands its partner the sum
      // the partner keeps, so every bead row and twiddle is read once per pair
      const bool edge = (m == 0 || m == P - 1);
      const int k = m == 0 ? P / 2 : (m == P - 1 ? P / 2 : (m + 1) / 2);
      const int h = (m != 0) && (m == P - 1 || !(m & 1));  // half of the beads
      const int pm = m == 0 ? P - 1 : (m == P - 1 ? 0 : ((m & 1) ? m + 1 : m - 1));
      double3 uA = make_double3(0.0, 0.0, 0.0), uB = uA, FA = uA, FB = uA;
#pragma unroll
      for (int j = 0; j < PT / 2; j++) {
        const int b = h * (PT / 2) + j;
        const double2 cs = stw[(k * b) & (PT - 1)];
        const double wa = edge ? 1.0 : cs.x, wb = edge ? cs.x : cs.y;
        const double3 xv = px[b], fv = pf[b];
        uA.x += xv.x * wa;
        uA.y += xv.y * wa;
        uA.z += xv.z * wa;
        uB.x += xv.x * wb;
        uB.y += xv.y * wb;
        uB.z += xv.z * wb;
        FA.x += fv.x * wa;
        FA.y += fv.y * wa;
        FA.z += fv.z * wa;
        FB.x += fv.x * wb;
        FB.y += fv.y * wb;
        FB.z += fv.z * wb;
      }
      // lane 0 and the cosine lanes keep the A sums, lane P-1 and the sine
      // lanes the B sums
      const bool keepA = (m == 0) || ((m & 1) && m != P - 1);
      const double3 um = keepA ? uA : uB, ut = keepA ? uB : uA;
      const double3 Fm = keepA ? FA : FB, Ft = keepA ? FB : FA;
      const int src = lead + pm;
      u = make_double3(um.x + __shfl(ut.x, src, 64), um.y + __shfl(ut.y, src, 64), um.z + __shfl(ut.z, src, 64));
      F = make_double3(Fm.x + __shfl(Ft.x, src, 64), Fm.y + __shfl(Ft.y, src, 64), Fm.z + __shfl(Ft.z, src, 64));
      const double dP = (double)P;
      u.x /= dP;
      u.y /= dP;
      u.z /= dP;
      if (!edge) {
        F.x *= 2.0;
        F.y *= 2.0;
        F.z *= 2.0;
      }
    } else {
      int k = 0, im = 0;  // frequency and real (0) / imaginary (1) part of mode m
      if (m == P - 1) k = P / 2;
      else if (m > 0) {
        k = (m + 1) / 2;
        im = (m % 2 == 0);
      }
      u = make_double3(0.0, 0.0, 0.0);
      F = u;
#pragma unroll
      for (int b = 0; b < (PT > 0 ? PT : 64); b++) {
        if (PT == 0 && b >= P) break;
        // exp(+2 pi i k b / P): its real or imaginary part
        const int ph = PT > 0 ? ((k * b) & (PT - 1)) : (k * b) % P;
        const double2 cs = stw[ph];
        const double wgt = im ? cs.y : cs.x;
        const double3 xv = px[b], fv = pf[b];
        u.x += xv.x * wgt;
        u.y += xv.y * wgt;
        u.z += xv.z * wgt;
        F.x += fv.x * wgt;
        F.y += fv.y * wgt;
        F.z += fv.z * wgt;
      }
      const double dP = (double)P;
      u.x /= dP;
      u.y /= dP;
      u.z /= dP;
      if (m != 0 && m != P - 1) {
        F.x *= 2.0;
        F.y *= 2.0;
        F.z *= 2.0;
      }
    }
    const double fk = a.prekf[i] * a.veig[m + 1];
    ake = fk * (u.x * u.x + u.y * u.y + u.z * u.z);
    a.fx[t] = F.x + (-2.0 * fk * u.x);
    a.fy[t] = F.y + (-2.0 * fk * u.y);
    a.fz[t] = F.z + (-2.0 * fk * u.z);
  }
  for (int off = 32; off >= 1; off >>= 1) ake += __shfl_xor(ake, off, 64);
  if (lane == 0) red[w] = ake;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int v = 0; v < (int)(blockDim.x >> 6); v++) s += red[v];
    a.part[blockIdx.x] = s;
  }
}

// ---------------------------------------------------------------------------
// velocity Verlet pieces (int_0_to_dt2.c:83-104, int_dt2_to_dt.c:76-80):
// v += f dt/2 / m and x += v dt, in the reference's order of operations
__global__ void k_vv_kick(int n, double dt2, const double *__restrict__ mass, const double *__restrict__ fx,
                          const double *__restrict__ fy, const double *__restrict__ fz, double *vx, double *vy,
                          double *vz) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double m = mass[i];
  vx[i] += fx[i] * dt2 / m;
  vy[i] += fy[i] * dt2 / m;
  vz[i] += fz[i] * dt2 / m;
}

__global__ void k_vv_drift(int n, double dt, const double *__restrict__ vx, const double *__restrict__ vy,
                           const double *__restrict__ vz, double *x, double *y, double *z) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] += vx[i] * dt;
  y[i] += vy[i] * dt;
  z[i] += vz[i] * dt;
}

// kinetic energy 1/2 sum m v^2 (get_tvten), one atomic per wavefront
__global__ void k_kinetic(int n, const double *__restrict__ mass, const double *__restrict__ vx,
                          const double *__restrict__ vy, const double *__restrict__ vz, double *ke) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (i < n) e = 0.5 * mass[i] * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
  for (int off = 32; off >= 1; off >>= 1) e += __shfl_xor(e, off, 64);
  if ((threadIdx.x & 63) == 0) atomic_add(ke, e);
}

// ---------------------------------------------------------------------------
// launch wrappers

// Workgroups (4 wavefronts, one site per wavefront at a time) of the pair
// kernels for N sites and at most max_waves wavefronts.
// ---------------------------------------------------------------------------
// Molecule-cluster form of the mixed-precision pair kernel (CluArgs).

// The shift s of the minimum image of a separation d: s = -nint(h^-1 d), as
// period and period_pimd form it (period.c:27-206), and d + h s.
__device__ __forceinline__ void clu_image(const Cell &c, double dx, double dy, double dz, double (&s)[3],
                                          double &ix, double &iy, double &iz) {
#pragma clang fp contract(off)
  if (c.cubic) {
    s[0] = -nint_d(dx * c.hi[0]);
    s[1] = -nint_d(dy * c.hi[4]);
    s[2] = -nint_d(dz * c.hi[8]);
    ix = dx + s[0] * c.h[0];
    iy = dy + s[1] * c.h[4];
    iz = dz + s[2] * c.h[8];
  } else {
    const double fa = dx * c.hi[0] + dy * c.hi[3] + dz * c.hi[6];
    const double fb = dx * c.hi[1] + dy * c.hi[4] + dz * c.hi[7];
    const double fc = dx * c.hi[2] + dy * c.hi[5] + dz * c.hi[8];
    s[0] = -nint_d(fa);
    s[1] = -nint_d(fb);
    s[2] = -nint_d(fc);
    ix = dx + (s[0] * c.h[0] + s[1] * c.h[3] + s[2] * c.h[6]);
    iy = dy + (s[0] * c.h[1] + s[1] * c.h[4] + s[2] * c.h[7]);
    iz = dz + (s[0] * c.h[2] + s[1] * c.h[5] + s[2] * c.h[8]);
  }
}

// Positions relative to the molecule's site-0 reference, in FP32, and for
// beads the largest displacement of every site from its xmod: a group of P
// lanes holds one site's beads (64 / P sites per wavefront, P <= 64).
__global__ void __launch_bounds__(256) k_clu_pack(CluArgs a) {
  const int lane = threadIdx.x & 63;
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int P = a.P, S = 64 / P, g = lane / P, k = lane - g * P, lead = g * P;
  const int i = wave * S + g;
  const bool act = g < S && i < a.N;
  float m = 0.0f;
  if (act) {
    const int i0 = (i / a.ns) * a.ns;
    const size_t t = (size_t)i * P + k;
    const double x = a.x[t], y = a.y[t], z = a.z[t];
    a.vw[t] = make_float4((float)(x - a.xr[i0]), (float)(y - a.yr[i0]), (float)(z - a.zr[i0]), 0.0f);
    if (a.pimd) {
      const float ux = (float)(x - a.xr[i]), uy = (float)(y - a.yr[i]), uz = (float)(z - a.zr[i]);
      m = ux * ux + uy * uy + uz * uz;
    }
  }
  if (!a.pimd) return;
  float mx = m;
  if ((P & (P - 1)) == 0) {
    for (int off = 1; off < P; off <<= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  } else {
    for (int o = 1; o < P; o++) mx = fmaxf(mx, __shfl(m, lead + (k + o) % P, 64));
  }
  if (act && k == 0) a.umax_w[i] = sqrt((double)mx);
}

// Molecule rows from the force list's site rows: one wavefront per molecule
// I gathers the partner sites of its ns sites into an LDS hash on the
// partner molecule, OR-ing the site-pair bits, then writes the partners in
// ascending order of their site-0 distance at the build (a bitonic sort of
// the compacted keys, ties by partner index), so every build gives the same
// rows and neighbouring entries, which the pair kernel's slots take side by
// side, have alike site-pair sets.  flag[0] receives the longest row (the host grows
// the rows and builds again when it exceeds mcap); a molecule with more
// than CLU_HASH / 2 partners sets flag[0] to CLU_HASH + 1, and the host then
// keeps the per-site kernel.
__global__ void __launch_bounds__(64) k_clu_rows(CluArgs a) {
  __shared__ int hk[CLU_HASH];
  __shared__ int hm[CLU_HASH];
  const int lane = threadIdx.x;
  const int I = blockIdx.x;
  if (I >= a.nmol) return;
  const int ns = a.ns;
  for (int k = lane; k < CLU_HASH; k += 64) {
    hk[k] = -1;
    hm[k] = 0;
  }
  __syncthreads();
  if (lane < ns) {  // the site's distance from site 0 (reference positions): the image margin
    const int i = I * ns + lane, i0 = I * ns;
    const double dx = a.xr[i] - a.xr[i0], dy = a.yr[i] - a.yr[i0], dz = a.zr[i] - a.zr[i0];
    atomicMax(a.flag + 2, __float_as_int((float)sqrt(dx * dx + dy * dy + dz * dz) * 1.0001f));
  }
  int inserted = 0;
  for (int s = 0; s < ns; s++) {
    const int i = I * ns + s;
    const int n = a.nforce[i];
    const int *row = a.jforce + (size_t)i * a.cap;
    for (int k = lane; k < n; k += 64) {
      const int j = row[k] & ENTRY_MASK;
      const int J = j / ns, b = j - J * ns;
      unsigned h = ((unsigned)J * 2654435761u) >> (32 - 11);
      for (int probe = 0; probe < CLU_HASH; probe++) {
        const int old = atomicCAS(&hk[h], -1, J);
        if (old == -1 || old == J) {
          atomicOr(&hm[h], 1 << (s * ns + b));
          inserted += (old == -1);
          break;
        }
        h = (h + 1) & (CLU_HASH - 1);
      }
    }
  }
  __syncthreads();
  for (int off = 32; off >= 1; off >>= 1) inserted += __shfl_xor(inserted, off, 64);
  const int nu = inserted;
  if (nu > CLU_HASH / 2) {
    if (lane == 0) atomicMax(a.flag, CLU_HASH + 1);
    return;
  }
  // compact the occupied slots into keys (J, bits) at the front
  __shared__ int ck[CLU_HASH / 2], cm[CLU_HASH / 2], cj[CLU_HASH / 2];
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  int pos = 0;
  for (int b0 = 0; b0 < CLU_HASH; b0 += 64) {
    const int J = hk[b0 + lane];
    const bool occ = J >= 0;
    const uint64_t msk = __ballot(occ);
    if (occ) {
      const int p = pos + __popcll(msk & below);
      double s0[3], dx, dy, dz;
      clu_image(a.cell, a.xr[I * ns] - a.xr[J * ns], a.yr[I * ns] - a.yr[J * ns], a.zr[I * ns] - a.zr[J * ns], s0,
                dx, dy, dz);
      // the key: the distance's float bits (ordered like the distances),
      // low bits replaced by nothing; ties go by the partner index below
      ck[p] = __float_as_int((float)sqrt(dx * dx + dy * dy + dz * dz));
      cj[p] = J;
      cm[p] = hm[b0 + lane];
    }
    pos += __popcll(msk);
  }
  int n2 = 1;
  while (n2 < nu) n2 <<= 1;
  for (int k = nu + lane; k < n2; k += 64) {
    ck[k] = 0x7fffffff;
    cj[k] = 0x7fffffff;
  }
  __syncthreads();
  // bitonic sort of (ck, cm) by ck over n2 keys
  for (int size = 2; size <= n2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = lane; t < (n2 >> 1); t += 64) {
        const int lo = 2 * stride * (t / stride) + (t % stride), hi = lo + stride;
        const bool up = ((lo & size) == 0);
        const int kl = ck[lo], kh = ck[hi], jl = cj[lo], jh = cj[hi];
        if ((kl > kh || (kl == kh && jl > jh)) == up) {
          ck[lo] = kh;
          ck[hi] = kl;
          cj[lo] = jh;
          cj[hi] = jl;
          const int ml = cm[lo];
          cm[lo] = cm[hi];
          cm[hi] = ml;
        }
      }
      __syncthreads();
    }
  }
  if (nu <= a.mcap) {
    int2 *out = a.mrow_w + (size_t)I * a.mcap;
    for (int k = lane; k < nu; k += 64) out[k] = make_int2(cj[k], cm[k]);
  }
  if (lane == 0) {
    a.mcnt_w[I] = nu;
    atomicMax(a.flag, nu);
  }
}

// A wavefront owns molecule I; its lanes are (slot, bead) with B beads per
// slot.  The row is taken in chunks of 64 entries: each lane stages one, in
// FP64: the site-0 separation and its image, then for every listed site pair
// the check that its own image is the same and (beads) the cull of pairs
// that no bead can bring inside the cutoff (|D_ab| - max|u_a| - max|u_b| >=
// r_c); entries left without site pairs are dropped and the chunk
// compacted.  The (slot, bead) lanes then evaluate the chunk's entries in
// FP32: for each listed site pair one cubic piece of the whole switched,
// shifted function (erfc term per unit charge product, LJ term) from LDS;
// r beyond the cutoff reads the zero piece past the grid's end.  Forces of
// I's sites are summed per lane in FP32 over the row (a few hundred terms
// per site) and over the slots in FP64 at the end of the row; energies and
// virial per lane in FP32 over the lane's rows, then in FP64 over lanes and
// workgroups, halved for the full list's double count.
template <int B, int NS, bool PIMD, bool COUNT>
__global__ void __launch_bounds__(256) k_pair_clu(CluArgs a) {
  extern __shared__ float4 sft[];
  __shared__ float4 sD[4][64];
  __shared__ int2 sE[4][64];
  for (int k = threadIdx.x; k < a.npieces; k += blockDim.x) sft[k] = a.ftab[k];
  __syncthreads();
  constexpr int S = 64 / B;
  constexpr int NP2 = NS * NS;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int slot = lane / B;
  const int bead = (lane % B) + blockIdx.y * B;
  const bool bact = bead < a.P;
  const int bsel = bact ? bead : 0;
  const int P = a.P;
  // the molecule index is uniform over the wavefront: its loads are scalar
  const int wave = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const int nwave = (int)((gridDim.x * blockDim.x) >> 6);
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const Cell &c = a.cell;
  const float idr = a.idr, r0i = a.r0i;
  int mism = 0;
  unsigned npair = 0;
  // energies and virial in FP32 over the lane's rows (a few thousand terms),
  // FP64 across lanes and workgroups
  float cvl = 0.f, cvc = 0.f, c11 = 0.f, c22 = 0.f, c33 = 0.f, c12 = 0.f, c13 = 0.f, c23 = 0.f;
  for (int I = wave; I < a.nmol; I += nwave) {
    float vix[NS], viy[NS], viz[NS];
    float fix[NS], fiy[NS], fiz[NS];
    double xri[NS], yri[NS], zri[NS], umi[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const float4 v = a.v[(size_t)(I * NS + s) * P + bsel];
      vix[s] = v.x;
      viy[s] = v.y;
      viz[s] = v.z;
      fix[s] = fiy[s] = fiz[s] = 0.f;
      xri[s] = a.xr[I * NS + s];
      yri[s] = a.yr[I * NS + s];
      zri[s] = a.zr[I * NS + s];
      umi[s] = PIMD ? a.umax[I * NS + s] : 0.0;
    }
    const int cnt = a.mcnt[I];
    const int2 *row = a.mrow + (size_t)I * a.mcap;
    for (int k0 = 0; k0 < cnt; k0 += 64) {
      const int ks = k0 + lane;
      int J = 0, m = 0;
      float4 dd = make_float4(0.f, 0.f, 0.f, 0.f);
      if (ks < cnt) {
        const int2 e = row[ks];
        J = e.x;
        m = e.y;
        double s0[3], dx0, dy0, dz0;
        clu_image(c, xri[0] - a.xr[J * NS], yri[0] - a.yr[J * NS], zri[0] - a.zr[J * NS], s0, dx0, dy0, dz0);
        dd = make_float4((float)dx0, (float)dy0, (float)dz0, 0.f);
#pragma unroll
        for (int sa = 0; sa < NS; sa++) {
#pragma unroll
          for (int sb = 0; sb < NS; sb++) {
            const int bit = 1 << (sa * NS + sb);
            if (!(m & bit)) continue;
            const int j = J * NS + sb;
            double sab[3], ex, ey, ez;
            clu_image(c, xri[sa] - a.xr[j], yri[sa] - a.yr[j], zri[sa] - a.zr[j], sab, ex, ey, ez);
            mism += (sab[0] != s0[0]) | (sab[1] != s0[1]) | (sab[2] != s0[2]);
            const double lim = PIMD ? (a.cut + umi[sa] + a.umax[j]) * (1.0 + 1e-5) : a.cut;
            if (!(ex * ex + ey * ey + ez * ez < lim * lim)) m &= ~bit;
          }
        }
      }
      const bool keep = m != 0;
      const uint64_t mk = __ballot(keep);
      if (keep) {
        const int pos = __popcll(mk & below);
        sD[w][pos] = dd;
        sE[w][pos] = make_int2(J, m);
      }
      wave_sync();
      const int nk = __popcll(mk);
      float(&cfx)[NS] = fix, (&cfy)[NS] = fiy, (&cfz)[NS] = fiz;
      // the partner's positions for entry p (past the chunk's end: molecule 0,
      // unused), loaded on every path so the loads pipeline
      auto fetch = [&](float (&jx)[NS], float (&jy)[NS], float (&jz)[NS], int p) {
        const int e = sE[w][min(p, 63)].x;
        const size_t base = (size_t)(p < nk ? e : 0) * NS;
#pragma unroll
        for (int sb = 0; sb < NS; sb++) {
          const float4 v = a.v[(base + sb) * P + bsel];
          jx[sb] = v.x;
          jy[sb] = v.y;
          jz[sb] = v.z;
        }
      };
      auto entry = [&](int p, const float (&jx)[NS], const float (&jy)[NS], const float (&jz)[NS]) {
        const float4 D = sD[w][p];
        const int mm = sE[w][p].y;
        float gx[NS], gy[NS], gz[NS];
#pragma unroll
        for (int sa = 0; sa < NS; sa++) {
          gx[sa] = D.x + vix[sa];
          gy[sa] = D.y + viy[sa];
          gz[sa] = D.z + viz[sa];
        }
#pragma unroll
        for (int sa = 0; sa < NS; sa++) {
#pragma unroll
          for (int sb = 0; sb < NS; sb++) {
            const int q = sa * NS + sb;
            const bool hc = (a.cmask >> q) & 1, hl = (a.lmask >> q) & 1;  // uniform
            if (!hc && !hl) continue;  // no function for this site pair
            if (!(mm & (1 << q))) continue;
            const float dx = gx[sa] - jx[sb], dy = gy[sa] - jy[sb], dz = gz[sa] - jz[sb];
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (COUNT) {  // pairs inside the cutoff, for pmx_get_pair_count
              const float ct = (float)a.cut;
              npair += (bact && r2 < ct * ct) ? 1 : 0;
              continue;
            }
            const float rinv = __frsqrt_rn(r2);
            const float u = (r2 * rinv) * idr - r0i;
            const unsigned k = min(__float2uint_rz(u), (unsigned)CT_N);
            const float t = u - (float)k;
            float dv = 0.f;
            if (hc) {
              const float4 cc = sft[a.tcoul + k];
              const float qq = a.qq[q];
              cvc += qq * (cc.x + t * (cc.y + t * (cc.z + t * cc.w)));
              dv = qq * (cc.y + t * (2.f * cc.z + 3.f * t * cc.w));
            }
            if (hl) {
              const float4 cl = sft[a.toff_l[q] + k];
              cvl += cl.x + t * (cl.y + t * (cl.z + t * cl.w));
              dv += cl.y + t * (2.f * cl.z + 3.f * t * cl.w);
            }
            const float dvd = -(dv * idr) * rinv;  // -(1/r) dV/dr
            const float fx1 = dx * dvd, fy1 = dy * dvd, fz1 = dz * dvd;
            cfx[sa] += fx1;
            cfy[sa] += fy1;
            cfz[sa] += fz1;
            c11 += dx * fx1;
            c22 += dy * fy1;
            c33 += dz * fz1;
            c12 += dx * fy1;
            c13 += dx * fz1;
            c23 += dy * fz1;
          }
        }
      };
      float jxA[NS], jyA[NS], jzA[NS], jxB[NS], jyB[NS], jzB[NS];
      fetch(jxA, jyA, jzA, slot);
      int p = slot;
      for (; p + S < nk; p += 2 * S) {
        fetch(jxB, jyB, jzB, p + S);
        entry(p, jxA, jyA, jzA);
        fetch(jxA, jyA, jzA, p + 2 * S);
        entry(p + S, jxB, jyB, jzB);
      }
      if (p < nk) entry(p, jxA, jyA, jzA);
      wave_sync();
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
      double gx = fix[s], gy = fiy[s], gz = fiz[s];
      for (int off = 32; off >= B; off >>= 1) {
        gx += __shfl_xor(gx, off, 64);
        gy += __shfl_xor(gy, off, 64);
        gz += __shfl_xor(gz, off, 64);
      }
      if (!COUNT && slot == 0 && bact) {  // the first writer of this step's forces
        const size_t ib = (size_t)(I * NS + s) * P + bead;
        a.fx[ib] = gx;
        a.fy[ib] = gy;
        a.fz[ib] = gz;
      }
    }
  }
  for (int off = 32; off >= 1; off >>= 1) mism += __shfl_xor(mism, off, 64);
  if (lane == 0 && mism) atomicAdd(a.flag + 1, mism);
  if (COUNT) {
    unsigned long long n = npair;
    for (int off = 32; off >= 1; off >>= 1) n += __shfl_xor(n, off, 64);
    if (lane == 0) atomicAdd(a.count, n);
    return;
  }
  const double vals[PAIR_NV] = {cvl * 0.5, cvc * 0.5, c11 * 0.5, c22 * 0.5, c33 * 0.5,
                                c12 * 0.5, c13 * 0.5, c23 * 0.5, 0.0};
  block_partials<PAIR_NV>(vals, B, blockIdx.y * B, P, a.part, blockIdx.x, gridDim.x);
}

hipError_t launch_clu_pack(const CluArgs &a, hipStream_t st) {
  const int S = 64 / a.P;
  const size_t waves = ((size_t)a.N + S - 1) / S;
  hipLaunchKernelGGL(k_clu_pack, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_clu_rows(const CluArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(k_clu_rows, dim3(a.nmol), dim3(64), 0, st, a);
  return hipGetLastError();
}

template <int B, int NS, bool PIMD, bool COUNT>
static void launch_pair_clu_k(const CluArgs &a, dim3 grid, size_t lds, hipStream_t st) {
  (void)hipFuncSetAttribute((const void *)k_pair_clu<B, NS, PIMD, COUNT>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pair_clu<B, NS, PIMD, COUNT>), grid, dim3(256), lds, st, a);
}

template <int B, int NS>
static void launch_pair_clu_b(const CluArgs &a, dim3 grid, size_t lds, hipStream_t st) {
  if (a.count) {
    if (a.pimd) launch_pair_clu_k<B, NS, true, true>(a, grid, lds, st);
    else launch_pair_clu_k<B, NS, false, true>(a, grid, lds, st);
  } else {
    if (a.pimd) launch_pair_clu_k<B, NS, true, false>(a, grid, lds, st);
    else launch_pair_clu_k<B, NS, false, false>(a, grid, lds, st);
  }
}

template <int B>
static void launch_pair_clu_ns(const CluArgs &a, dim3 grid, size_t lds, hipStream_t st) {
  switch (a.ns) {
    case 1: launch_pair_clu_b<B, 1>(a, grid, lds, st); break;
    case 2: launch_pair_clu_b<B, 2>(a, grid, lds, st); break;
    case 3: launch_pair_clu_b<B, 3>(a, grid, lds, st); break;
    default: launch_pair_clu_b<B, 4>(a, grid, lds, st); break;
  }
}

hipError_t launch_pair_clu(const CluArgs &a, int gx, hipStream_t st) {
  int B = 1;
  while (B < a.P && B < 64) B <<= 1;
  const dim3 grid(gx, (a.P + B - 1) / B);
  const size_t lds = (size_t)a.npieces * sizeof(float4);
  switch (B) {
    case 1: launch_pair_clu_ns<1>(a, grid, lds, st); break;
    case 2: launch_pair_clu_ns<2>(a, grid, lds, st); break;
    case 4: launch_pair_clu_ns<4>(a, grid, lds, st); break;
    case 8: launch_pair_clu_ns<8>(a, grid, lds, st); break;
    case 16: launch_pair_clu_ns<16>(a, grid, lds, st); break;
    case 32: launch_pair_clu_ns<32>(a, grid, lds, st); break;
    default: launch_pair_clu_ns<64>(a, grid, lds, st); break;
  }
  return hipGetLastError();
}

hipError_t launch_mode_forces(const ModeArgs &a, hipStream_t st) {
  const dim3 grid(mode_blocks(a.N, a.P)), block(256);
  switch (a.P) {
    case 2: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mode_forces<2>), grid, block, 0, st, a); break;
    case 4: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mode_forces<4>), grid, block, 0, st, a); break;
    case 8: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mode_forces<8>), grid, block, 0, st, a); break;
    case 16: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mode_forces<16>), grid, block, 0, st, a); break;
    case 32: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mode_forces<32>), grid, block, 0, st, a); break;
    case 64: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mode_forces<64>), grid, block, 0, st, a); break;
    default: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mode_forces<0>), grid, block, 0, st, a); break;
  }
  return hipGetLastError();
}

int pair_grid_x(int N, int max_waves) {
  const int waves = N < max_waves ? N : max_waves;
  return (waves + 3) / 4;
}

// k_pair_mixed with the erfc term formed as TM says, with or without beads
// and RESPA weights
template <int B, int TM>
void launch_pair_mixed_tm(const PairArgs &a, dim3 grid, size_t lds, hipStream_t st) {
  const dim3 block(256);
  if (a.respa) {
    if (a.pimd) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pair_mixed<B, true, TM, true>), grid, block, lds, st, a);
    else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pair_mixed<B, false, TM, true>), grid, block, lds, st, a);
  } else {
    if (a.pimd) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pair_mixed<B, true, TM, false>), grid, block, lds, st, a);
    else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pair_mixed<B, false, TM, false>), grid, block, lds, st, a);
  }
}

template <int B>
void launch_pair_b(const PairArgs &a, int mode, bool full, dim3 grid, hipStream_t st) {
  const dim3 block(256);
  if (full) {
    if (mode == 0) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pair_staged_tab<B>), grid, block, 0, st, a);
    else if (mode == 1) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pair_staged_dir<B>), grid, block, 0, st, a);
    else if (a.ftab && !a.respa) {
      const size_t lds = (size_t)a.nftab * CT_N * sizeof(float4);
      if (a.pimd) {
        (void)hipFuncSetAttribute((const void *)k_pair_mixed<B, true, 1, false>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pair_mixed<B, true, 1, false>), grid, block, lds, st, a);
      } else {
        (void)hipFuncSetAttribute((const void *)k_pair_mixed<B, false, 1, false>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pair_mixed<B, false, 1, false>), grid, block, lds, st, a);
      }
    } else {
      const size_t lds = (size_t)(CT_N + 1) * sizeof(float4);
      if (a.cswtab) launch_pair_mixed_tm<B, 2>(a, grid, lds, st);
      else launch_pair_mixed_tm<B, 0>(a, grid, lds, st);
    }
  } else {
    if (mode == 0) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pair_half<B, 0>), grid, block, 0, st, a);
    else if (mode == 1) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pair_half<B, 1>), grid, block, 0, st, a);
    else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pair_half<B, 2>), grid, block, 0, st, a);
  }
}

hipError_t launch_pair(const PairArgs &a, int mode, bool full, hipStream_t st) {
  // beads per slot: the smallest power of two covering P, at most 64
  int B = 1;
  while (B < a.P && B < 64) B <<= 1;
  const int nbg = (a.P + B - 1) / B;
  const dim3 grid(a.gx, nbg);
  switch (B) {
    case 1: launch_pair_b<1>(a, mode, full, grid, st); break;
    case 2: launch_pair_b<2>(a, mode, full, grid, st); break;
    case 4: launch_pair_b<4>(a, mode, full, grid, st); break;
    case 8: launch_pair_b<8>(a, mode, full, grid, st); break;
    case 16: launch_pair_b<16>(a, mode, full, grid, st); break;
    case 32: launch_pair_b<32>(a, mode, full, grid, st); break;
    default: launch_pair_b<64>(a, mode, full, grid, st); break;
  }
  return hipGetLastError();
}

template <int B>
void launch_pair_grad_b(const PairArgs &a, int mode, dim3 grid, hipStream_t st) {
  if (mode == 0)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pair_grad<B, 0>), grid, dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pair_grad<B, 1>), grid, dim3(256), 0, st, a);
}

hipError_t launch_pair_grad(const PairArgs &a, int mode, hipStream_t st) {
  int B = 1;
  while (B < a.P && B < 64) B <<= 1;
  const int nbg = (a.P + B - 1) / B;
  const dim3 grid(a.gx, nbg);
  switch (B) {
    case 1: launch_pair_grad_b<1>(a, mode, grid, st); break;
    case 2: launch_pair_grad_b<2>(a, mode, grid, st); break;
    case 4: launch_pair_grad_b<4>(a, mode, grid, st); break;
    case 8: launch_pair_grad_b<8>(a, mode, grid, st); break;
    case 16: launch_pair_grad_b<16>(a, mode, grid, st); break;
    case 32: launch_pair_grad_b<32>(a, mode, grid, st); break;
    default: launch_pair_grad_b<64>(a, mode, grid, st); break;
  }
  return hipGetLastError();
}

}  // namespace pmx

namespace pmx {

// ---------------------------------------------------------------------------
// PIMD NVT integrator: int_NVT_pimd (pimd/int_NVT_pimd.c:36-380) with RESPA
// off.  A step: the chains and the first half kick (int_0_to_dt2_nvt_pimd,
// int_0_to_dt2_pimd.c:34-196), the drift in mode coordinates, the force call
// (energy_control_pimd, mode forces plus springs), the second half kick and
// the chains (int_dt2_to_dt_nvt_pimd, int_dt2_to_dt_pimd.c:31-172).

// One Nose-Hoover chain of apply_NHC_par (md/int_NVT.c:144-330) over its
// degrees of freedom, whose sum of m v^2 is kin; v, x are the chain's
// velocities and positions (1-based, len of them).  Returns the factor the
// degrees of freedom's velocities are scaled by.
__device__ __forceinline__ double nhc_chain(const NhcArgs &a, const double *mass, const double *gkt, double *v,
                                            double *x, double kin) {
#pragma clang fp contract(off)
  const int L = a.len;
  double f[8];
  double sc = 1.0;
  f[1] = (kin - gkt[1]) / mass[1];
  for (int ic = 1; ic <= L - 1; ic++) f[ic + 1] = (mass[ic] * v[ic] * v[ic] - gkt[ic + 1]) / mass[ic + 1];
  for (int ir = 1; ir <= a.nres; ir++) {
    for (int iy = 1; iy <= a.nyosh; iy++) {
      v[L] += f[L] * a.wdti4[iy];
      for (int ic = 1; ic <= L - 1; ic++) {
        const double aa = exp(-a.wdti8[iy] * v[L + 1 - ic]);
        v[L - ic] = v[L - ic] * aa * aa + a.wdti4[iy] * f[L - ic] * aa;
      }
      const double aa = exp(-a.wdti2[iy] * v[1]);
      sc *= aa;
      kin *= aa * aa;
      for (int ic = 1; ic <= L; ic++) x[ic] += v[ic] * a.wdti2[iy];
      f[1] = (kin - gkt[1]) / mass[1];
      for (int ic = 1; ic <= L - 1; ic++) {
        const double a2 = exp(-a.wdti8[iy] * v[ic + 1]);
        v[ic] = v[ic] * a2 * a2 + a.wdti4[iy] * f[ic] * a2;
        f[ic + 1] = (mass[ic] * v[ic] * v[ic] - gkt[ic + 1]) / mass[ic + 1];
      }
      v[L] += f[L] * a.wdti4[iy];
    }
  }
  return sc;
}

// Per-workgroup sums of m v^2 over the real sites' centroid-mode velocities
// (the centroid chain's kinetic term, apply_NHC_par section I).
__global__ void __launch_bounds__(256) k_nhc_ckin(NhcArgs a) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.N; i += gridDim.x * blockDim.x) {
    if (a.ghost[i]) continue;
    const size_t t = (size_t)i * a.P;
    s += a.mass[i] * (a.vx[t] * a.vx[t] + a.vy[t] * a.vy[t] + a.vz[t] * a.vz[t]);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = blockDim.x >> 1; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) a.part[blockIdx.x] = red[0];
}

// The centroid chain (one thread): the kinetic sum in a fixed order, the
// chain's evolution, and the scale factor in cls[2 len].
__global__ void k_nhc_class(NhcArgs a, int G) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double kin = 0.0;
  for (int g = 0; g < G; g++) kin += a.part[g];
  const int L = a.len;
  double v[8], x[8];
  for (int j = 1; j <= L; j++) {
    v[j] = a.cls[j - 1];
    x[j] = a.cls[L + j - 1];
  }
  const double sc = nhc_chain(a, a.cmass, a.cgkt, v, x, kin);
  for (int j = 1; j <= L; j++) {
    a.cls[j - 1] = v[j];
    a.cls[L + j - 1] = x[j];
  }
  a.cls[2 * L] = sc;
}

// One thread per (site, mode): phase 0 is the first half step (chains, then
// the kick: int_0_to_dt2_nvt_pimd sections 1-2), phase 1 the second (the
// kick, then the chains of modes 2..P; the centroid mode is scaled by
// k_nhc_scale0 once its chain has run).  The centroid mode's velocities are
// scaled by the centroid chain's factor in phase 0; every other mode's
// degree of freedom runs its own chain with the mode mass.
__global__ void __launch_bounds__(256) k_nhc_kick(NhcArgs a, int phase) {
#pragma clang fp contract(off)
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x, NP = (size_t)a.N * a.P;
  if (t >= NP) return;
  const int i = (int)(t / a.P), m = (int)(t - (size_t)i * a.P);
  const double mm = a.mass[i] * a.mfac[m];
  double *vv[3] = {a.vx, a.vy, a.vz};
  const double *ff[3] = {a.fx, a.fy, a.fz};
  const bool real = a.ghost[i] == 0;
  auto chains = [&]() {
    if (!real) return;
    if (m == 0) {
      if (phase == 0) {
        const double sc = a.cls[2 * a.len];
        for (int d = 0; d < 3; d++) vv[d][t] *= sc;
      }
      return;
    }
    for (int d = 0; d < 3; d++) {
      double v[8], x[8];
      for (int j = 1; j <= a.len; j++) {
        const size_t o = ((size_t)((j - 1) * 3 + d)) * NP + t;
        v[j] = a.vnhc[o];
        x[j] = a.xnhc[o];
      }
      const double vd = vv[d][t];
      const double sk = nhc_chain(a, a.bmass, a.bgkt, v, x, mm * vd * vd);
      vv[d][t] = vd * sk;
      for (int j = 1; j <= a.len; j++) {
        const size_t o = ((size_t)((j - 1) * 3 + d)) * NP + t;
        a.vnhc[o] = v[j];
        a.xnhc[o] = x[j];
      }
    }
  };
  if (phase == 0) chains();
  for (int d = 0; d < 3; d++) vv[d][t] += ff[d][t] * a.dt2 / mm;
  if (phase == 1) chains();
}

// the centroid mode's velocities scaled by its chain's factor (phase 1)
__global__ void __launch_bounds__(256) k_nhc_scale0(NhcArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N || a.ghost[i]) return;
  const size_t t = (size_t)i * a.P;
  const double sc = a.cls[2 * a.len];
  a.vx[t] *= sc;
  a.vy[t] *= sc;
  a.vz[t] *= sc;
}

// The drift of int_0_to_dt2_nvt_pimd section 3: the bead positions of one
// site to mode coordinates (control_pimd_trans_mode: transform_cnt.c:240-316
// normal modes, transform_stg.c:94-160 staging), u += v dt, and back
// (control_pimd_trans_pos: transform_cnt.c:320-410, transform_stg.c:168-226),
// with mode 1 kept as xmod (control_pimd_trans.c:175-181).  A group of P
// lanes holds one site's beads (64 / P sites per wavefront); the modes and
// the positions go through the wavefront's LDS rows.
__global__ void __launch_bounds__(256) k_mode_drift(NhcArgs a) {
#pragma clang fp contract(off)
  __shared__ double3 sp[4][64], su[4][64];
  __shared__ double2 stw[64];
  const int P = a.P;
  for (int k = threadIdx.x; k < P; k += blockDim.x) stw[k] = make_double2(a.cs[2 * k], a.cs[2 * k + 1]);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int S = 64 / P, g = lane / P, m = lane - g * P, lead = g * P;
  const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * S + g;
  const bool act = g < S && i < a.N;
  const size_t t = (size_t)(act ? i : 0) * P + m;
  double3 xb = make_double3(0.0, 0.0, 0.0);
  if (act) xb = make_double3(a.x[t], a.y[t], a.z[t]);
  sp[w][lane] = xb;
  __syncthreads();
  const double3 *px = sp[w] + lead;
  double3 u = make_double3(0.0, 0.0, 0.0);
  if (act) {
    if (a.stag) {
      const double3 x1 = px[0];
      if (m == 0) {
        u = x1;
      } else {
        const double r1 = ((double)m) / ((double)(m + 1)), r2 = 1.0 / ((double)(m + 1));
        const double3 nx = (m + 1 < P) ? px[m + 1] : x1;
        u = make_double3(xb.x - (r1 * nx.x + r2 * x1.x), xb.y - (r1 * nx.y + r2 * x1.y),
                         xb.z - (r1 * nx.z + r2 * x1.z));
      }
    } else {
      int k = 0, im = 0;
      if (m == P - 1) k = P / 2;
      else if (m > 0) {
        k = (m + 1) / 2;
        im = (m % 2 == 0);
      }
      int ph = 0;
      for (int b = 0; b < P; b++) {
        const double2 cs = stw[ph];
        const double wgt = im ? cs.y : cs.x;
        const double3 xv = px[b];
        u.x += xv.x * wgt;
        u.y += xv.y * wgt;
        u.z += xv.z * wgt;
        ph += k;
        if (ph >= P) ph -= P;
      }
      const double dP = (double)P;
      u.x /= dP;
      u.y /= dP;
      u.z /= dP;
    }
    u.x += a.vx[t] * a.dt;
    u.y += a.vy[t] * a.dt;
    u.z += a.vz[t] * a.dt;
  }
  su[w][lane] = u;
  __syncthreads();
  const double3 *pu = su[w] + lead;
  if (act) {
    double3 xn;
    if (a.stag) {
      // x_P = u_P + x_1, then x_ip = u_ip + (rat1 x_(ip+1) + x_1 rat2) for
      // ip = P-1 down to 2: each lane runs the recurrence down to its bead
      const double3 x1 = pu[0];
      if (m == 0) {
        xn = x1;
      } else {
        double3 xr = make_double3(pu[P - 1].x + x1.x, pu[P - 1].y + x1.y, pu[P - 1].z + x1.z);
        for (int j = P - 2; j >= m; j--) {
          const double r1 = ((double)j) / ((double)(j + 1)), r2 = 1.0 / ((double)(j + 1));
          const double3 uj = pu[j];
          xr = make_double3(uj.x + (r1 * xr.x + x1.x * r2), uj.y + (r1 * xr.y + x1.y * r2),
                            uj.z + (r1 * xr.z + x1.z * r2));
        }
        xn = xr;
      }
    } else {
      // x_b = u_0 + u_(P-1) cos(pi b) + 2 sum_k (u_(2k-1) cos + u_(2k) sin)(2 pi k b / P)
      xn = pu[0];
      const double sgn = (m & 1) ? -1.0 : 1.0;
      xn.x += pu[P - 1].x * sgn;
      xn.y += pu[P - 1].y * sgn;
      xn.z += pu[P - 1].z * sgn;
      int ph = 0;
      for (int k = 1; k < P / 2; k++) {
        ph += m;
        if (ph >= P) ph -= P;
        const double2 cs = stw[ph];
        const double3 re = pu[2 * k - 1], imv = pu[2 * k];
        xn.x += 2.0 * (re.x * cs.x + imv.x * cs.y);
        xn.y += 2.0 * (re.y * cs.x + imv.y * cs.y);
        xn.z += 2.0 * (re.z * cs.x + imv.z * cs.y);
      }
    }
    a.x[t] = xn.x;
    a.y[t] = xn.y;
    a.z[t] = xn.z;
    if (m == 0) {
      a.xm[i] = u.x;
      a.ym[i] = u.y;
      a.zm[i] = u.z;
    }
  }
}

// The extended-system energy's kinetic and chain parts, per workgroup:
// part[2 blk] = sum 1/2 m v^2 over every mode (get_tvten_pimd), part[2 blk
// + 1] = sum over the bead chains of 1/2 Q v^2 + gkt x.
__global__ void __launch_bounds__(256) k_pimd_energy(NhcArgs a) {
  __shared__ double red[2][256];
  const size_t NP = (size_t)a.N * a.P;
  double ke = 0.0, en = 0.0;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < NP; t += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(t / a.P), m = (int)(t - (size_t)i * a.P);
    const double mm = a.mass[i] * a.mfac[m];
    ke += 0.5 * mm * (a.vx[t] * a.vx[t] + a.vy[t] * a.vy[t] + a.vz[t] * a.vz[t]);
    if (m > 0 && !a.ghost[i])
      for (int d = 0; d < 3; d++)
        for (int j = 1; j <= a.len; j++) {
          const size_t o = ((size_t)((j - 1) * 3 + d)) * NP + t;
          en += 0.5 * a.bmass[j] * a.vnhc[o] * a.vnhc[o] + a.bgkt[j] * a.xnhc[o];
        }
  }
  red[0][threadIdx.x] = ke;
  red[1][threadIdx.x] = en;
  __syncthreads();
  for (int h = blockDim.x >> 1; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      red[0][threadIdx.x] += red[0][threadIdx.x + h];
      red[1][threadIdx.x] += red[1][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.part[2 * blockIdx.x] = red[0][0];
    a.part[2 * blockIdx.x + 1] = red[1][0];
  }
}

// out[0] kinetic, out[1] chain energy with the centroid chain (one thread)
__global__ void k_pimd_energy_sum(NhcArgs a, int G, double *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double ke = 0.0, en = 0.0;
  for (int g = 0; g < G; g++) {
    ke += a.part[2 * g];
    en += a.part[2 * g + 1];
  }
  const int L = a.len;
  for (int j = 1; j <= L; j++) en += 0.5 * a.cmass[j] * a.cls[j - 1] * a.cls[j - 1] + a.cgkt[j] * a.cls[L + j - 1];
  out[0] = ke;
  out[1] = en;
}

}  // namespace pmx
