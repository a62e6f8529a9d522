// Context, device memory and the C ABI (include/piny_amd.h).
//
// pmx_energy_control follows the stage order of energy_control
// (energy_control.c:49-74) and energy_control_pimd (energy_control_pimd.c:
// 91-273): initial (zero forces) -> real space (list check/build, pair
// forces) -> reciprocal space (self/background, PME, excluded-pair
// corrections) -> intramolecular -> PIMD /P and virial correction -> ghost
// force redistribution -> optional normal-mode forces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include <chrono>

#include "pmx_internal.h"
#include "pmx_kernels.h"

using namespace pmx;

struct pmx_ctx {
  std::string err;
  int device = 0;
  Setup s;
  pmx_params par{};
  hipStream_t st = nullptr;
  hipEvent_t ev[6] = {};
  double ms[5] = {0, 0, 0, 0, 0};
  std::vector<void *> allocs;
  // positions / forces, atom-major [N*P]
  double *x = nullptr, *y = nullptr, *z = nullptr, *stage = nullptr;
  double *fx = nullptr, *fy = nullptr, *fz = nullptr;
  double *fxt = nullptr, *fyt = nullptr, *fzt = nullptr;  // bead forces kept when modes are returned
  double *vx = nullptr, *vy = nullptr, *vz = nullptr;     // velocities, atom-major
  double *ke = nullptr;                                   // kinetic energy scratch
  bool last_modes = false, forces_valid = false;
  double *xm = nullptr, *ym = nullptr, *zm = nullptr;
  double *b0x = nullptr, *b0y = nullptr, *b0z = nullptr, *x0 = nullptr, *y0 = nullptr, *z0 = nullptr;
  double *maxes = nullptr;  // list check: per-workgroup maxima [2][LIST_CHECK_BLOCKS]
  int check_blocks = 0;     // workgroups of the last list check
  int *type = nullptr, *ichrg = nullptr, *ex_off = nullptr, *ex_num = nullptr, *ex_j = nullptr;
  double *q = nullptr, *mass = nullptr;
  double2 *tab_lj = nullptr, *tab_c = nullptr, *ecor_tab = nullptr;
  double *pair_pp = nullptr;   // [npair][PP_NUM]
  double *pair_pp_res = nullptr;  // the same with the RESPA cutoff as cutoff (short-range calls)
  pmx_respa respa{};           // RESPA settings (zeros: off)
  bool last_weighted = false;  // the last call's forces carried RESPA weights other than 1
  bool fxt_kept = false;       // fxt holds the last call's unweighted forces (a weighted full call's first pass)
  bool keep_fxt = false;       // this sequence must not overwrite fxt
  float4 *ctab = nullptr;      // mixed precision: erfc(alp r)/r as cubic pieces [CT_N + 1]
  float4 *ftab = nullptr;      // mixed precision: the switched pair functions [nftab][CT_N] (null if too many)
  float4 *cswtab = nullptr;    // mixed precision: the switched, shifted erfc term [CT_N + 1] when every type pair shares it
  int nftab = 0;
  std::vector<int2> ftab_off;  // per type pair: (erfc table, LJ table or -1), offsets in pieces
  float ct_r0 = 0.f, ct_idr = 0.f;
  double4 *site = nullptr;     // [N] (xmod, q)
  float4 *disp = nullptr;      // mixed precision with beads: x - xmod in FP32 [N*P]
  double *umax = nullptr;      // the largest |x - xmod| of every site over its beads [N]
  int pair_mode = 0;       // 0 FP64 tables, 1 FP64 direct, 2 FP32 direct
  // molecule-cluster form of the mixed pair kernel (k_pair_clu)
  bool clu_ok = false;      // the box qualifies: one molecule template of <= 4 sites, one cutoff
  bool clu_enabled = false; // chosen (PMX_CLUSTER=1 or the autotuner): rows are built with the list
  bool clu_on = false;      // the current list has molecule rows and the image margin holds
  bool clu_used = false;    // the last force sequence ran the cluster kernel
  bool disp_packed = false; // this call's pre_sequence packed the bead displacements
  int clu_ns = 0, clu_mcap = 0, clu_npieces = 0;
  int2 *mrow = nullptr;
  int *mcnt = nullptr, *clu_flag = nullptr, *clu_pin = nullptr;
  float4 *clu_v = nullptr, *clu_tab = nullptr;
  int clu_cmask = 0, clu_lmask = 0, clu_tcoul = 0, clu_tl[16];
  float clu_qq[16];
  double clu_cut = 0.0;
  std::vector<float4> ft_host;  // the whole switched pair functions [nftab][CT_N]
  bool full_list = true;   // force list with both directions (no atomics)
  int cap = 0, *jfull = nullptr, *jforce = nullptr, *nfull = nullptr, *nforce = nullptr, *overflow = nullptr;
  int *jsplit = nullptr;  // the force list's second buffer (rows split by the cutoff)
  // link cells of the list build
  int *site_cell = nullptr, *site_pos = nullptr, *cell_cnt = nullptr, *cell_start = nullptr, *cell_sites = nullptr;
  bool list_cells = false;  // the last build used link cells
  bool ref_valid = false;   // jfull / nfull hold the reference's half list of the last build
  double *xm0 = nullptr, *ym0 = nullptr, *zm0 = nullptr;  // xmod at the last build
  PairTypes pt_build{};     // cut+skin at the last build
  double4 *cell_p = nullptr, *cell_m = nullptr;  // sites in cell order (link-cell build)
  int *cell_i = nullptr;
  int *b1 = nullptr, *b2 = nullptr, *btyp = nullptr, *a1 = nullptr, *a2 = nullptr, *a3 = nullptr,
      *atyp = nullptr, *e1 = nullptr, *e2 = nullptr;
  double *bc = nullptr, *beq = nullptr, *ac = nullptr, *as = nullptr;
  int nghost = 0, *gatom = nullptr, *gncomp = nullptr, *gcomp = nullptr;
  std::vector<int> ghost_h;  // the ghost sites (host copy)
  // molecules of k_intra_mol (nmol = 0: the topology does not split into
  // molecules of at most MOL_MAX consecutive sites; per-term kernels then)
  int nmol = 0, mol_max = 0, ntpl = 0;
  bool mol_regular = false;  // every molecule has mol_max sites, molecule m at sites m*mol_max ...
  int2 *mol_dev = nullptr;     // [nmol] (first site, template)
  MolTpl *tpl_dev = nullptr;   // [TPL_MAX]
  double *gcoef = nullptr;
  // PME
  int pme = 0, npass[3] = {0, 0, 0}, radix[3][FFT_MAX_PASSES] = {};
  size_t gstride = 0, sstride = 0;
  double *grid = nullptr, *infl = nullptr;
  double *frec = nullptr;  // k-space forces, bead-major [3][P][N] (zero for uncharged sites)
  double2 *spec = nullptr, *tw[3] = {nullptr, nullptr, nullptr};
  float2 *twf[3] = {nullptr, nullptr, nullptr};
  int pme_f32 = 0;  // grids, B-splines and FFT in FP32 (the mixed precision)
  // standard Ewald (ewald_pme_opt off): k list, structure factors
  int ewald = 0;
  double4 *ew_kv = nullptr, *ew_kg = nullptr;
  double2 *ew_sk = nullptr;
  // binned PME: tiling, per-item data, bins, halo boxes
  PmeArgs pa{};
  int *item_s = nullptr, *item_tile = nullptr, *item_pos = nullptr, *bin_cnt = nullptr, *bin_off = nullptr,
      *sorted = nullptr;
  double *item_f = nullptr, *hbuf = nullptr;
  // per-block partial sums of every stage, then accumulators [P][ACC_NUM] + ake
  double *part = nullptr;
  size_t part_n = 0;       // doubles of part
  bool no_alloc = false;   // inside pmx_energy_control_noalloc: lists may not grow
  size_t off_pair = 0, off_fft = 0, off_ecor = 0, off_bond = 0, off_bend = 0, off_fin = 0, off_mode = 0;
  int g_intra = 1024;
  double *acc = nullptr, *ake = nullptr;  // ake: the entry after acc's P * ACC_NUM
  std::vector<double> acc_h;
  // pinned host staging of the asynchronous readbacks: list-check maxima,
  // per-bead sums [P*ACC_NUM] followed by the spring energy
  double *mx_pin = nullptr;
  double *acc_pin = nullptr;
  // springs
  double *prekf = nullptr, *veig = nullptr, *cs = nullptr;
  // state
  bool have_pos = false, have_list = false, xmod_given = false;
  double spread = 0.0;
  int nver = 0;
  PairTypes pt{};
  Cell cell{};
  int nbond = 0, nbend = 0, nbond_type = 0, nbend_type = 0;
  // parameter gradients: LJ slot per type pair, ecor partners per site,
  // buffers (allocated by the first pmx_get_param_grad), and whether the
  // device state is that of a finished force call at the current positions
  // and parameters (with the k-space grid)
  std::vector<int> lj_slot;
  std::vector<double> bc_h;  // bond coefficients [nbond_type*7]
  int *ecs_off = nullptr, *ecs_num = nullptr, *ecs_j = nullptr;
  double *phi = nullptr, *gpart = nullptr, *gacc = nullptr, *gbond = nullptr, *gbend = nullptr;
  bool grad_ready = false;
  // bumped whenever an argument captured by a force-sequence graph changes
  // (list width and buffers, launch shapes)
  unsigned epoch = 0;
  // launch shapes (set at creation, changed by pmx_autotune): PME tile
  // caps, wavefronts of the pair kernels, list build (0 by size, 1 plain
  // scan, 2 link cells)
  int tile_cap[3] = {16, 16, 8};
  int pair_waves = 4096;
  int list_waves = 8;         // wavefronts per cell of the link-cell list build (4: 448 us, 8: 322 us, 16: slower)
  int list_mode = 0;
  hipEvent_t ready = nullptr;  // list stage done (batched calls)
  // PME on a second stream, concurrent with the pair kernel (the pair
  // kernel is bound by vector ALU issue, PME by memory and LDS)
  int overlap = 0;
  bool last_overlap = false;  // the last force sequence ran PME on sk
  bool pme_deferred = false;  // PME on sk still to be launched, by main_sequence after the pair kernel
  bool fuse_pack = false;     // k_list_check writes site, disp and umax (mixed precision, P a power of 2 <= 64)
  hipStream_t sk = nullptr;
  hipEvent_t kfork = nullptr, kjoin = nullptr;
  hipEvent_t mxev = nullptr;  // the list check's maxima are in mx_pin
  // PIMD NVT integrator: chain state (device layout of NhcArgs), mode mass
  // factors, ghost flags, partial sums and the energies' readback
  int nhc_len = 0;
  double *vnhc = nullptr, *xnhc = nullptr, *nhc_cls = nullptr, *mfac = nullptr, *nhc_part = nullptr;
  double *nhc_out = nullptr;
  int *ghostflag = nullptr;
};

struct pmx_batch {
  std::string err;
  int device = 0;
  std::vector<pmx_ctx *> ctx;
  hipStream_t cs = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  unsigned flags = 0;
  std::vector<unsigned> epochs;
};

// The static scratch of pmx_energy_control_noalloc: the per-workgroup
// partial sums of one force call (energies, virial, the k-space and bonded
// sums, the springs), which pmx_energy_control keeps in the context's own
// allocation.  One copy per device, held until the process ends; every call
// on a device shares it, and each call writes every entry it reads before
// reading it.
static __device__ __align__(256) double scratch[PMX_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex scratch_mu;
double *scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// The device address of the scratch array on the current device, taken once
// per device (the host-side name of the array is not a device address).
// Called by pmx_create with its device current.
int scratch_setup(int device) {
  if (device < 0 || device >= SCRATCH_MAX_DEVICES) return PMX_ERR_RANGE;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_ptr[device]) return PMX_OK;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return PMX_ERR_DEVICE;
  scratch_ptr[device] = (double *)p;
  return PMX_OK;
}

double *scratch_address(int device) {
  if (device < 0 || device >= SCRATCH_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(scratch_mu);
  return scratch_ptr[device];
}

void set_err(pmx_ctx *c, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  c->err = buf;
}

#define HIPCHK(expr)                                                               \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      set_err(c, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
      return PMX_ERR_DEVICE;                                                       \
    }                                                                              \
  } while (0)

template <typename T>
int dalloc(pmx_ctx *c, T **p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) {
    set_err(c, "hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
    return PMX_ERR_NOMEM;
  }
  c->allocs.push_back((void *)*p);
  return PMX_OK;
}

template <typename T>
int upload(pmx_ctx *c, T **p, const T *src, size_t n) {
  int rc = dalloc(c, p, n);
  if (rc) return rc;
  if (n) HIPCHK(hipMemcpy(*p, src, n * sizeof(T), hipMemcpyHostToDevice));
  return PMX_OK;
}

// Radices 8 (FP32 grids only: in FP64 its butterflies need more than 128
// VGPRs), 4, 2, 3, 5, in that order of preference, whose product is k; the
// number of passes, or -1 when k has another prime factor.
int fft_factor(int k, int *radix, bool radix8) {
  int n = 0;
  while (k > 1 && n < FFT_MAX_PASSES) {
    int r = (radix8 && k % 8 == 0) ? 8
            : (k % 4 == 0)         ? 4
            : (k % 2 == 0)         ? 2
            : (k % 3 == 0)         ? 3
            : (k % 5 == 0)         ? 5
                                   : 0;
    if (r == 0) return -1;
    radix[n++] = r;
    k /= r;
  }
  return k == 1 ? n : -1;
}

template <typename T>
void dfree(pmx_ctx *c, T *&p) {
  if (!p) return;
  (void)hipFree((void *)p);
  c->allocs.erase(std::remove(c->allocs.begin(), c->allocs.end(), (void *)p), c->allocs.end());
  p = nullptr;
}

unsigned blocks(size_t n, unsigned bs = 256) { return (unsigned)((n + bs - 1) / bs); }

int alloc_lists(pmx_ctx *c, int cap) {
  const size_t n = (size_t)c->s.N * cap;
  if (c->no_alloc) {
    set_err(c, "the Verlet list needs rows of %d entries (now %d), which takes an allocation: "
               "one call of pmx_energy_control grows them", cap, c->cap);
    return PMX_ERR_RANGE;
  }
  if (c->jfull) {
    for (int *p : {c->jfull, c->jforce, c->jsplit}) {
      (void)hipFree(p);
      c->allocs.erase(std::remove(c->allocs.begin(), c->allocs.end(), (void *)p), c->allocs.end());
    }
  }
  c->cap = cap;
  c->epoch++;
  int rc = dalloc(c, &c->jfull, n);
  if (!rc) rc = dalloc(c, &c->jforce, n);
  if (!rc) rc = dalloc(c, &c->jsplit, n);
  return rc;
}

// Initial ELL width: 1.3 x the mean full-neighbour count within the largest
// list radius, plus a margin, in multiples of 64.
int estimate_cap(const Setup &s, double rskin) {
  const double dens = (double)s.N / s.vol;
  const double nb = dens * 4.0 / 3.0 * M_PI * rskin * rskin * rskin;
  int cap = (int)(1.3 * nb) + 64;
  cap = ((cap + 63) / 64) * 64;
  return std::min(std::max(cap, 64), ((s.N + 63) / 64) * 64);
}

constexpr int MAX_LIST_CELLS = 1 << 18;

#define TRY_RC(x)      \
  do {                 \
    int rc_ = (x);     \
    if (rc_) return rc_; \
  } while (0)

// Chooses the PME tile (the largest divisors of the grid below the caps
// whose halo box fits the LDS of one workgroup), allocates the binning
// buffers and
// fills the constant part of the PME argument block.
int setup_pme_tiles(pmx_ctx *c) {
  Setup &s = c->s;
  PmeArgs &a = c->pa;
  const int n = c->par.n_interp;
  if (n < 4 || n > MAX_INTERP) {
    set_err(c, "PME interpolation order %d: the device path supports orders 4..%d", n, MAX_INTERP);
    return PMX_ERR_NOTIMPL;
  }
  for (int d = 0; d < 3; d++) {
    if (s.ngrid[d] < n) {
      set_err(c, "PME grid %d is smaller than the interpolation order %d", s.ngrid[d], n);
      return PMX_ERR_RANGE;
    }
  }
  a.N = s.N;
  a.P = s.P;
  a.nchrg = (int)s.ichrg.size();
  a.n = n;
  a.f32 = c->pme_f32;
  // tiles divide the grid, so every tile is whole and a support of n points
  // starting in a tile reaches at most ceil((n-1)/T) tiles beyond it
  int cap[3] = {c->tile_cap[0], c->tile_cap[1], c->tile_cap[2]};
  auto divisor_below = [](int K, int m) {
    int t = std::max(1, std::min(m, K));
    while (K % t) t--;
    return t;
  };
  int T[3];
  for (;;) {
    for (int d = 0; d < 3; d++) T[d] = divisor_below(s.ngrid[d], cap[d]);
    for (int d = 0; d < 3; d++) {
      a.K[d] = s.ngrid[d];
      a.T[d] = T[d];
      a.nt[d] = (s.ngrid[d] + T[d] - 1) / T[d];
      a.H[d] = T[d] + n - 1;
    }
    a.hvol = (size_t)a.H[0] * a.H[1] * a.H[2];
    auto magic = [](int d) { return (unsigned)((1ull << 32) / (unsigned)d + 1ull); };
    a.mH0 = magic(a.H[0]);
    a.mH1 = magic(a.H[1]);
    a.mT0 = magic(a.T[0]);
    a.mT1 = magic(a.T[1]);
    bool ok = true;
    for (int d = 0; d < 3; d++) a.nbr[d] = 1 + (n - 1 + a.T[d] - 1) / a.T[d];
    ok = ok && (size_t)a.nt[0] * a.nt[1] * a.nt[2] * sizeof(int) <= 65536;  // LDS histogram of k_pme_bin
    ok = ok && pme_lds_bytes(a, true) <= FFT_LDS_MAX && pme_lds_bytes(a, false) <= FFT_LDS_MAX;
    // 32-bit element offsets into one bead's halo boxes
    ok = ok && (size_t)a.nt[0] * a.nt[1] * a.nt[2] * a.hvol < (1ull << 31);
    if (ok) break;
    int dm = 0;
    for (int d = 1; d < 3; d++)
      if (T[d] > T[dm]) dm = d;
    if (T[dm] == 1) {
      set_err(c, "PME order %d on a %dx%dx%d grid does not fit the tiled spread", n, s.ngrid[0], s.ngrid[1],
              s.ngrid[2]);
      return PMX_ERR_NOTIMPL;
    }
    cap[dm] = T[dm] - 1;
  }
  a.ntile = a.nt[0] * a.nt[1] * a.nt[2];
  a.gstride = c->gstride;
  const size_t nitem = (size_t)s.P * a.nchrg;
  if (!c->item_s) {
    TRY_RC(dalloc(c, &c->item_s, 3 * nitem));
    TRY_RC(dalloc(c, &c->item_f, 3 * nitem));
    TRY_RC(dalloc(c, &c->item_tile, nitem));
    TRY_RC(dalloc(c, &c->item_pos, nitem));
    TRY_RC(dalloc(c, &c->sorted, nitem));
  }
  dfree(c, c->bin_cnt);  // a second tiling (autotuner) replaces the tile-sized buffers
  dfree(c, c->bin_off);
  dfree(c, c->hbuf);
  TRY_RC(dalloc(c, &c->bin_cnt, (size_t)s.P * a.ntile));
  TRY_RC(dalloc(c, &c->bin_off, (size_t)s.P * a.ntile));
  TRY_RC(dalloc(c, &c->hbuf, (size_t)s.P * a.ntile * a.hvol));
  a.hbuf = c->hbuf;
  a.item_s = c->item_s;
  a.item_f = c->item_f;
  a.item_tile = c->item_tile;
  a.item_pos = c->item_pos;
  a.sorted = c->sorted;
  a.bin_cnt = c->bin_cnt;
  a.bin_off = c->bin_off;
  a.grid = c->grid;
  a.q = c->q;
  a.ichrg = c->ichrg;
  return PMX_OK;
}

// Per type pair constants of the force kernels (PP_* entries).  LJ type
// pairs get gradient slots in order, up to MAX_LJ_GRAD.
std::vector<double> pair_constants(pmx_ctx *c) {
  const Setup &s = c->s;
  std::vector<double> pp((size_t)s.npair * PP_NUM, 0.0);
  c->lj_slot.assign(s.npair, -1);
  int nslot = 0;
  for (int t = 0; t < s.npair; t++) {
    double *e = &pp[(size_t)t * PP_NUM];
    const bool lj = (s.label[t] == PMX_INTER_LJ);
    e[PP_CUT] = s.cutoff[t];
    e[PP_CUT2] = s.cutoff[t] * s.cutoff[t];
    e[PP_CUTRES] = s.cutoff_res[t];
    e[PP_CULL] = s.cutoff[t];
    e[PP_RMIN] = s.rmin[t];
    e[PP_DRI] = s.dri[t];
    e[PP_DR] = s.dr[t];
    e[PP_EPS] = s.lj_eps[t];
    e[PP_SIG6] = s.lj_sig6[t];
    e[PP_VSLJ] = s.vshift_lj[t];
    e[PP_VSC] = s.vshift_c[t];
    e[PP_LJ] = lj ? 1.0 : 0.0;
    e[PP_C6] = 4.0 * s.lj_eps[t] * s.lj_sig6[t];
    e[PP_C12] = e[PP_C6] * s.lj_sig6[t];
    if (lj && c->par.ishift == PMX_SHIFT_SHIFT) {
      // the shift is the table's value at node nsplin - 2 (spline_fit.c:28-113)
      const double rs = s.dr[t] * (double)(s.nsplin - 5) + s.rmin[t];
      const double r2 = rs * rs, s6 = s.lj_sig6[t] / (r2 * r2 * r2);
      e[PP_VSLJ1] = 4.0 * s6 * (s6 - 1.0);
      e[PP_DSLJ1] = 4.0 * (12.0 * s6 * s6 - 6.0 * s6);
    }
    e[PP_GSLOT] = -1.0;
    if (lj && nslot < MAX_LJ_GRAD) {
      c->lj_slot[t] = nslot;
      e[PP_GSLOT] = (double)nslot++;
    }
  }
  return pp;
}

// The constants of the short-range RESPA calls: the RESPA cutoff as the
// call's cutoff (CULL, CUT2), so that the pair kernels cull there, as
// npol_shave_skin does for iget_full_inter = 0 (force_npol.c:1316-1327).
// The functions keep their own cutoff (their switch); past the RESPA cutoff
// the RESPA switch is 0, so whatever a kernel still evaluates there adds 0.
std::vector<double> pair_constants_res(const pmx_ctx *c, const std::vector<double> &pp) {
  std::vector<double> r = pp;
  for (int t = 0; t < c->s.npair; t++) {
    double *e = &r[(size_t)t * PP_NUM];
    e[PP_CULL] = c->s.cutoff_res[t];
    e[PP_CUT2] = c->s.cutoff_res[t] * c->s.cutoff_res[t];
  }
  return r;
}

// Molecules for k_intra_mol: the connected components of the bonds, bends,
// excluded pairs and ghost constructions.  Used when every component is a
// run of at most MOL_MAX consecutive sites and the components follow at most
// TPL_MAX templates (terms by local site index, charge products, ghost
// coefficients), as water and other small molecules do; otherwise nmol
// stays 0 and the per-term kernels run.  Called again when the charges
// change (the templates hold charge products).
int setup_molecules(pmx_ctx *c, const pmx_system *sys) {
  c->nmol = 0;
  c->mol_regular = false;
  if (const char *env = getenv("PMX_INTRA"))  // tuning override: "terms" keeps the per-term kernels
    if (!strcmp(env, "terms")) return PMX_OK;
  const Setup &s = c->s;
  const int N = s.N;
  std::vector<int> par(N);
  for (int i = 0; i < N; i++) par[i] = i;
  auto find = [&](int i) {
    while (par[i] != i) i = par[i] = par[par[i]];
    return i;
  };
  auto unite = [&](int i, int j) {
    i = find(i);
    j = find(j);
    if (i != j) par[std::max(i, j)] = std::min(i, j);
  };
  for (int k = 0; k < sys->nbond; k++) unite(sys->bond_j1[k], sys->bond_j2[k]);
  for (int k = 0; k < sys->nbend; k++) {
    unite(sys->bend_j1[k], sys->bend_j2[k]);
    unite(sys->bend_j3[k], sys->bend_j2[k]);
  }
  for (size_t k = 0; k < s.e1.size(); k++) unite(s.e1[k], s.e2[k]);
  for (int g = 0; g < sys->nghost; g++)
    for (int k = 0; k < sys->ghost_ncomp[g]; k++) unite(sys->ghost_atom[g], sys->ghost_comp[4 * g + k]);
  // components in order of their first site (the root); each must be a run
  std::vector<int> mol(N), lo;
  std::vector<MolTpl> mt;
  for (int i = 0; i < N; i++) {
    const int r = find(i);
    if (r == i) {
      mol[i] = (int)lo.size();
      lo.push_back(i);
      MolTpl t;
      memset(&t, 0, sizeof t);
      t.n = 1;
      mt.push_back(t);
    } else {
      const int m = mol[i] = mol[r];
      if (i != lo[m] + mt[m].n || mt[m].n == MOL_MAX) return PMX_OK;  // not consecutive, or too large
      mt[m].n++;
    }
  }
  // the terms of each molecule, in term order, by local index
  for (size_t k = 0; k < s.e1.size(); k++) {
    const int i = s.e1[k], j = s.e2[k], m = mol[i];
    MolTpl &t = mt[m];
    if (t.ne == TPL_MAXT) return PMX_OK;
    t.e[t.ne][0] = i - lo[m];
    t.e[t.ne][1] = j - lo[m];
    t.eq[t.ne++] = sys->q[i] * sys->q[j];
  }
  for (int k = 0; k < sys->nbond; k++) {
    const int i = sys->bond_j1[k], m = mol[i];
    MolTpl &t = mt[m];
    if (t.nb == TPL_MAXT) return PMX_OK;
    t.bd[t.nb][0] = i - lo[m];
    t.bd[t.nb][1] = sys->bond_j2[k] - lo[m];
    t.bd[t.nb++][2] = sys->bond_type[k];
  }
  for (int k = 0; k < sys->nbend; k++) {
    const int m = mol[sys->bend_j2[k]];
    MolTpl &t = mt[m];
    if (t.na == TPL_MAXT) return PMX_OK;
    t.an[t.na][0] = sys->bend_j1[k] - lo[m];
    t.an[t.na][1] = sys->bend_j2[k] - lo[m];
    t.an[t.na][2] = sys->bend_j3[k] - lo[m];
    t.an[t.na++][3] = sys->bend_type[k];
  }
  for (int g = 0; g < sys->nghost; g++) {
    const int m = mol[sys->ghost_atom[g]];
    MolTpl &t = mt[m];
    if (t.ng == MOL_MAX) return PMX_OK;
    t.gh[t.ng][0] = sys->ghost_atom[g] - lo[m];
    t.gh[t.ng][1] = sys->ghost_ncomp[g];
    for (int k = 0; k < sys->ghost_ncomp[g]; k++) {
      t.gh[t.ng][2 + k] = sys->ghost_comp[4 * g + k] - lo[m];
      t.gc[t.ng][k] = sys->ghost_coef[4 * g + k];
    }
    t.ng++;
  }
  // distinct templates (byte-wise; the structs are zero-filled)
  const int M = (int)lo.size();
  std::vector<MolTpl> tpl;
  std::vector<int2> md(M);
  int mmax = 0;
  for (int m = 0; m < M; m++) {
    int id = -1;
    for (size_t u = 0; u < tpl.size() && id < 0; u++)
      if (!memcmp(&tpl[u], &mt[m], sizeof(MolTpl))) id = (int)u;
    if (id < 0) {
      if ((int)tpl.size() == TPL_MAX) return PMX_OK;
      id = (int)tpl.size();
      tpl.push_back(mt[m]);
    }
    md[m] = make_int2(lo[m], id);
    mmax = std::max(mmax, mt[m].n);
  }
  if (!c->mol_dev) {
    int rc;
    if ((rc = dalloc(c, &c->mol_dev, (size_t)N)) || (rc = dalloc(c, &c->tpl_dev, (size_t)TPL_MAX))) return rc;
  }
  HIPCHK(hipMemcpy(c->mol_dev, md.data(), M * sizeof(int2), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->tpl_dev, tpl.data(), tpl.size() * sizeof(MolTpl), hipMemcpyHostToDevice));
  c->ntpl = (int)tpl.size();
  c->mol_max = mmax;
  c->mol_regular = true;
  for (int m = 0; m < M; m++) c->mol_regular = c->mol_regular && md[m].x == m * mmax && mt[m].n == mmax;
  c->mol_regular = c->mol_regular && (size_t)M * mmax == (size_t)N;
  c->nmol = M;
  return PMX_OK;
}

// The mixed-precision pair functions as FP32 cubic pieces on one grid in r
// (CT_N intervals of width dr from r0, built in FP64):
// * ctab: erfc(alp r)/r with PINY's polynomial (spline_fit.c:327-339), for
//   the direct FP32 evaluation (pair_direct_f32), which applies the switch
//   and shift itself;
// * ftab: every tabulated function whole, as control_inter_params.c:643-772
//   tabulates it: the erfc term per unit charge product and the LJ term, each
//   times the switch of width rheal below the type pair's cutoff (swit) and
//   minus its shift (shift_inter_pe = shift).  Type pairs with the same
//   function share one table; ftab_off[t] = (erfc table, LJ table or -1) in
//   units of CT_N pieces.  The grid ends at the largest cutoff with
//   cutoff - rheal a node, so the switch's kink (its second derivative jumps
//   there) falls on a node when all type pairs share one cutoff.
// Each piece is the cubic Hermite interpolant of the function and its r
// derivative at the two nodes: c0 + c1 t + c2 t^2 + c3 t^3, t in [0, 1].
// Interpolation error of the erfc term: 3e-6 relative at the spline start
// (min_dist 0.1 A), 6e-9 from 1 Bohr on (1024 pieces over 17 Bohr).  Below
// the start the first piece extrapolates, as PINY's table does below its
// rmin (force_npol.c:742-751).
int build_pair_tables(pmx_ctx *c) {
  const Setup &s = c->s;
  double rend = 0.0, rlo = 1e300;
  for (int t = 0; t < s.npair; t++) {
    rend = std::max(rend, s.cutoff[t]);
    if (s.rmin[t] > 0.0) rlo = std::min(rlo, s.rmin[t]);
  }
  if (!(rlo < rend)) rlo = 0.5;
  const bool swit = (c->par.ishift == PMX_SHIFT_SWIT) && c->par.rheal > 0.0;
  // the grid reaches down to PINY's spline start (the smallest min_dist)
  double dr = (rend - rlo) / CT_N;
  if (swit && c->par.rheal >= dr)  // rheal a whole number of intervals, the start kept at or below rlo
    dr = c->par.rheal / std::floor(c->par.rheal / dr);
  const double r0 = rend - CT_N * dr, alp = s.alp_ewd;
  auto erfc_r = [&](double r, double &v, double &d) {  // erfc(alp r)/r and its r derivative
    const double x = alp * r, tt = 1.0 / (1.0 + kErfP * x), ee = exp(-x * x);
    double g = kErfE[9], dg = 9.0 * kErfE[9];
    for (int k = 8; k >= 1; k--) {
      g = g * tt + kErfE[k];
      dg = dg * tt + k * kErfE[k];
    }
    g = g * tt * ee;
    dg = dg * tt * tt * ee * kErfP * alp + 2.0 * alp * alp * g * r;  // -d g / dr
    v = g / r;
    d = -(g / (r * r) + dg / r);
  };
  // the switch of type pair t (spline_fit.c:65-73) and its r derivative
  auto sw = [&](int t, double r, double &w, double &dw) {
    w = 1.0;
    dw = 0.0;
    if (!swit) return;
    const double rs = r - s.cutoff[t] + c->par.rheal;
    if (rs <= 0.0) return;
    const double rp = std::min(rs / c->par.rheal, 1.0);
    w = 1.0 + rp * rp * (2.0 * rp - 3.0);
    dw = rs < c->par.rheal ? 6.0 * rp * (rp - 1.0) / c->par.rheal : 0.0;
  };
  auto pieces = [&](auto fn, std::vector<float4> &out) {
    double v0, d0;
    fn(r0, v0, d0);
    for (int k = 0; k < CT_N; k++) {
      double v1, d1;
      fn(r0 + dr * (k + 1), v1, d1);
      out.push_back(make_float4((float)v0, (float)(dr * d0), (float)(3.0 * (v1 - v0) - dr * (2.0 * d0 + d1)),
                                (float)(2.0 * (v0 - v1) + dr * (d0 + d1))));
      v0 = v1;
      d0 = d1;
    }
  };
  std::vector<float4> ct;
  pieces(erfc_r, ct);
  // whole functions, deduplicated by their parameters
  std::vector<float4> ft;
  std::vector<std::array<double, 5>> keys;
  auto table_of = [&](const std::array<double, 5> &key, auto fn) {
    for (size_t k = 0; k < keys.size(); k++)
      if (keys[k] == key) return (int)k;
    keys.push_back(key);
    pieces(fn, ft);
    return (int)keys.size() - 1;
  };
  c->ftab_off.assign(s.npair, make_int2(-1, -1));
  for (int t = 0; t < s.npair; t++) {
    c->ftab_off[t].x = table_of({0.0, s.cutoff[t], s.vshift_c[t], 0.0, 0.0}, [&](double r, double &v, double &d) {
      double e, de, w, dw;
      erfc_r(r, e, de);
      sw(t, r, w, dw);
      v = e * w - s.vshift_c[t];
      d = de * w + e * dw;
    });
    if (s.label[t] == PMX_INTER_LJ && s.lj_eps[t] != 0.0) {
      c->ftab_off[t].y =
          table_of({1.0, s.cutoff[t], s.vshift_lj[t], s.lj_eps[t], s.lj_sig6[t]}, [&](double r, double &v, double &d) {
            const double r2i = 1.0 / (r * r), s6 = s.lj_sig6[t] * r2i * r2i * r2i;
            const double e = 4.0 * s.lj_eps[t] * s6 * (s6 - 1.0);
            const double de = -24.0 * s.lj_eps[t] * s6 * (2.0 * s6 - 1.0) / r;
            double w, dw;
            sw(t, r, w, dw);
            v = e * w - s.vshift_lj[t];
            d = de * w + e * dw;
          });
    }
  }
  c->nftab = (int)keys.size();
  c->ft_host = ft;
  for (int t = 0; t < s.npair; t++) {
    c->ftab_off[t].x *= CT_N;
    if (c->ftab_off[t].y >= 0) c->ftab_off[t].y *= CT_N;
  }
  dfree(c, c->ctab);
  dfree(c, c->ftab);
  dfree(c, c->cswtab);
  // the erfc tables of the per-site mixed kernel hold one more piece, the
  // last repeated, so an r that rounds to the grid's end indexes a piece
  // without a clamp
  ct.push_back(ct.back());
  int rc = upload(c, &c->ctab, ct.data(), ct.size());
  // One erfc function for every type pair (one cutoff and shift): the mixed
  // pair kernel reads it switched and shifted from its table and evaluates
  // the switch only for the LJ term.
  bool one_c = s.npair > 0;
  for (int t = 1; t < s.npair; t++) one_c = one_c && c->ftab_off[t].x == c->ftab_off[0].x;
  if (!rc && one_c) {
    std::vector<float4> cs(ft.begin() + c->ftab_off[0].x, ft.begin() + c->ftab_off[0].x + CT_N);
    cs.push_back(cs.back());
    rc = upload(c, &c->cswtab, cs.data(), cs.size());
  }
  // The whole-function tables are built but used only on request
  // (PMX_MIXED_TAB=1): on the 4096-molecule, 32-bead box the pair kernel
  // takes 0.656 ms with them against 0.373 ms with LJ and the switch
  // evaluated directly beside the one erfc table.
  const char *env = getenv("PMX_MIXED_TAB");
  const bool use_ftab = env && atoi(env) == 1;
  if (!rc && use_ftab && c->nftab <= MAX_FTAB) rc = upload(c, &c->ftab, ft.data(), ft.size());
  if (rc) return rc;
  c->ct_r0 = (float)r0;
  c->ct_idr = (float)(1.0 / dr);
  c->epoch++;
  return PMX_OK;
}

// The molecule-cluster pair kernel (mixed precision), when the box
// qualifies: every molecule one template of ns <= CLU_MAX_NS consecutive
// sites with the same types and charges, and every site pair with a pair
// function at the grid's end as its cutoff (past the end the tables hold a
// zero piece, which stands for the cutoff).  Tables: the whole switched
// functions of build_pair_tables, CT_N + 1 pieces each.
int setup_cluster(pmx_ctx *c, const pmx_system *sys) {
  const bool was_on = c->clu_on;
  c->clu_ok = false;
  c->clu_on = false;
  c->epoch++;
  const Setup &s = c->s;
  const int N = s.N, P = s.P, ns = c->mol_max;
  if (c->pair_mode != 2 || !c->full_list || c->nmol <= 0 || c->ntpl != 1 || !c->mol_regular) return PMX_OK;
  if (ns < 1 || ns > CLU_MAX_NS || P > 64) return PMX_OK;
  if ((size_t)c->nftab * (CT_N + 1) * sizeof(float4) > 65 * 1024) return PMX_OK;
  for (int a = 0; a < ns; a++)
    for (int m = 1; m < c->nmol; m++)
      if (sys->atm_type[m * ns + a] != sys->atm_type[a] || sys->q[m * ns + a] != sys->q[a]) return PMX_OK;
  double rend = 0.0;
  for (int t = 0; t < s.npair; t++) rend = std::max(rend, s.cutoff[t]);
  c->clu_cmask = c->clu_lmask = 0;
  c->clu_tcoul = -1;
  c->clu_cut = rend;
  for (int q = 0; q < 16; q++) {
    c->clu_tl[q] = 0;
    c->clu_qq[q] = 0.f;
  }
  for (int a = 0; a < ns; a++)
    for (int b = 0; b < ns; b++) {
      const int q = a * ns + b, t = pair_index(s.ntype, sys->atm_type[a], sys->atm_type[b]);
      const double qq = sys->q[a] * sys->q[b];
      const int2 o = c->ftab_off[t];
      if (qq != 0.0) {
        const int tc = (o.x / CT_N) * (CT_N + 1);
        if (c->clu_tcoul >= 0 && tc != c->clu_tcoul) return PMX_OK;  // one erfc table for all
        c->clu_tcoul = tc;
        c->clu_cmask |= 1 << q;
      }
      if (o.y >= 0) {
        c->clu_tl[q] = (o.y / CT_N) * (CT_N + 1);
        c->clu_lmask |= 1 << q;
      }
      if (((c->clu_cmask | c->clu_lmask) >> q & 1) && s.cutoff[t] != rend) return PMX_OK;
      c->clu_qq[q] = (float)qq;
    }
  if (c->clu_tcoul < 0) c->clu_tcoul = 0;
  std::vector<float4> tab;
  for (int k = 0; k < c->nftab; k++) {
    tab.insert(tab.end(), c->ft_host.begin() + (size_t)k * CT_N, c->ft_host.begin() + (size_t)(k + 1) * CT_N);
    tab.push_back(make_float4(0.f, 0.f, 0.f, 0.f));
  }
  dfree(c, c->clu_tab);
  TRY_RC(upload(c, &c->clu_tab, tab.data(), tab.size()));
  c->clu_npieces = (int)tab.size();
  if (!c->mcnt) {
    TRY_RC(dalloc(c, &c->mcnt, (size_t)c->nmol));
    TRY_RC(dalloc(c, &c->clu_flag, 4));
    TRY_RC(dalloc(c, &c->clu_v, (size_t)N * P));
    if (P > 1 && !c->umax) TRY_RC(dalloc(c, &c->umax, (size_t)N));
    HIPCHK(hipHostMalloc((void **)&c->clu_pin, 4 * sizeof(int), hipHostMallocDefault));
    c->clu_pin[1] = 0;
  }
  c->clu_ns = ns;
  c->clu_ok = true;
  c->clu_on = was_on && c->mrow;  // a new force field keeps the force list, and so the rows
  return PMX_OK;
}

CluArgs clu_args(pmx_ctx *c) {
  const Setup &s = c->s;
  CluArgs a{};
  a.nmol = c->nmol;
  a.ns = c->clu_ns;
  a.N = s.N;
  a.P = s.P;
  a.pimd = s.P > 1;
  a.mcap = c->clu_mcap;
  a.mrow = c->mrow;
  a.mcnt = c->mcnt;
  a.jforce = c->jforce;
  a.nforce = c->nforce;
  a.cap = c->cap;
  a.mrow_w = c->mrow;
  a.mcnt_w = c->mcnt;
  a.flag = c->clu_flag;
  a.xr = s.P > 1 ? c->xm : c->x;
  a.yr = s.P > 1 ? c->ym : c->y;
  a.zr = s.P > 1 ? c->zm : c->z;
  a.x = c->x;
  a.y = c->y;
  a.z = c->z;
  a.vw = c->clu_v;
  a.v = c->clu_v;
  a.umax_w = c->umax;
  a.umax = c->umax;
  a.cell = c->cell;
  a.ftab = c->clu_tab;
  a.npieces = c->clu_npieces;
  a.idr = c->ct_idr;
  a.r0i = c->ct_r0 * c->ct_idr;
  a.cmask = c->clu_cmask;
  a.lmask = c->clu_lmask;
  a.tcoul = c->clu_tcoul;
  a.cut = c->clu_cut;
  for (int q = 0; q < 16; q++) {
    a.toff_l[q] = c->clu_tl[q];
    a.qq[q] = c->clu_qq[q];
  }
  a.fx = c->fx;
  a.fy = c->fy;
  a.fz = c->fz;
  a.part = c->part + c->off_pair;
  return a;
}

// Molecule rows of the force list just built, and whether the cluster kernel
// may use them.  A listed pair (a, b) lies within r_list at the build, so in
// fractional coordinates its separation is within r_list / w of an integer
// (w the cell's smallest width); the site-0 pair's differs from it by the
// two sites' offsets from their site 0, at most 2 ext / w, so both round to
// the same image when r_list + 2 ext < w / 2 (ext: the largest distance of
// a site from its molecule's site 0, reference positions).  Motion after the
// build is covered by the kernel's own check of every listed pair.
int build_clu_rows(pmx_ctx *c) {
  const bool was = c->clu_on;
  c->clu_on = false;
  for (int attempt = 0; attempt < 4; attempt++) {
    if (!c->mrow) {
      c->clu_mcap = std::max(64, ((c->cap + 63) / 64) * 64 / 2);
      TRY_RC(dalloc(c, &c->mrow, (size_t)c->nmol * c->clu_mcap));
    }
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->st, c->clu_flag, 0, (size_t)4);
    CluArgs a = clu_args(c);
    HIPCHK(launch_clu_rows(a, c->st));
    int fl[4];
    HIPCHK(hipMemcpyAsync(fl, c->clu_flag, sizeof fl, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (fl[0] > CLU_HASH) break;  // a molecule with too many partners: the per-site kernel
    if (fl[0] > c->clu_mcap) {
      if (c->no_alloc) {
        set_err(c, "the molecule rows need %d entries (now %d), which takes an allocation: "
                   "one call of pmx_energy_control grows them", fl[0], c->clu_mcap);
        return PMX_ERR_RANGE;
      }
      (void)hipFree(c->mrow);
      c->allocs.erase(std::remove(c->allocs.begin(), c->allocs.end(), (void *)c->mrow), c->allocs.end());
      c->mrow = nullptr;
      const int want = ((fl[0] * 5 / 4 + 63) / 64) * 64;
      c->clu_mcap = want;
      TRY_RC(dalloc(c, &c->mrow, (size_t)c->nmol * c->clu_mcap));
      c->epoch++;
      continue;
    }
    float ext;
    memcpy(&ext, &fl[2], sizeof ext);
    double rl = 0.0, wmin = 1e300;
    for (int t = 0; t < c->s.npair; t++) rl = std::max(rl, sqrt(c->pt.cutskin2[t]));
    for (int d = 0; d < 3; d++) {
      const double *hi = c->cell.hi;
      wmin = std::min(wmin, 1.0 / sqrt(hi[d] * hi[d] + hi[d + 3] * hi[d + 3] + hi[d + 6] * hi[d + 6]));
    }
    c->clu_on = rl + 2.0 * (double)ext < 0.5 * wmin * (1.0 - 1e-9);
    break;
  }
  if (c->clu_on != was) c->epoch++;
  return PMX_OK;
}

int build_device(pmx_ctx *c, const pmx_system *sys) {
  Setup &s = c->s;
  const int N = s.N, P = s.P;
  const size_t NP = (size_t)N * P;
  int rc = 0;
  if (N > ENTRY_MASK) {  // force-list entries hold the site index in 24 bits
    set_err(c, "%d sites exceed the %d sites of one context", N, ENTRY_MASK);
    return PMX_ERR_RANGE;
  }
  if (NP > (size_t)INT32_MAX) {  // kernels index (site, bead) values with 32-bit offsets
    set_err(c, "%d sites x %d beads exceed 2^31 - 1 site-bead values", N, P);
    return PMX_ERR_RANGE;
  }
#define TRY(x)        \
  do {                \
    rc = (x);         \
    if (rc) return rc; \
  } while (0)
  TRY(dalloc(c, &c->x, NP));
  TRY(dalloc(c, &c->y, NP));
  TRY(dalloc(c, &c->z, NP));
  TRY(dalloc(c, &c->stage, NP));
  TRY(dalloc(c, &c->fx, NP));
  TRY(dalloc(c, &c->fy, NP));
  TRY(dalloc(c, &c->fz, NP));
  TRY(dalloc(c, &c->fxt, NP));
  TRY(dalloc(c, &c->fyt, NP));
  TRY(dalloc(c, &c->fzt, NP));
  TRY(dalloc(c, &c->vx, NP));
  TRY(dalloc(c, &c->vy, NP));
  TRY(dalloc(c, &c->vz, NP));
  TRY(dalloc(c, &c->ke, 1));
  HIPCHK(hipMemset(c->vx, 0, NP * sizeof(double)));
  HIPCHK(hipMemset(c->vy, 0, NP * sizeof(double)));
  HIPCHK(hipMemset(c->vz, 0, NP * sizeof(double)));
  TRY(dalloc(c, &c->xm, N));
  TRY(dalloc(c, &c->ym, N));
  TRY(dalloc(c, &c->zm, N));
  TRY(dalloc(c, &c->b0x, N));
  TRY(dalloc(c, &c->b0y, N));
  TRY(dalloc(c, &c->b0z, N));
  TRY(dalloc(c, &c->x0, N));
  TRY(dalloc(c, &c->y0, N));
  TRY(dalloc(c, &c->z0, N));
  TRY(dalloc(c, &c->maxes, 2 * LIST_CHECK_BLOCKS));
  TRY(upload(c, &c->type, sys->atm_type, N));
  TRY(upload(c, &c->q, sys->q, N));
  TRY(upload(c, &c->mass, sys->mass, N));
  TRY(upload(c, &c->ichrg, s.ichrg.data(), s.ichrg.size()));
  // the list kernel tests exclusions in both directions
  TRY(upload(c, &c->ex_off, s.exs_off.data(), N));
  TRY(upload(c, &c->ex_num, s.exs_num.data(), N));
  TRY(upload(c, &c->ex_j, s.exs_j.data(), s.exs_j.size()));
  {
    const char *env = getenv("PMX_PAIR_LIST");  // tuning override: "half" or "full"
    if (env) c->full_list = strcmp(env, "half") != 0;
  }
  // tables as (V, -V'/r) pairs
  {
    const size_t nt = (size_t)s.npair * s.nsplin;
    std::vector<double2> lj(nt), cc(nt);
    for (size_t k = 0; k < nt; k++) {
      lj[k] = make_double2(s.tab[4 * k + 0], s.tab[4 * k + 1]);
      cc[k] = make_double2(s.tab[4 * k + 2], s.tab[4 * k + 3]);
    }
    TRY(upload(c, &c->tab_lj, lj.data(), nt));
    TRY(upload(c, &c->tab_c, cc.data(), nt));
    std::vector<double2> et(s.nsplin);
    for (int k = 0; k < s.nsplin; k++) et[k] = make_double2(s.ecor_tab[2 * k], s.ecor_tab[2 * k + 1]);
    TRY(upload(c, &c->ecor_tab, et.data(), et.size()));
  }
  {
    const std::vector<double> pp = pair_constants(c);
    TRY(upload(c, &c->pair_pp, pp.data(), pp.size()));
    const std::vector<double> ppr = pair_constants_res(c, pp);
    TRY(upload(c, &c->pair_pp_res, ppr.data(), ppr.size()));
    TRY(dalloc(c, &c->site, N));
    c->pair_mode = (c->par.precision == PMX_PREC_MIXED) ? 2 : (c->par.pair_eval == PMX_PAIR_DIRECT ? 1 : 0);
    if (c->pair_mode == 2 && P > 1) {
      if (NP * sizeof(float4) > (size_t)UINT32_MAX) {  // the pair kernel reads them at 32-bit byte offsets
        set_err(c, "%d sites x %d beads exceed the 2^28 bead displacements of the mixed precision", N, P);
        return PMX_ERR_RANGE;
      }
      TRY(dalloc(c, &c->disp, NP));
      c->fuse_pack = P <= 64 && (64 % P) == 0;  // the list check packs the sites and displacements
    }
    if (c->pair_mode == 2) TRY(build_pair_tables(c));
    if (P > 1) TRY(dalloc(c, &c->umax, (size_t)N));
  }
  TRY(dalloc(c, &c->site_cell, N));
  TRY(dalloc(c, &c->site_pos, N));
  TRY(dalloc(c, &c->cell_sites, N));
  TRY(dalloc(c, &c->cell_cnt, (size_t)MAX_LIST_CELLS));
  TRY(dalloc(c, &c->cell_start, (size_t)MAX_LIST_CELLS));
  TRY(dalloc(c, &c->nfull, N));
  TRY(dalloc(c, &c->nforce, N));
  TRY(dalloc(c, &c->overflow, 4));
  TRY(dalloc(c, &c->cell_p, N));
  TRY(dalloc(c, &c->cell_m, N));
  TRY(dalloc(c, &c->cell_i, N));
  TRY(dalloc(c, &c->xm0, N));
  TRY(dalloc(c, &c->ym0, N));
  TRY(dalloc(c, &c->zm0, N));
  TRY(alloc_lists(c, estimate_cap(s, s.cutoff_max + c->par.skin + 1.0 / BOHR * (P > 1 ? 1.0 : 0.0))));
  // bonded
  c->nbond = sys->nbond;
  c->nbend = sys->nbend;
  c->nbond_type = sys->nbond_type;
  c->nbend_type = sys->nbend_type;
  c->bc_h.assign(sys->bond_c, sys->bond_c + 7 * (size_t)sys->nbond_type);
  TRY(upload(c, &c->b1, sys->bond_j1, sys->nbond));
  TRY(upload(c, &c->b2, sys->bond_j2, sys->nbond));
  TRY(upload(c, &c->btyp, sys->bond_type, sys->nbond));
  TRY(upload(c, &c->bc, sys->bond_c, (size_t)7 * sys->nbond_type));
  TRY(upload(c, &c->beq, sys->bond_eq, sys->nbond_type));
  TRY(upload(c, &c->a1, sys->bend_j1, sys->nbend));
  TRY(upload(c, &c->a2, sys->bend_j2, sys->nbend));
  TRY(upload(c, &c->a3, sys->bend_j3, sys->nbend));
  TRY(upload(c, &c->atyp, sys->bend_type, sys->nbend));
  TRY(upload(c, &c->ac, sys->bend_c, (size_t)7 * sys->nbend_type));
  TRY(upload(c, &c->as, sys->bend_s, (size_t)7 * sys->nbend_type));
  TRY(upload(c, &c->e1, s.e1.data(), s.e1.size()));
  TRY(upload(c, &c->e2, s.e2.data(), s.e2.size()));
  {  // the Ewald-corrected partners of every site, both directions
    std::vector<int> off(N, 0), num(N, 0), js(2 * s.e1.size());
    for (size_t k = 0; k < s.e1.size(); k++) {
      num[s.e1[k]]++;
      num[s.e2[k]]++;
    }
    for (int i = 1; i < N; i++) off[i] = off[i - 1] + num[i - 1];
    std::vector<int> fill(N, 0);
    for (size_t k = 0; k < s.e1.size(); k++) {
      const int i = s.e1[k], j = s.e2[k];
      js[off[i] + fill[i]++] = j;
      js[off[j] + fill[j]++] = i;
    }
    TRY(upload(c, &c->ecs_off, off.data(), N));
    TRY(upload(c, &c->ecs_num, num.data(), N));
    TRY(upload(c, &c->ecs_j, js.data(), js.size()));
  }
  c->nghost = sys->nghost;
  c->ghost_h.assign(sys->ghost_atom, sys->ghost_atom + sys->nghost);
  TRY(upload(c, &c->gatom, sys->ghost_atom, sys->nghost));
  TRY(upload(c, &c->gncomp, sys->ghost_ncomp, sys->nghost));
  TRY(upload(c, &c->gcomp, sys->ghost_comp, (size_t)4 * sys->nghost));
  TRY(upload(c, &c->gcoef, sys->ghost_coef, (size_t)4 * sys->nghost));
  TRY(setup_molecules(c, sys));
  TRY(setup_cluster(c, sys));
  // PME grids, twiddles and the influence function on the sphere list
  c->pme = (!s.ichrg.empty() && c->par.pme_on);
  c->pme_f32 = (c->par.precision == PMX_PREC_MIXED) ? 1 : 0;
  c->ewald = (!s.ichrg.empty() && !c->par.pme_on);
  if (c->ewald) {  // ewald3d_recip's k terms (ewald3d.c:27-373)
    std::vector<double4> kv(s.nktot), kg(s.nktot);
    const double *hi = s.hmati;
    const double falp2 = 4.0 * s.alp_ewd * s.alp_ewd, tpi = 2.0 * M_PI, pivol = s.vol / (4.0 * M_PI);
    for (int k = 0; k < s.nktot; k++) {
      const double ga = tpi * ((double)s.ka[k]), gb = tpi * ((double)s.kb[k]), gc = tpi * ((double)s.kc[k]);
      const double xk = (ga * hi[1] + gb * hi[2] + gc * hi[3]);
      const double yk = (ga * hi[4] + gb * hi[5] + gc * hi[6]);
      const double zk = (ga * hi[7] + gb * hi[8] + gc * hi[9]);
      const double g2 = xk * xk + yk * yk + zk * zk;
      const double preg = exp((-g2 / falp2)) / (g2 * pivol);
      const double prep = -2.0 * preg * ((g2 / falp2) + 1.0) / g2;
      kv[k] = make_double4(ga, gb, gc, preg);
      kg[k] = make_double4(xk, yk, zk, prep);
    }
    TRY(upload(c, &c->ew_kv, kv.data(), kv.size()));
    TRY(upload(c, &c->ew_kg, kg.data(), kg.size()));
    TRY(dalloc(c, &c->ew_sk, (size_t)s.nktot * P));
  }
  if (c->pme) {
    const int Ka = s.ngrid[0], Kb = s.ngrid[1], Kc = s.ngrid[2], Ha = Ka / 2 + 1;
    for (int d = 0; d < 3; d++) {
      c->npass[d] = fft_factor(s.ngrid[d], c->radix[d], c->pme_f32 != 0);
      if (c->npass[d] < 0 || (s.ngrid[d] & 1)) {
        set_err(c, "PME grid %dx%dx%d: the device FFT takes even sizes with prime factors 2, 3, 5", Ka, Kb, Kc);
        return PMX_ERR_NOTIMPL;
      }
    }
    const size_t cx = c->pme_f32 ? 2 * sizeof(float) : 2 * sizeof(double);
    if (((size_t)Ka * Kb + Ka + Kb) * cx > FFT_LDS_DYN || ((size_t)Kc * Ha + Kc) * cx > FFT_LDS_DYN) {
      set_err(c, "PME grid %dx%dx%d: a plane and its twiddles exceed the %zu bytes of LDS an FFT workgroup may use",
              Ka, Kb, Kc, (size_t)FFT_LDS_DYN);
      return PMX_ERR_NOTIMPL;
    }
    c->gstride = (size_t)Ka * Kb * Kc;
    c->sstride = (size_t)Kc * Kb * Ha;
    TRY(dalloc(c, &c->grid, c->gstride * P));
    TRY(dalloc(c, &c->spec, c->sstride * P));
    for (int d = 0; d < 3; d++) {
      const int K = s.ngrid[d];
      std::vector<double2> t(K);
      for (int m = 0; m < K; m++) t[m] = make_double2(cos(2.0 * M_PI * m / K), -sin(2.0 * M_PI * m / K));
      TRY(upload(c, &c->tw[d], t.data(), K));
      std::vector<float2> tf(K);
      for (int m = 0; m < K; m++) tf[m] = make_float2((float)t[m].x, (float)t[m].y);
      TRY(upload(c, &c->twf[d], tf.data(), K));
    }
    std::vector<double> W(c->sstride, 0.0);
    const double *hi = s.hmati;
    const double falp2 = 4.0 * s.alp_ewd * s.alp_ewd, tpi = 2.0 * M_PI, pivol = s.vol / (4.0 * M_PI);
    for (int k = 0; k < s.nktot; k++) {
      const double ga = tpi * ((double)s.ka[k]), gb = tpi * ((double)s.kb[k]), gc = tpi * ((double)s.kc[k]);
      const double xk = (ga * hi[1] + gb * hi[2] + gc * hi[3]);
      const double yk = (ga * hi[4] + gb * hi[5] + gc * hi[6]);
      const double zk = (ga * hi[7] + gb * hi[8] + gc * hi[9]);
      const double g2 = xk * xk + yk * yk + zk * zk;
      const double preg = exp((-g2 / falp2)) / (g2 * pivol);
      const double w = preg * s.bweight[k];
      const int ia = s.ka[k], ib = ((s.kb[k] % Kb) + Kb) % Kb, ic = ((s.kc[k] % Kc) + Kc) % Kc;
      W[((size_t)ic * Kb + ib) * Ha + ia] = w;
      if (ia == 0) {
        const int jb = ((-s.kb[k] % Kb) + Kb) % Kb, jc = ((-s.kc[k] % Kc) + Kc) % Kc;
        W[((size_t)jc * Kb + jb) * Ha] = w;
      }
    }
    TRY(upload(c, &c->infl, W.data(), W.size()));
    TRY(setup_pme_tiles(c));
  }
  {
    const size_t pair = (size_t)pair_grid_x(N, PAIR_MAX_WAVES) * P * PAIR_NV;
    const size_t fft = (size_t)std::max(c->pme ? s.ngrid[1] : 0,
                                        c->ewald ? (s.nktot + EWALD_BLOCK - 1) / EWALD_BLOCK : 0) *
                       P * FFT_NV;
    const size_t intra = (size_t)c->g_intra * P * INTRA_NV;
    const size_t fin = (size_t)c->g_intra * P * FIN_NV;
    c->off_pair = 0;
    c->off_fft = pair;
    c->off_ecor = c->off_fft + fft;
    c->off_bond = c->off_ecor + intra;
    c->off_bend = c->off_bond + intra;
    c->off_fin = c->off_bend + intra;
    c->off_mode = c->off_fin + fin;
    c->part_n = c->off_mode + (P > 1 ? mode_blocks(N, P) : 1);
    TRY(dalloc(c, &c->part, c->part_n));
  }
  if (c->pme || c->ewald) {
    // the k-space kernels store the force of every (charge, bead); the
    // entries of the other sites stay zero
    TRY(dalloc(c, &c->frec, 3 * NP));
    HIPCHK(hipMemset(c->frec, 0, 3 * NP * sizeof(double)));
  }
  // the per-bead sums and, right after them, the spring energy: one readback
  TRY(dalloc(c, &c->acc, (size_t)P * ACC_NUM + 1));
  c->ake = c->acc + (size_t)P * ACC_NUM;
  c->acc_h.assign((size_t)P * ACC_NUM, 0.0);
  HIPCHK(hipHostMalloc((void **)&c->mx_pin, 2 * LIST_CHECK_BLOCKS * sizeof(double), hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void **)&c->acc_pin, ((size_t)P * ACC_NUM + 1) * sizeof(double), hipHostMallocDefault));
  TRY(upload(c, &c->prekf, s.prekf.data(), N));
  TRY(upload(c, &c->veig, s.veig.data(), s.veig.size()));
  {
    std::vector<double> csv(2 * P);
    for (int m = 0; m < P; m++) {
      csv[2 * m] = cos(2.0 * M_PI * m / P);
      csv[2 * m + 1] = sin(2.0 * M_PI * m / P);
    }
    TRY(upload(c, &c->cs, csv.data(), csv.size()));
  }
  // kernel parameter blocks
  for (int k = 0; k < 9; k++) {
    c->cell.h[k] = s.hmat[k + 1];
    c->cell.hi[k] = s.hmati[k + 1];
  }
  c->cell.cubic = s.cubic;
  c->pt.ntype = s.ntype;
  for (int t = 0; t < s.npair; t++) {
    c->pt.lj[t] = (s.label[t] == PMX_INTER_LJ);
    c->pt.rmin[t] = s.rmin[t];
    c->pt.dr[t] = s.dr[t];
    c->pt.dri[t] = s.dri[t];
    c->pt.cut[t] = s.cutoff[t];
    const double cs = s.cutoff[t] + c->par.skin;
    c->pt.cutskin2[t] = cs * cs;
  }
  return PMX_OK;
#undef TRY
}

// Link cells for the list build (make_lnk_lst.c / make_lnk_map.c): cells of
// fractional space at least r_list / H wide and at least 2H + 1 per axis,
// so that the (2H + 1)^3 cells around a site's cell hold every site with an
// image inside the list radius, each once.  H = 2 when the box allows it
// (a 5^3 stencil of half-width cells: about 3.7 candidates per list pair),
// else H = 1 (27 cells); 0 when the box is too small for either, or the
// plain scan is asked for (PMX_LIST=brute, or the autotuner's choice), and
// the ascending O(N^2) scan is used.
int choose_cells(const pmx_ctx *c, int nc[3]) {
  const char *env = getenv("PMX_LIST");  // tuning override: "brute" or "cells"
  if ((env && strcmp(env, "brute") == 0) || c->list_mode == 1) return 0;
  double rl = 0.0;
  for (int t = 0; t < c->s.npair; t++) rl = std::max(rl, sqrt(c->pt.cutskin2[t]));
  for (int H = 2; H >= 1; H--) {
    size_t ncell = 1;
    bool ok = true;
    for (int d = 0; d < 3; d++) {
      const double *hi = c->cell.hi;
      const double w = 1.0 / sqrt(hi[d] * hi[d] + hi[d + 3] * hi[d + 3] + hi[d + 6] * hi[d + 6]);
      nc[d] = std::min(64, (int)floor(w * H / rl));
      ok = ok && nc[d] >= 2 * H + 1;
      ncell *= (size_t)std::max(nc[d], 1);
    }
    if (ok && ncell <= (size_t)MAX_LIST_CELLS) return H;
  }
  return 0;
}

ListArgs list_args(pmx_ctx *c) {
  Setup &s = c->s;
  ListArgs a{};
  a.N = s.N;
  a.cap = c->cap;
  a.pimd = (s.P > 1);
  a.bx = c->b0x;
  a.by = c->b0y;
  a.bz = c->b0z;
  a.xm = c->xm;
  a.ym = c->ym;
  a.zm = c->zm;
  a.type = c->type;
  a.q = c->q;
  a.ex_off = c->ex_off;
  a.ex_num = c->ex_num;
  a.ex_j = c->ex_j;
  a.pt = c->pt;
  a.cell = c->cell;
  a.full = c->full_list ? 1 : 0;
  a.jfull = c->jfull;
  a.jforce = c->jforce;
  a.nfull = c->nfull;
  a.nforce = c->nforce;
  a.overflow = c->overflow;
  a.site_cell = c->site_cell;
  a.site_pos = c->site_pos;
  a.cell_cnt = c->cell_cnt;
  a.cell_start = c->cell_start;
  a.cell_sites = c->cell_sites;
  a.cell_p = c->cell_p;
  a.cell_m = c->cell_m;
  a.cell_i = c->cell_i;
  return a;
}

// (Re)build the Verlet list from the bead-1 copy; grows the ELL width when a
// row overflows.  The force list is built by link cells where the box
// allows; the reference's half list in its own order (pmx_get_verlet_list)
// then follows on demand from the build's saved state.
int build_list(pmx_ctx *c) {
  Setup &s = c->s;
  bool cells_failed = false;
  for (int attempt = 0; attempt < 8; attempt++) {
    int nc[3] = {1, 1, 1};
    const int H = cells_failed ? 0 : choose_cells(c, nc);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->st, c->overflow, 0, (size_t)4);
    ListArgs a = list_args(c);
    a.use_cells = 0;
    for (int d = 0; d < 3; d++) a.nc[d] = nc[d];
    if (H > 0) {
      const int ncell = nc[0] * nc[1] * nc[2];
      hipLaunchKernelGGL(k_fill_int, dim3(blocks(ncell)), dim3(256), 0, c->st, c->cell_cnt, 0, (size_t)ncell);
      hipLaunchKernelGGL(k_cell_count, dim3(blocks(s.N)), dim3(256), 0, c->st, a);
      hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(256), 0, c->st, ncell, c->cell_cnt, c->cell_start);
      hipLaunchKernelGGL(k_cell_scatter, dim3(blocks(s.N)), dim3(256), 0, c->st, a);
      hipLaunchKernelGGL(k_cell_order, dim3(ncell), dim3(256), 0, c->st, a);
      HIPCHK(hipGetLastError());
      HIPCHK(launch_list_cells(a, H, ncell, c->list_waves, c->st));
    } else {
      hipLaunchKernelGGL(k_list_build, dim3(blocks((size_t)s.N * 64)), dim3(256), 0, c->st, a);
      HIPCHK(hipGetLastError());
    }
    int ov[4];
    HIPCHK(hipMemcpyAsync(ov, c->overflow, sizeof ov, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (ov[2]) {  // a cell too full for the cell kernels: the plain scan
      cells_failed = true;
      continue;
    }
    if (ov[0] == 0) {
      // force-list rows: pairs inside the cutoff first
      hipLaunchKernelGGL(k_split_rows, dim3(blocks((size_t)s.N * 64)), dim3(256), 0, c->st, a, c->jsplit);
      HIPCHK(hipGetLastError());
      // copied back rather than swapped: captured force sequences keep
      // pointing at jforce
      HIPCHK(hipMemcpyAsync(c->jforce, c->jsplit, (size_t)s.N * c->cap * sizeof(int), hipMemcpyDeviceToDevice,
                            c->st));
      // the state the reference list is built from on demand
      HIPCHK(hipMemcpyAsync(c->xm0, c->xm, s.N * sizeof(double), hipMemcpyDeviceToDevice, c->st));
      HIPCHK(hipMemcpyAsync(c->ym0, c->ym, s.N * sizeof(double), hipMemcpyDeviceToDevice, c->st));
      HIPCHK(hipMemcpyAsync(c->zm0, c->zm, s.N * sizeof(double), hipMemcpyDeviceToDevice, c->st));
      c->pt_build = c->pt;
      c->ref_valid = (H == 0);  // the plain scan also wrote the reference's rows
      c->list_cells = H > 0;
      c->nver = ov[1];
      if (c->clu_ok && c->clu_enabled) {
        if (int rc = build_clu_rows(c)) return rc;
      } else if (c->clu_on) {
        c->clu_on = false;
        c->epoch++;
      }
      return PMX_OK;
    }
    int rc = alloc_lists(c, ((ov[0] * 5 / 4 + 63) / 64) * 64);
    if (rc) return rc;
  }
  set_err(c, "Verlet list does not fit after growing its rows");
  return PMX_ERR_RANGE;
}

// The reference's half list in its order (verlist_create.c:373-503), from
// the positions, xmod and cut+skin of the last build, by the ascending scan
// over j < i.
int build_ref_list(pmx_ctx *c) {
  ListArgs a = list_args(c);
  a.bx = c->x0;
  a.by = c->y0;
  a.bz = c->z0;
  a.xm = c->xm0;
  a.ym = c->ym0;
  a.zm = c->zm0;
  a.pt = c->pt_build;
  a.full = 0;
  a.ref_only = 1;
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->st, c->overflow, 0, (size_t)4);
  hipLaunchKernelGGL(k_list_build, dim3(blocks((size_t)c->s.N * 64)), dim3(256), 0, c->st, a);
  HIPCHK(hipGetLastError());
  int ov[4];
  HIPCHK(hipMemcpyAsync(ov, c->overflow, sizeof ov, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  if (ov[0] != 0 || ov[1] != c->nver) {
    set_err(c, "reference list rebuild: %d pairs against %d of the force list's build", ov[1], c->nver);
    return PMX_ERR_STATE;
  }
  c->ref_valid = true;
  return PMX_OK;
}

void add_stage(RedArgs &r, size_t off, int G, int nv, const std::vector<int> &slots) {
  RedStage &st = r.st[r.nstage++];
  st.off = off;
  st.G = G;
  st.nv = nv;
  st.wv = 1.0;
  for (int v = 0; v < nv; v++) st.slot[v] = slots[v];
}

IntraArgs intra_args(pmx_ctx *c, bool pimd) {
  Setup &s = c->s;
  IntraArgs a{};
  a.N = s.N;
  a.P = s.P;
  a.pimd = pimd && s.P > 1;
  a.intra_perds = c->par.intra_perds;
  a.nsplin = s.nsplin;
  a.x = c->x;
  a.y = c->y;
  a.z = c->z;
  a.xm = c->xm;
  a.ym = c->ym;
  a.zm = c->zm;
  a.q = c->q;
  a.cell = c->cell;
  a.fx = c->fx;
  a.fy = c->fy;
  a.fz = c->fz;
  a.nbond = c->nbond;
  a.b1 = c->b1;
  a.b2 = c->b2;
  a.btyp = c->btyp;
  a.bc = c->bc;
  a.beq = c->beq;
  a.nbend = c->nbend;
  a.a1 = c->a1;
  a.a2 = c->a2;
  a.a3 = c->a3;
  a.atyp = c->atyp;
  a.ac = c->ac;
  a.as = c->as;
  a.necor = (int)s.e1.size();
  a.e1 = c->e1;
  a.e2 = c->e2;
  a.ecor_tab = c->ecor_tab;
  a.e_rmin = s.e_rmin;
  a.e_dr = s.e_dr;
  a.e_dri = s.e_dri;
  a.w_ecor = a.w_rec = a.w_intra = 1.0;
  return a;
}

bool respa_on(const pmx_ctx *c) {
  const pmx_respa &r = c->respa;
  return r.respa_steps_lrf > 0 || r.respa_steps_torsion > 0 || r.respa_steps_intra > 0 || r.respa_steps_pimd > 0;
}

// A call's RESPA reading: iget_full_inter, iget_res_inter and
// iget_full_intra from the flags; iget_res_intra 1 when RESPA is on, as the
// RESPA integrators make every call (int_NVT_pimd.c:174-187,
// int_NVE_res.c:106-116), 0 otherwise; and the force weights of
// energy_control_initial.c:212-243.
struct RespaCall {
  bool on, full, res, full_intra, res_intra, weighted;
  double wt, wtr, wtrr;  // wght_ter, wght_ter_res, wght_tra_res
};
RespaCall respa_call(const pmx_ctx *c, unsigned flags) {
  RespaCall r{};
  r.on = respa_on(c);
  r.full = !(flags & PMX_EVAL_NO_FULL_INTER);
  r.res = (flags & PMX_EVAL_RES_INTER) != 0;
  r.full_intra = !(flags & PMX_EVAL_NO_FULL_INTRA);
  r.res_intra = r.on;
  const pmx_respa &p = c->respa;
  const int nter = std::max(p.respa_steps_lrf, 1), ntor = std::max(p.respa_steps_torsion, 1);
  const int ntra = std::max(p.respa_steps_intra, 1), npimd = std::max(p.respa_steps_pimd, 1);
  if (flags & PMX_EVAL_PIMD) {
    r.wt = (double)(npimd * ntra * ntor * nter);
    r.wtr = (double)(npimd * ntra * ntor);
    r.wtrr = (double)npimd;
  } else {
    r.wt = (double)(ntra * ntor * nter);
    r.wtr = (double)(ntra * ntor);
    r.wtrr = 1.0;
  }
  r.weighted = r.wt != 1.0 || r.wtr != 1.0 || r.wtrr != 1.0;
  return r;
}

// the molecule-cluster pair kernel runs for this call (not under RESPA)
bool clu_active(const pmx_ctx *c) { return c->clu_on && !respa_on(c); }

}  // namespace

namespace {

// How the list check sets xmod: 0 given (or no beads), 1 the bead centroid
// (normal modes), 2 bead 1 (staging): mode 1 of the transform, which
// control_pimd_trans_pos copies into xmod (control_pimd_trans.c:175-181).
int xmod_kind(const pmx_ctx *c) {
  if (c->s.P == 1 || c->xmod_given) return 0;
  return c->par.pimd_typ == PMX_PIMD_STAGING ? 2 : 1;
}

// The list stage of a force call, first half: centroid reference and list
// check (nbr_list_control.c:32-118) launched on st, with the per-workgroup
// maxima copied to pinned host memory (event mxev).  The caller waits for
// mxev before the second half, which takes the maxima over the workgroups.
int list_stage_launch(pmx_ctx *c, hipStream_t st, bool timing) {
  Setup &s = c->s;
  const int N = s.N, P = s.P;
  if (timing) HIPCHK(hipEventRecord(c->ev[0], st));
  const int sites_per_wave = P <= 64 ? 64 / P : 1;
  const size_t check_waves = ((size_t)N + sites_per_wave - 1) / sites_per_wave;
  const unsigned check_blocks = (unsigned)std::min<size_t>(LIST_CHECK_BLOCKS, blocks(check_waves * 64));
  c->check_blocks = (int)check_blocks;
  hipLaunchKernelGGL(k_list_check, dim3(check_blocks), dim3(256), 0, st, N, P, c->x, c->y, c->z, c->x0,
                     c->y0, c->z0,
                     c->b0x, c->b0y, c->b0z, c->maxes, xmod_kind(c), c->xm, c->ym, c->zm,
                     c->fuse_pack ? CheckPack{c->disp, c->umax, c->site, c->q} : CheckPack{});
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(c->mx_pin, c->maxes, 2 * check_blocks * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipEventRecord(c->mxev, st));
  return PMX_OK;
}

// Second half: the bead spread and cut+skin (get_cut_skin,
// nbr_list_control.c:249-283), the skin criterion, and the list build when
// it is due.
int list_stage_finish(pmx_ctx *c, unsigned flags, pmx_energies *out) {
  Setup &s = c->s;
  const int N = s.N, P = s.P;
  hipStream_t st = c->st;
  double dmax = 0.0, smax = 0.0;
  for (int k = 0; k < c->check_blocks; k++) {
    dmax = std::max(dmax, c->mx_pin[k]);
    smax = std::max(smax, c->mx_pin[c->check_blocks + k]);
  }
  const double spread_now = (P > 1) ? sqrt(smax) : 0.0;
  if (P > 1) {
    for (int t = 0; t < s.npair; t++) {
      const double cs = s.cutoff[t] + c->par.skin + spread_now;
      c->pt.cutskin2[t] = cs * cs;
    }
  }
  bool rebuild = !c->have_list || (flags & PMX_EVAL_FORCE_LIST);
  if (!rebuild) {
    const double dspread = (P > 1) ? std::max(0.0, spread_now - c->spread) : 0.0;
    rebuild = sqrt(dmax) > c->par.skin / 2.0 + dspread;
  }
  if (rebuild) {
    if (P > 1) c->spread = spread_now;
    HIPCHK(hipMemcpyAsync(c->x0, c->b0x, N * sizeof(double), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(c->y0, c->b0y, N * sizeof(double), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(c->z0, c->b0z, N * sizeof(double), hipMemcpyDeviceToDevice, st));
    int rc = build_list(c);
    if (rc) return rc;
    c->have_list = true;
  }
  out->list_rebuilt = rebuild;
  out->nver = c->nver;
  out->spread = c->spread;
  return PMX_OK;
}

// The PME chain of a force call (energy_control_inter_recip.c:119-355 with
// ewald3d_pme.c): binning, spread, FFT and convolution, gather into frec,
// launched on sk.
int launch_pme(pmx_ctx *c, hipStream_t sk) {
  Setup &s = c->s;
  const int P = s.P;
  const size_t NP = (size_t)s.N * P;
  const int nchrg = (int)s.ichrg.size();
  PmeArgs a = c->pa;
  a.x = c->x;
  a.y = c->y;
  a.z = c->z;
  a.cell = c->cell;
  a.rx = c->frec;
  a.ry = c->frec + NP;
  a.rz = c->frec + 2 * NP;
  const size_t nitem = (size_t)P * nchrg, nbin = (size_t)P * a.ntile;
  hipLaunchKernelGGL(k_fill_int, dim3(blocks(nbin)), dim3(256), 0, sk, c->bin_cnt, 0, nbin);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_pme_bin, dim3(blocks(nchrg), P), dim3(256), (size_t)a.ntile * sizeof(int), sk, a);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_pme_scan, dim3(P), dim3(256), 0, sk, a);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_pme_scatter, dim3(blocks(nitem)), dim3(256), 0, sk, a);
  HIPCHK(hipGetLastError());
  HIPCHK(launch_pme_spread(a, sk));
  FftArgs f{};
  f.P = P;
  f.f32 = c->pme_f32;
  for (int d = 0; d < 3; d++) {
    f.K[d] = s.ngrid[d];
    f.npass[d] = c->npass[d];
    for (int k = 0; k < c->npass[d]; k++) f.radix[d][k] = c->radix[d][k];
    f.tw[d] = c->tw[d];
    f.twf[d] = c->twf[d];
  }
  f.gstride = c->gstride;
  f.sstride = c->sstride;
  f.grid = c->grid;
  f.spec = c->spec;
  f.infl = c->infl;
  f.falp2 = 4.0 * s.alp_ewd * s.alp_ewd;
  f.cell = c->cell;
  f.part = c->part + c->off_fft;
  f.hbuf = a.hbuf;
  for (int d = 0; d < 3; d++) {
    f.T[d] = a.T[d];
    f.nt[d] = a.nt[d];
    f.H[d] = a.H[d];
    f.nbr[d] = a.nbr[d];
  }
  f.ntile = a.ntile;
  f.hvol = a.hvol;
  f.mKa = (unsigned)((1ull << 32) / (unsigned)s.ngrid[0] + 1ull);
  f.mT0 = a.mT0;
  f.mT1 = a.mT1;
  HIPCHK(launch_pme_fft(f, sk));
  HIPCHK(launch_pme_gather(a, sk));
  return PMX_OK;
}

// The device work of a force call that does not depend on the Verlet list,
// launched on st before the host reads the list check, so the device keeps
// working meanwhile: the force and kinetic clears, the site packing and
// bead displacements of the pair kernels, and the k space (PME or standard
// Ewald, energy_control_inter_recip.c:119-355), whose forces go to frec.
// With fork and c->overlap PME runs on the context's second stream, joined
// in main_sequence.  A captured sequence (batched calls, whose boxes are
// already parallel branches of the graph) stays on st: with a second fork
// inside the capture the batched call ended in a host-side segmentation
// fault on MI355X (ROCm 7).  No host synchronization and no memset, so it
// can be captured into a graph.
int pre_sequence(pmx_ctx *c, unsigned flags, hipStream_t st, bool timing, bool fork) {
  Setup &s = c->s;
  const int N = s.N, P = s.P;
  const size_t NP = (size_t)N * P;
  // --- initial (energy_control_initial.c:168-208)
  if (!c->full_list) {  // with the full list the pair kernel stores every force first
    for (double *f : {c->fx, c->fy, c->fz}) {
      hipLaunchKernelGGL(k_fill_double, dim3(blocks(NP)), dim3(256), 0, st, f, 0.0, NP);
      HIPCHK(hipGetLastError());
    }
  }
  // with fuse_pack the list check has packed the sites and displacements
  // (list_stage_launch precedes every force sequence; a weighted call's
  // second pass reuses the first pass's, the positions being the same)
  if (!c->fuse_pack) {
    hipLaunchKernelGGL(k_pack_sites, dim3(blocks(N)), dim3(256), 0, st, N, P > 1 ? 1 : 0, c->xm, c->ym, c->zm,
                       c->q, c->site);
    HIPCHK(hipGetLastError());
  }
  c->disp_packed = c->fuse_pack;
  if (c->disp && !clu_active(c) && !c->fuse_pack) {
    c->disp_packed = true;
    if (P <= 64) {
      const size_t waves = ((size_t)N + 64 / P - 1) / (64 / P);
      hipLaunchKernelGGL(k_pack_disp_umax, dim3(blocks(waves * 64)), dim3(256), 0, st, N, P, c->x, c->y, c->z, c->xm,
                         c->ym, c->zm, c->disp, c->umax);
    } else {
      hipLaunchKernelGGL(k_pack_disp, dim3(blocks(NP)), dim3(256), 0, st, N, P, c->x, c->y, c->z, c->xm, c->ym, c->zm,
                         c->disp);
      hipLaunchKernelGGL(k_site_umax, dim3(blocks(N)), dim3(256), 0, st, N, P, c->disp, c->umax);
    }
    HIPCHK(hipGetLastError());
  } else if (!c->disp && c->umax && c->full_list) {
    hipLaunchKernelGGL(k_site_umax64, dim3(blocks(N)), dim3(256), 0, st, N, P, c->x, c->y, c->z, c->xm, c->ym,
                       c->zm, c->umax);
    HIPCHK(hipGetLastError());
  }
  if (timing) HIPCHK(hipEventRecord(c->ev[1], st));
  const int nchrg = (int)s.ichrg.size();
  const bool full = respa_call(c, flags).full;  // the k space only in full calls
  const bool pme = c->pme && !(flags & PMX_EVAL_NO_RECIP) && full;
  const bool overlap = pme && c->overlap && fork;
  c->last_overlap = overlap;
  c->pme_deferred = overlap;
  if (overlap) {  // PME follows the pair kernel's launch (main_sequence), from this point of st
    HIPCHK(hipEventRecord(c->kfork, st));
    return PMX_OK;
  }
  if (pme)
    if (int rc = launch_pme(c, st)) return rc;
  if (c->ewald && !(flags & PMX_EVAL_NO_RECIP) && full) {
    EwaldArgs a{};
    a.N = N;
    a.P = P;
    a.nchrg = nchrg;
    a.nk = s.nktot;
    a.x = c->x;
    a.y = c->y;
    a.z = c->z;
    a.q = c->q;
    a.ichrg = c->ichrg;
    a.cell = c->cell;
    a.kv = c->ew_kv;
    a.kg = c->ew_kg;
    a.sk = c->ew_sk;
    a.rx = c->frec;
    a.ry = c->frec + NP;
    a.rz = c->frec + 2 * NP;
    a.part = c->part + c->off_fft;
    HIPCHK(launch_ewald(a, st));
  }
  if (timing) HIPCHK(hipEventRecord(c->ev[2], st));
  return PMX_OK;
}

// The device work after the list stage, launched on st: real space,
// excluded-pair corrections, intramolecular terms, the k-space forces
// added, PIMD finish, ghosts, modes and the reduction of the partial sums.
int main_sequence(pmx_ctx *c, unsigned flags, hipStream_t st, bool timing) {
  Setup &s = c->s;
  const int N = s.N, P = s.P;
  const size_t NP = (size_t)N * P;
  const bool pimd = (flags & PMX_EVAL_PIMD) != 0;
  const int nchrg = (int)s.ichrg.size();
  const RespaCall rsp = respa_call(c, flags);
  const bool kspace = (c->pme || c->ewald) && !(flags & PMX_EVAL_NO_RECIP) && rsp.full;
  RedArgs red{};
  if (timing) HIPCHK(hipEventRecord(c->ev[3], st));
  // --- real space (force_control -> force_verlst -> force_npol); none
  // when neither the full nor the short-range inter forces are asked for
  // (energy_control_inter_real.c:112)
  c->clu_used = clu_active(c);
  if (!rsp.full && !rsp.res) {
    if (c->full_list)  // the pair kernel would have been the forces' first writer
      for (double *f : {c->fx, c->fy, c->fz}) {
        hipLaunchKernelGGL(k_fill_double, dim3(blocks(NP)), dim3(256), 0, st, f, 0.0, NP);
        HIPCHK(hipGetLastError());
      }
  } else if (c->clu_used) {
    // the molecule-cluster form of the mixed pair kernel
    CluArgs a = clu_args(c);
    HIPCHK(launch_clu_pack(a, st));
    const int gx = pair_grid_x(c->nmol, c->pair_waves);
    HIPCHK(launch_pair_clu(a, gx, st));
    add_stage(red, c->off_pair, gx, PAIR_NV,
              {ACC_VVDW, ACC_VCOUL, ACC_P11, ACC_P22, ACC_P33, ACC_P12, ACC_P13, ACC_P23, ACC_NPAIR});
  } else {
    if (c->disp && !c->disp_packed) {  // packing skipped in pre_sequence for the cluster kernel
      if (P <= 64) {
        const size_t waves = ((size_t)N + 64 / P - 1) / (64 / P);
        hipLaunchKernelGGL(k_pack_disp_umax, dim3(blocks(waves * 64)), dim3(256), 0, st, N, P, c->x, c->y, c->z,
                           c->xm, c->ym, c->zm, c->disp, c->umax);
      } else {
        hipLaunchKernelGGL(k_pack_disp, dim3(blocks(NP)), dim3(256), 0, st, N, P, c->x, c->y, c->z, c->xm, c->ym,
                           c->zm, c->disp);
        hipLaunchKernelGGL(k_site_umax, dim3(blocks(N)), dim3(256), 0, st, N, P, c->disp, c->umax);
      }
      HIPCHK(hipGetLastError());
    }
    PairArgs a{};
    a.N = N;
    a.P = P;
    a.cap = c->cap;
    a.nsplin = s.nsplin;
    a.pimd = (P > 1);
    a.npair = s.npair;
    a.gx = pair_grid_x(N, c->pair_waves);
    a.alp = s.alp_ewd;
    a.p_alpha = kErfP * s.alp_ewd;
    a.two_alpha2 = 2.0 * s.alp_ewd * s.alp_ewd;
    a.spalp = 2.0 * s.alp_ewd / sqrt(M_PI);
    a.ctab = c->ctab;
    a.cswtab = c->cswtab;
    a.ct_r0 = c->ct_r0;
    a.ct_idr = c->ct_idr;
    a.ftab = c->ftab;
    a.nftab = c->nftab;
    for (size_t t = 0; t < c->ftab_off.size(); t++) a.ftoff[t] = c->ftab_off[t];
    a.rheal = c->par.rheal;
    a.irheal = c->par.rheal > 0.0 ? 1.0 / c->par.rheal : 0.0;
    a.swit = (c->par.ishift == PMX_SHIFT_SWIT);
    a.pp = c->pair_pp;
    if (rsp.on) {
      a.respa = rsp.res ? (rsp.full ? RESPA_BOTH : RESPA_RES) : RESPA_FULL;
      a.w_full = rsp.wt;
      a.w_res = rsp.wtr;
      if (rsp.res && !rsp.full) a.pp = c->pair_pp_res;  // cut at the RESPA cutoff
    }
    a.x = c->x;
    a.y = c->y;
    a.z = c->z;
    a.site = c->site;
    a.u = c->disp;
    a.umax = c->umax;
    a.type = c->type;
    a.jlist = c->jforce;
    a.cnt = c->nforce;
    a.tab_lj = c->tab_lj;
    a.tab_c = c->tab_c;
    a.cell = c->cell;
    a.fx = c->fx;
    a.fy = c->fy;
    a.fz = c->fz;
    a.part = c->part + c->off_pair;
    HIPCHK(launch_pair(a, c->pair_mode, c->full_list, st));
    add_stage(red, c->off_pair, pair_grid_x(N, c->pair_waves), PAIR_NV,
              {ACC_VVDW, ACC_VCOUL, ACC_P11, ACC_P22, ACC_P33, ACC_P12, ACC_P13, ACC_P23, ACC_NPAIR});
  }
  if (c->pme_deferred) {
    // PME on the second stream, launched once the pair kernel is queued:
    // the pair kernel, the longest kernel of the call, starts as soon as
    // the list decision allows, and the k space fills the CUs beside it
    c->pme_deferred = false;
    HIPCHK(hipStreamWaitEvent(c->sk, c->kfork, 0));
    if (int rc = launch_pme(c, c->sk)) return rc;
    if (timing) HIPCHK(hipEventRecord(c->ev[2], c->sk));
    HIPCHK(hipEventRecord(c->kjoin, c->sk));
  }
  if (timing) HIPCHK(hipEventRecord(c->ev[4], st));
  if (kspace) {
    const int G = c->pme ? s.ngrid[1] : (s.nktot + EWALD_BLOCK - 1) / EWALD_BLOCK;
    add_stage(red, c->off_fft, G, FFT_NV, {ACC_VRECIP, ACC_P11, ACC_P22, ACC_P33, ACC_P12, ACC_P13, ACC_P23});
    red.st[red.nstage - 1].wv = rsp.wt;  // ewald3d_pme.c:420-422
  }
  IntraArgs ia = intra_args(c, pimd);
  // RESPA weights: the excluded-pair terms and the k-space forces by
  // wght_ter in full calls only (energy_control_inter_recip.c:278-312), the
  // bonds and bends by wght_tra_res when the intramolecular terms are
  // asked for (energy_control_intra.c:124-230)
  ia.w_ecor = ia.w_rec = rsp.wt;
  ia.w_intra = (rsp.full_intra || rsp.res_intra) ? rsp.wtrr : 0.0;
  if (!rsp.full) ia.necor = 0;
  const int bsP = bead_block(P);
  auto igrid = [&](size_t items) { return (unsigned)std::min<size_t>(c->g_intra, (items + bsP - 1) / bsP); };
  const std::initializer_list<int> intra_slots = {0, ACC_P11, ACC_P22, ACC_P33, ACC_P12, ACC_P13, ACC_P23};
  auto intra_stage = [&](size_t off, unsigned G, int eslot) {
    std::vector<int> sl(intra_slots);
    sl[0] = eslot;
    add_stage(red, off, (int)G, INTRA_NV, sl);
    red.st[red.nstage - 1].wv = eslot == ACC_VECOR ? ia.w_ecor : ia.w_intra;
  };
  const double *frec = kspace ? c->frec : nullptr;
  if (c->nmol > 0) {
    // --- intramolecular (energy_control_intra.c:33-331), PIMD finish and
    // ghosts in one pass over the molecules
    if (c->last_overlap) HIPCHK(hipStreamWaitEvent(st, c->kjoin, 0));
    const unsigned G = igrid((size_t)c->nmol * P);
    ia.necor = nchrg > 0 ? ia.necor : 0;
    ia.nmol = c->nmol;
    ia.mol_max = c->mol_max;
    ia.ntpl = c->ntpl;
    ia.mol = c->mol_dev;
    ia.tpl = c->tpl_dev;
    ia.rec = frec;
    ia.part_ecor = c->part + c->off_ecor;
    ia.part_bond = c->part + c->off_bond;
    ia.part_bend = c->part + c->off_bend;
    ia.part_fin = c->part + c->off_fin;
    const size_t lds = (size_t)6 * c->mol_max * bsP * sizeof(double);  // at most 96 KB (MOL_MAX 8)
    if (lds > 65536)
      HIPCHK(hipFuncSetAttribute((const void *)k_intra_mol, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_intra_mol, dim3(G), dim3(bsP), lds, st, ia);
    HIPCHK(hipGetLastError());
    intra_stage(c->off_ecor, G, ACC_VECOR);
    intra_stage(c->off_bond, G, ACC_VBOND);
    intra_stage(c->off_bend, G, ACC_VBEND);
    if (pimd)
      add_stage(red, c->off_fin, (int)G, FIN_NV,
                {ACC_C11, ACC_C22, ACC_C33, ACC_C12, ACC_C21, ACC_C13, ACC_C31, ACC_C23, ACC_C32, ACC_AKVIR});
  } else {
    if (nchrg > 0 && ia.necor > 0) {
      const unsigned G = igrid((size_t)ia.necor * P);
      ia.part = c->part + c->off_ecor;
      hipLaunchKernelGGL(k_ecor, dim3(G), dim3(bsP), 0, st, ia);
      HIPCHK(hipGetLastError());
      intra_stage(c->off_ecor, G, ACC_VECOR);
    }
    // --- intramolecular (energy_control_intra.c:33-331)
    if (c->nbond) {
      const unsigned G = igrid((size_t)c->nbond * P);
      ia.part = c->part + c->off_bond;
      hipLaunchKernelGGL(k_bond, dim3(G), dim3(bsP), 0, st, ia);
      HIPCHK(hipGetLastError());
      intra_stage(c->off_bond, G, ACC_VBOND);
    }
    if (c->nbend) {
      const unsigned G = igrid((size_t)c->nbend * P);
      ia.part = c->part + c->off_bend;
      hipLaunchKernelGGL(k_bend, dim3(G), dim3(bsP), 0, st, ia);
      HIPCHK(hipGetLastError());
      intra_stage(c->off_bend, G, ACC_VBEND);
    }
    // --- PIMD: f/P and virial correction, then ghosts, then modes
    // the k-space forces join the others here, with f/P for PIMD
    if (c->last_overlap) HIPCHK(hipStreamWaitEvent(st, c->kjoin, 0));
    if (pimd) {
      FinishArgs a{};
      a.N = N;
      a.P = P;
      a.x = c->x;
      a.y = c->y;
      a.z = c->z;
      a.xm = c->xm;
      a.ym = c->ym;
      a.zm = c->zm;
      a.fx = c->fx;
      a.fy = c->fy;
      a.fz = c->fz;
      a.rec = frec;
      a.w_rec = ia.w_rec;
      a.part = c->part + c->off_fin;
      const unsigned G = igrid(NP);
      hipLaunchKernelGGL(k_pimd_finish, dim3(G), dim3(bsP), 0, st, a);
      HIPCHK(hipGetLastError());
      add_stage(red, c->off_fin, (int)G, FIN_NV,
                {ACC_C11, ACC_C22, ACC_C33, ACC_C12, ACC_C21, ACC_C13, ACC_C31, ACC_C23, ACC_C32, ACC_AKVIR});
    } else if (frec) {
      hipLaunchKernelGGL(k_add_rec, dim3(blocks(NP)), dim3(256), 0, st, N, P, frec, ia.w_rec, c->fx, c->fy,
                         c->fz);
      HIPCHK(hipGetLastError());
    }
    if (c->nghost) {
      hipLaunchKernelGGL(k_ghost_force, dim3(blocks((size_t)c->nghost * P)), dim3(256), 0, st, c->nghost, P,
                         c->gatom, c->gncomp, c->gcomp, c->gcoef, c->fx, c->fy, c->fz);
      HIPCHK(hipGetLastError());
    }
  }
  const bool modes = pimd && (flags & PMX_EVAL_MODE_FORCES) && P > 1;
  if (modes) {
    // the bead forces kept as fxt (energy_control_pimd.c:116-144), then mode
    // forces plus springs in place
    ModeArgs a{};
    a.N = N;
    a.P = P;
    a.stag = (c->par.pimd_typ == PMX_PIMD_STAGING) ? 1 : 0;
    a.x = c->x;
    a.y = c->y;
    a.z = c->z;
    a.fx = c->fx;
    a.fy = c->fy;
    a.fz = c->fz;
    a.fxt = c->keep_fxt ? nullptr : c->fxt;
    a.fyt = c->fyt;
    a.fzt = c->fzt;
    a.prekf = c->prekf;
    a.veig = c->veig;
    a.cs = c->cs;
    a.part = c->part + c->off_mode;
    HIPCHK(launch_mode_forces(a, st));
    HIPCHK(hipGetLastError());
  }
  red.P = P;
  red.nslot = ACC_NUM;
  red.part = c->part;
  red.acc = c->acc;
  red.spart = c->part + c->off_mode;  // the spring energy (0 without modes)
  red.sG = modes ? mode_blocks(N, P) : 0;
  red.sacc = c->ake;
  hipLaunchKernelGGL(k_reduce_partials, dim3((unsigned)P * ACC_NUM + 1), dim3(256), 0, st, red);
  HIPCHK(hipGetLastError());
  if (timing) HIPCHK(hipEventRecord(c->ev[5], st));
  return PMX_OK;
}


// The whole device sequence of a force call on st (batched calls capture it).
int force_sequence(pmx_ctx *c, unsigned flags, hipStream_t st) {
  if (int rc = pre_sequence(c, flags, st, false, false)) return rc;
  return main_sequence(c, flags, st, false);
}

// Readback of the per-bead sums into pinned host memory, launched on st.
int readback_launch(pmx_ctx *c, hipStream_t st) {
  if (c->clu_used) HIPCHK(hipMemcpyAsync(c->clu_pin + 1, c->clu_flag + 1, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(c->acc_pin, c->acc, (c->acc_h.size() + 1) * sizeof(double), hipMemcpyDeviceToHost, st));
  return PMX_OK;
}

// Energies and pressure tensors from the per-bead sums, after the readback
// has completed.
void assemble(pmx_ctx *c, unsigned flags, pmx_energies *out, bool timing) {
  Setup &s = c->s;
  const int N = s.N, P = s.P;
  const bool pimd = (flags & PMX_EVAL_PIMD) != 0;
  const bool modes = pimd && (flags & PMX_EVAL_MODE_FORCES) && P > 1;
  const int nchrg = (int)s.ichrg.size();
  memcpy(c->acc_h.data(), c->acc_pin, c->acc_h.size() * sizeof(double));
  const double ake = c->acc_pin[c->acc_h.size()];
  if (timing) {
    // events: 0 list check, 1 k space starts, 2 k space ends (on the second
    // stream with overlap), 3 after the list decision, 4 after the pair
    // kernel, 5 end
    auto el = [&](int i, int j) {
      float t = 0.f;
      (void)hipEventElapsedTime(&t, c->ev[i], c->ev[j]);
      return (double)t;
    };
    c->ms[0] = c->last_overlap ? el(0, 3) : el(0, 1) + el(2, 3);
    c->ms[1] = el(3, 4);
    c->ms[2] = el(1, 2);
    c->ms[3] = el(4, 5);
    c->ms[4] = el(0, 5);
  }
  // --- assemble (energy_control_inter_real.c:188-197, _recip.c:340-347, _intra.c)
  const double dP = (double)P;
  const RespaCall rsp = respa_call(c, flags);
  double sum[ACC_NUM] = {0};
  for (int b = 0; b < P; b++)
    for (int k = 0; k < ACC_NUM; k++) sum[k] += c->acc_h[(size_t)b * ACC_NUM + k];
  // energies the reference keeps: the intermolecular ones in full calls
  // (force_npol.c:243-256), the intramolecular ones when the intra or the
  // inter terms are full (energy_control_intra.c:284)
  const bool keep_e = rsp.full_intra || rsp.full;
  const bool intra = rsp.full_intra || rsp.res_intra;
  out->vvdw = rsp.full ? sum[ACC_VVDW] / dP : 0.0;
  out->vcoul = rsp.full ? sum[ACC_VCOUL] / dP : 0.0;
  out->vreal = out->vvdw + out->vcoul;
  out->vlong = rsp.full ? s.clong / s.vol : 0.0;
  out->vrecip = sum[ACC_VRECIP] / dP;
  out->vecor = sum[ACC_VECOR] / dP;
  out->vbond = keep_e && intra ? sum[ACC_VBOND] / dP : 0.0;
  out->vbend = keep_e && intra ? sum[ACC_VBEND] / dP : 0.0;
  double pv[10] = {0};
  pv[1] = sum[ACC_P11];
  pv[5] = sum[ACC_P22];
  pv[9] = sum[ACC_P33];
  pv[2] = pv[4] = sum[ACC_P12];
  pv[3] = pv[7] = sum[ACC_P13];
  pv[6] = pv[8] = sum[ACC_P23];
  {
    // long_range_corr with its RESPA terms (energy_control_inter_real.c:
    // 206-318): the bead, intramolecular and short-range steps' guesses and
    // the dispersion tail, each as the difference to the previous level
    // times that level's weight
    const double amult = (c->par.ishift == PMX_SHIFT_SWIT) ? 1.0 : 2.0, amult_bead = amult * dP;
    const double T = c->par.temperature > 0 ? c->par.temperature : 300.0;
    double pd_bead = 0.0, pd_tra, pd_ter, d = 0.0;
    if (P > 1 && rsp.wtrr != 1.0) {
      const double pkin = (double)(c->respa.nfree / 3) * T / (BOLTZ * s.vol);
      pd_bead = (c->respa.pext - pkin) * s.vol * dP;
      d += pd_bead;
    }
    if (rsp.res_intra && rsp.wt != 1.0) {
      pd_tra = c->respa.pten_inter_guess * s.vol * dP;
      d += (pd_tra - pd_bead) * rsp.wtrr;
    } else {
      pd_tra = pd_bead;
    }
    if (rsp.res) {
      pd_ter = (s.clong_res / s.vol) * amult_bead;
      d += (pd_ter - pd_tra) * rsp.wtr;
    } else {
      pd_ter = pd_tra;
    }
    if (rsp.full) d += ((s.clong / s.vol) * amult_bead - pd_ter) * rsp.wt;
    pv[1] += d;
    pv[5] += d;
    pv[9] += d;
  }
  out->vself = out->vbgr = 0.0;
  if (nchrg > 0 && (rsp.full_intra || !rsp.res_intra)) {
    // ewald3d_selfbgr (ewald3d_self_bgr.c:29-83) over all sites, its
    // pressure times wght_tra_res
    const double qq = s.qsq, qs = s.qsum;
    const double bgr = -0.5 * qs * qs * M_PI / (s.alp_ewd * s.alp_ewd * s.vol);
    if (keep_e) {
      out->vself = -qq * (s.alp_ewd) * (s.self_erf) / sqrt(M_PI);
      out->vbgr = bgr;
    }
    pv[1] += bgr * rsp.wtrr * dP;
    pv[5] += bgr * rsp.wtrr * dP;
    pv[9] += bgr * rsp.wtrr * dP;
  }
  out->vintert = rsp.full ? out->vreal + out->vlong + out->vrecip + out->vself + out->vbgr + out->vecor : 0.0;
  out->vintrat = out->vbond + out->vbend;
  if (pimd) {
    for (int k = 1; k <= 9; k++) pv[k] /= dP;
    if (P > 1) {
      pv[1] += sum[ACC_C11];
      pv[5] += sum[ACC_C22];
      pv[9] += sum[ACC_C33];
      pv[2] += sum[ACC_C12];
      pv[4] += sum[ACC_C21];
      pv[3] += sum[ACC_C13];
      pv[7] += sum[ACC_C31];
      pv[6] += sum[ACC_C23];
      pv[8] += sum[ACC_C32];
      const double T = c->par.temperature > 0 ? c->par.temperature : 300.0;
      const double beta = BOLTZ / T;
      // the virial estimator needs the unweighted forces (fxt,
      // pimd_estimators.c:259-303): under RESPA weights it is not formed
      out->pi_ke_vir = !(rsp.full_intra && rsp.full) ? 0.0
                       : rsp.weighted ? NAN
                                      : 0.5 * (3.0 * ((double)(N - c->nghost)) / beta + sum[ACC_AKVIR]);
    }
  }
  // pvten_tot: the unweighted tensor of full calls (zero in the others);
  // without weights it is pvten
  for (int k = 0; k < 9; k++) {
    out->pvten[k] = pv[k + 1];
    out->pvten_tot[k] = !rsp.full ? 0.0 : rsp.weighted ? NAN : pv[k + 1];
  }
  c->last_weighted = rsp.weighted || !rsp.full;
  if (modes) {
    const double T = c->par.temperature > 0 ? c->par.temperature : 300.0;
    const double tau = (BOLTZ / T) / dP;
    out->kin_harm = ake;
    out->pi_ke_prim = 1.5 * ((double)(N - c->nghost)) / tau - ake;
  }
  c->last_modes = modes;
  c->forces_valid = true;
  c->grad_ready = !(c->pme && (flags & PMX_EVAL_NO_RECIP));
}

// After a force call that ran the cluster kernel: true when a listed site
// pair's image differed from its molecule pair's (the margin of
// build_clu_rows makes this a safety net); the cluster kernel is then off
// until the next list build and the caller repeats the call.
bool clu_recheck(pmx_ctx *c) {
  if (!c->clu_used || c->clu_pin[1] == 0) return false;
  c->clu_pin[1] = 0;
  c->clu_on = false;
  c->clu_used = false;
  c->epoch++;
  return true;
}

int check_call(pmx_ctx *c, unsigned flags, pmx_energies *out, const char *who) {
  if (!c || !out) return PMX_ERR_ARG;
  if (!c->have_pos) {
    set_err(c, "%s: no positions set", who);
    return PMX_ERR_STATE;
  }
  if (c->s.P > 1 && !(flags & PMX_EVAL_PIMD)) {
    set_err(c, "%s: more than one bead needs PMX_EVAL_PIMD", who);
    return PMX_ERR_ARG;
  }
  return PMX_OK;
}

}  // namespace

extern "C" {

int pmx_abi_version(void) { return PMX_ABI_VERSION; }

int pmx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

pmx_ctx *pmx_create(const pmx_system *sys, const pmx_params *par, const double *ix, const double *iy,
                    const double *iz, int device, char *err, size_t errlen) {
  auto fail = [&](const std::string &m) -> pmx_ctx * {
    if (err && errlen) snprintf(err, errlen, "%s", m.c_str());
    return nullptr;
  };
  if (!sys || !par) return fail("pmx_create: null argument");
  if (sys->natm < 2 || sys->ntype < 1 || par->pi_beads < 1 || par->nsplin < 10)
    return fail("pmx_create: invalid sizes");
  if (par->pi_beads > 64) return fail("pmx_create: more than 64 beads per context is not supported");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail("pmx_create: no HIP device");
  if (device < 0 || device >= ndev) return fail("pmx_create: device index out of range");
  pmx_ctx *c = new pmx_ctx();
  c->device = device;
  c->par = *par;
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return fail("pmx_create: hipSetDevice failed");
  }
  std::string m = build_setup(sys, par, ix, iy, iz, c->s);
  if (!m.empty()) {
    delete c;
    return fail("pmx_create: " + m);
  }
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail("pmx_create: stream creation failed");
  }
  for (auto &e : c->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      pmx_destroy(c);
      return fail("pmx_create: event creation failed");
    }
  }
  if (hipEventCreateWithFlags(&c->ready, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->kfork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->kjoin, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->mxev, hipEventDisableTiming) != hipSuccess) {
    pmx_destroy(c);
    return fail("pmx_create: event creation failed");
  }
  if (hipStreamCreateWithFlags(&c->sk, hipStreamNonBlocking) != hipSuccess) {
    pmx_destroy(c);
    return fail("pmx_create: stream creation failed");
  }
  if (const char *env = getenv("PMX_OVERLAP")) c->overlap = atoi(env) != 0;  // tuning override 0 / 1
  // tuning override: 1 runs the molecule-cluster form of the mixed pair
  // kernel where the box qualifies (the per-site form is the default)
  if (const char *env = getenv("PMX_CLUSTER")) c->clu_enabled = atoi(env) == 1;
  if (const char *env = getenv("PMX_PAIR_WAVES"))  // tuning override: wavefronts of the pair kernels
    c->pair_waves = std::max(64, std::min(PAIR_MAX_WAVES, atoi(env)));
  if (const char *env = getenv("PMX_LIST_WAVES"))  // tuning override: wavefronts per cell of the list build
    c->list_waves = atoi(env) >= 16 ? 16 : (atoi(env) >= 8 ? 8 : 4);
  if (const char *env = getenv("PMX_PME_TILE"))  // tuning override "Ta,Tb,Tc" (upper bounds)
    sscanf(env, "%d,%d,%d", &c->tile_cap[0], &c->tile_cap[1], &c->tile_cap[2]);
  if (build_device(c, sys) != PMX_OK) {
    std::string e = c->err;
    pmx_destroy(c);
    return fail("pmx_create: " + e);
  }
  if (scratch_setup(device) != PMX_OK) {
    pmx_destroy(c);
    return fail("pmx_create: hipGetSymbolAddress of the static scratch failed");
  }
  return c;
}

void pmx_destroy(pmx_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->st) (void)hipStreamSynchronize(c->st);
  if (c->sk) (void)hipStreamSynchronize(c->sk);
  for (void *p : c->allocs) (void)hipFree(p);
  if (c->mx_pin) (void)hipHostFree(c->mx_pin);
  if (c->acc_pin) (void)hipHostFree(c->acc_pin);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->ready) (void)hipEventDestroy(c->ready);
  if (c->kfork) (void)hipEventDestroy(c->kfork);
  if (c->kjoin) (void)hipEventDestroy(c->kjoin);
  if (c->mxev) (void)hipEventDestroy(c->mxev);
  if (c->st) (void)hipStreamDestroy(c->st);
  if (c->sk) (void)hipStreamDestroy(c->sk);
  delete c;
}

const char *pmx_last_error(const pmx_ctx *c) { return c ? c->err.c_str() : "null context"; }

static int set_pos_common(pmx_ctx *c, const double *x, const double *y, const double *z, hipMemcpyKind kind) {
  Setup &s = c->s;
  const size_t NP = (size_t)s.N * s.P;
  HIPCHK(hipSetDevice(c->device));
  const double *src[3] = {x, y, z};
  double *dst[3] = {c->x, c->y, c->z};
  for (int d = 0; d < 3; d++) {
    HIPCHK(hipMemcpyAsync(c->stage, src[d], NP * sizeof(double), kind, c->st));
    hipLaunchKernelGGL(k_bead_to_atom, dim3(blocks(NP)), dim3(256), 0, c->st, s.N, s.P, c->stage, dst[d]);
    HIPCHK(hipGetLastError());
  }
  if (c->nghost) {
    hipLaunchKernelGGL(k_ghost_pos, dim3(blocks((size_t)c->nghost * s.P)), dim3(256), 0, c->st, c->nghost, s.P,
                       c->gatom, c->gncomp, c->gcomp, c->gcoef, c->x, c->y, c->z);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(c->st));
  c->have_pos = true;
  c->grad_ready = false;
  return PMX_OK;
}

int pmx_set_positions(pmx_ctx *c, const double *x, const double *y, const double *z) {
  if (!c || !x || !y || !z) return PMX_ERR_ARG;
  return set_pos_common(c, x, y, z, hipMemcpyHostToDevice);
}

int pmx_set_positions_device(pmx_ctx *c, const double *x, const double *y, const double *z) {
  if (!c || !x || !y || !z) return PMX_ERR_ARG;
  return set_pos_common(c, x, y, z, hipMemcpyDeviceToDevice);
}

int pmx_set_xmod(pmx_ctx *c, const double *x, const double *y, const double *z) {
  if (!c) return PMX_ERR_ARG;
  if (!x) {
    c->xmod_given = false;
    c->grad_ready = false;
    return PMX_OK;
  }
  const size_t n = (size_t)c->s.N * sizeof(double);
  HIPCHK(hipMemcpy(c->xm, x, n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->ym, y, n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->zm, z, n, hipMemcpyHostToDevice));
  c->xmod_given = true;
  c->grad_ready = false;
  return PMX_OK;
}

int pmx_set_respa(pmx_ctx *c, const pmx_respa *r) {
  if (!c) return PMX_ERR_ARG;
  if (!r) {
    c->respa = pmx_respa{};
    c->epoch++;  // captured batch sequences hold the weights
    return PMX_OK;
  }
  if (r->respa_steps_lrf < 0 || r->respa_steps_torsion < 0 || r->respa_steps_intra < 0 || r->respa_steps_pimd < 0 ||
      r->nfree < 0) {
    set_err(c, "pmx_set_respa: negative step count or degrees of freedom");
    return PMX_ERR_ARG;
  }
  // set_sim_params.c:3346-3370
  if ((r->respa_steps_lrf > 0 && (r->respa_steps_intra == 0 || r->respa_steps_torsion == 0)) ||
      (r->respa_steps_torsion > 0 && r->respa_steps_intra == 0)) {
    set_err(c, "pmx_set_respa: the inter split needs the torsion and intramolecular splits, and the torsion "
               "split the intramolecular one (respa_steps_intra, respa_steps_torsion > 0)");
    return PMX_ERR_ARG;
  }
  if (r->respa_steps_lrf > 0 && !(c->par.rheal > 0.0)) {
    set_err(c, "pmx_set_respa: the inter split needs respa_rheal > 0");
    return PMX_ERR_ARG;
  }
  c->respa = *r;
  c->epoch++;
  return PMX_OK;
}

// A full call under RESPA weights other than 1 runs twice: first without
// RESPA, whose forces are the reference's unweighted fxt (every term enters
// fxt before its weight is applied: force_npol.c:1222-1287,
// ewald3d_pme.c:568-577, ecorr.c:210-231, bond.c:270-277, bend.c:360-394)
// and give the virial estimator and pvten_tot (pimd_estimators.c:223-303);
// then with the weights, for the forces and pvten.
static int weighted_full_call(pmx_ctx *c, unsigned flags, pmx_energies *out) {
  int rc;
  hipStream_t st = c->st;
  const size_t NP = (size_t)c->s.N * c->s.P;
  const pmx_respa saved = c->respa;
  const unsigned plain = flags & ~(unsigned)(PMX_EVAL_RES_INTER | PMX_EVAL_NO_FULL_INTER | PMX_EVAL_NO_FULL_INTRA);
  pmx_energies u{};
  c->respa = pmx_respa{};
  rc = list_stage_launch(c, st, true);
  if (!rc) rc = pre_sequence(c, plain, st, false, true);
  if (!rc && hipEventSynchronize(c->mxev) != hipSuccess) rc = PMX_ERR_DEVICE;
  if (!rc) rc = list_stage_finish(c, plain, out);
  if (!rc) rc = main_sequence(c, plain, st, false);
  if (!rc) rc = readback_launch(c, st);
  if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = PMX_ERR_DEVICE;
  if (!rc) assemble(c, plain, &u, false);  // still without RESPA: its pressure terms unweighted
  c->respa = saved;
  if (rc) return rc;
  const bool modes = (flags & PMX_EVAL_PIMD) && (flags & PMX_EVAL_MODE_FORCES) && c->s.P > 1;
  if (!modes) {  // with modes the mode kernel kept them
    double *dst[3] = {c->fxt, c->fyt, c->fzt};
    const double *src[3] = {c->fx, c->fy, c->fz};
    for (int d = 0; d < 3; d++)
      HIPCHK(hipMemcpyAsync(dst[d], src[d], NP * sizeof(double), hipMemcpyDeviceToDevice, st));
  }
  c->keep_fxt = true;
  rc = pre_sequence(c, flags, st, true, true);
  if (!rc) rc = main_sequence(c, flags, st, true);
  c->keep_fxt = false;
  if (rc) return rc;
  if ((rc = readback_launch(c, st))) return rc;
  HIPCHK(hipStreamSynchronize(st));
  const int nver = out->nver, rebuilt = out->list_rebuilt;
  const double spread = out->spread;
  assemble(c, flags, out, true);
  out->nver = nver;
  out->list_rebuilt = rebuilt;
  out->spread = spread;
  // with iget_full_intra = 0 (no reference integrator makes such a full
  // call) the estimator is 0 and pvten_tot lacks the bead correction,
  // which the unweighted pass does not hold apart
  const bool vir = !(flags & PMX_EVAL_NO_FULL_INTRA);
  out->pi_ke_vir = vir ? u.pi_ke_vir : 0.0;
  for (int k = 0; k < 9; k++) out->pvten_tot[k] = (vir || c->s.P == 1) ? u.pvten[k] : NAN;
  c->last_weighted = false;
  c->fxt_kept = true;
  return PMX_OK;
}

static int force_call(pmx_ctx *c, unsigned flags, pmx_energies *out) {
  int rc;
  HIPCHK(hipSetDevice(c->device));
  memset(out, 0, sizeof(*out));
  c->grad_ready = false;
  c->fxt_kept = false;
  hipStream_t st = c->st;
  {
    const RespaCall rsp = respa_call(c, flags);
    if (rsp.full && rsp.weighted) return weighted_full_call(c, flags, out);
  }
  if ((rc = list_stage_launch(c, st, true))) return rc;
  if ((rc = pre_sequence(c, flags, st, true, true))) return rc;
  HIPCHK(hipEventSynchronize(c->mxev));
  if ((rc = list_stage_finish(c, flags, out))) return rc;
  if (c->clu_on) hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, c->clu_flag + 1, 0, (size_t)1);
  if ((rc = main_sequence(c, flags, st, true))) return rc;
  if ((rc = readback_launch(c, st))) return rc;
  HIPCHK(hipStreamSynchronize(st));
  assemble(c, flags, out, true);
  if (clu_recheck(c)) {  // the call again with the per-site kernel
    const int rebuilt = out->list_rebuilt;
    rc = force_call(c, flags, out);
    out->list_rebuilt |= rebuilt;
    return rc;
  }
  return PMX_OK;
}

int pmx_energy_control(pmx_ctx *c, unsigned flags, pmx_energies *out) {
  int rc = check_call(c, flags, out, "pmx_energy_control");
  if (rc) return rc;
  return force_call(c, flags, out);
}

int pmx_energy_control_noalloc(pmx_ctx *c, unsigned flags, pmx_energies *out) {
  int rc = check_call(c, flags, out, "pmx_energy_control_noalloc");
  if (rc) return rc;
  if (c->part_n * sizeof(double) > PMX_NOALLOC_SCRATCH_BYTES) {
    set_err(c, "pmx_energy_control_noalloc: the context's partial sums take %zu bytes, more than the %zu of "
               "the static scratch", c->part_n * sizeof(double), (size_t)PMX_NOALLOC_SCRATCH_BYTES);
    return PMX_ERR_RANGE;
  }
  double *ptr = scratch_address(c->device);
  if (!ptr) {
    set_err(c, "pmx_energy_control_noalloc: no scratch address for device %d", c->device);
    return PMX_ERR_DEVICE;
  }
  double *own = c->part;
  c->part = ptr;
  c->no_alloc = true;
  rc = force_call(c, flags, out);
  c->part = own;
  c->no_alloc = false;
  return rc;
}

// --- batches of contexts --------------------------------------------------------

static void batch_release_graph(pmx_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

pmx_batch *pmx_batch_create(pmx_ctx *const *ctx, int n, char *err, size_t errlen) {
  auto fail = [&](const char *m, pmx_batch *b) -> pmx_batch * {
    if (err && errlen) snprintf(err, errlen, "pmx_batch_create: %s", m);
    if (b) pmx_batch_destroy(b);
    return nullptr;
  };
  if (!ctx || n < 1) return fail("no contexts", nullptr);
  for (int i = 0; i < n; i++) {
    if (!ctx[i]) return fail("null context", nullptr);
    if (ctx[i]->device != ctx[0]->device) return fail("the contexts of a batch share one device", nullptr);
    for (int k = 0; k < i; k++)
      if (ctx[k] == ctx[i]) return fail("a context appears twice", nullptr);
  }
  pmx_batch *b = new pmx_batch();
  b->device = ctx[0]->device;
  b->ctx.assign(ctx, ctx + n);
  if (hipSetDevice(b->device) != hipSuccess) return fail("hipSetDevice failed", b);
  if (hipStreamCreateWithFlags(&b->cs, hipStreamNonBlocking) != hipSuccess) return fail("stream creation failed", b);
  if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) return fail("event creation failed", b);
  b->join.assign(n, nullptr);
  for (auto &e : b->join)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return fail("event creation failed", b);
  return b;
}

void pmx_batch_destroy(pmx_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->cs) (void)hipStreamSynchronize(b->cs);
  batch_release_graph(b);
  for (auto e : b->join)
    if (e) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->cs) (void)hipStreamDestroy(b->cs);
  delete b;
}

const char *pmx_batch_last_error(const pmx_batch *b) { return b ? b->err.c_str() : "null batch"; }

static unsigned batch_flags(const pmx_ctx *c, unsigned flags) {
  return c->s.P > 1 ? (flags | PMX_EVAL_PIMD) : (flags & ~(unsigned)PMX_EVAL_MODE_FORCES);
}

// Captures every context's force sequence as a parallel branch of one graph.
static int batch_capture(pmx_batch *b, unsigned flags) {
  batch_release_graph(b);
  if (hipStreamBeginCapture(b->cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    b->err = "graph capture could not begin";
    return PMX_ERR_DEVICE;
  }
  int rc = PMX_OK;
  hipError_t e = hipEventRecord(b->fork, b->cs);
  for (size_t i = 0; i < b->ctx.size() && e == hipSuccess && rc == PMX_OK; i++) {
    pmx_ctx *c = b->ctx[i];
    e = hipStreamWaitEvent(c->st, b->fork, 0);
    if (e != hipSuccess) break;
    rc = force_sequence(c, batch_flags(c, flags), c->st);
    if (rc) {
      b->err = "context " + std::to_string(i) + ": " + c->err;
      break;
    }
    e = hipEventRecord(b->join[i], c->st);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->cs, b->join[i], 0);
  }
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(b->cs, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess || ee != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    b->err = std::string("graph capture failed: ") + hipGetErrorString(e != hipSuccess ? e : ee);
    return PMX_ERR_DEVICE;
  }
  b->graph = g;
  e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    b->exec = nullptr;
    b->err = std::string("graph instantiation failed: ") + hipGetErrorString(e);
    return PMX_ERR_DEVICE;
  }
  b->flags = flags;
  b->epochs.resize(b->ctx.size());
  for (size_t i = 0; i < b->ctx.size(); i++) b->epochs[i] = b->ctx[i]->epoch;
  return PMX_OK;
}

int pmx_batch_energy_control(pmx_batch *b, unsigned flags, pmx_energies *out) {
  if (!b || !out) return PMX_ERR_ARG;
  const size_t n = b->ctx.size();
  for (size_t i = 0; i < n; i++) {
    pmx_ctx *c = b->ctx[i];
    if (!c->have_pos) {
      b->err = "context " + std::to_string(i) + " has no positions";
      return PMX_ERR_STATE;
    }
  }
  if (hipSetDevice(b->device) != hipSuccess) return PMX_ERR_DEVICE;
  auto dev_fail = [&](hipError_t e, const char *what) {
    b->err = std::string(what) + ": " + hipGetErrorString(e);
    return PMX_ERR_DEVICE;
  };
  // list stages: all checks in flight, then the decisions and rebuilds
  for (size_t i = 0; i < n; i++) {
    pmx_ctx *c = b->ctx[i];
    memset(&out[i], 0, sizeof(pmx_energies));
    c->grad_ready = false;
    if (int rc = list_stage_launch(c, c->st, false)) {
      b->err = "context " + std::to_string(i) + ": " + c->err;
      return rc;
    }
  }
  for (size_t i = 0; i < n; i++) {
    pmx_ctx *c = b->ctx[i];
    hipError_t e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return dev_fail(e, "list check");
    if (int rc = list_stage_finish(c, batch_flags(c, flags), &out[i])) {
      b->err = "context " + std::to_string(i) + ": " + c->err;
      return rc;
    }
    if (c->clu_on) hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->st, c->clu_flag + 1, 0, (size_t)1);
    e = hipEventRecord(c->ready, c->st);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->cs, c->ready, 0);
    if (e != hipSuccess) return dev_fail(e, "list stage");
  }
  bool recapture = !b->exec || b->flags != flags;
  for (size_t i = 0; i < n && !recapture; i++) recapture = b->epochs[i] != b->ctx[i]->epoch;
  if (recapture)
    if (int rc = batch_capture(b, flags)) return rc;
  hipError_t e = hipGraphLaunch(b->exec, b->cs);
  if (e != hipSuccess) return dev_fail(e, "graph launch");
  for (size_t i = 0; i < n; i++)
    if (int rc = readback_launch(b->ctx[i], b->cs)) {
      b->err = "context " + std::to_string(i) + ": " + b->ctx[i]->err;
      return rc;
    }
  e = hipStreamSynchronize(b->cs);
  if (e != hipSuccess) return dev_fail(e, "batched force call");
  for (size_t i = 0; i < n; i++) assemble(b->ctx[i], batch_flags(b->ctx[i], flags), &out[i], false);
  bool again = false;
  for (size_t i = 0; i < n; i++) again = clu_recheck(b->ctx[i]) || again;
  if (again) {  // the call again, those boxes with the per-site kernel
    std::vector<int> rebuilt(n);
    for (size_t i = 0; i < n; i++) rebuilt[i] = out[i].list_rebuilt;
    const int rc = pmx_batch_energy_control(b, flags, out);
    for (size_t i = 0; i < n; i++) out[i].list_rebuilt |= rebuilt[i];
    return rc;
  }
  return PMX_OK;
}

// Launch shapes picked by measurement (the batching path's autotuner): the
// pair kernels' wavefront count, the PME tile caps and the list build, each
// in turn, timed with the stage events of reps force calls at the current
// positions (one untimed call first).  The force call's results at the end
// are those of the chosen shapes.
int pmx_autotune(pmx_ctx *c, unsigned flags, int reps, int *choice) {
  if (!c) return PMX_ERR_ARG;
  if (!c->have_pos) {
    set_err(c, "pmx_autotune: no positions set");
    return PMX_ERR_STATE;
  }
  if (reps < 1) reps = 1;
  pmx_energies e;
  auto stage_ms = [&](int stage, unsigned f, double *ms) -> int {
    int rc = pmx_energy_control(c, f, &e);
    if (rc) return rc;
    double t = 0.0;
    for (int r = 0; r < reps; r++) {
      if ((rc = pmx_energy_control(c, f, &e))) return rc;
      t += c->ms[stage];
    }
    *ms = t / reps;
    return PMX_OK;
  };
  int rc = PMX_OK;
  // the stages are timed one after the other, then PME beside the pair
  // kernel is tried last on the whole call
  c->overlap = 0;
  // pair kernels: wavefronts (stage 1, real space)
  {
    const int cand[] = {1024, 2048, 4096, 8192, 16384};
    int best = c->pair_waves;
    double tbest = 1e300;
    for (int w : cand) {
      if (w > PAIR_MAX_WAVES || (w > c->s.N && w != cand[0])) continue;
      c->pair_waves = w;
      double t;
      if ((rc = stage_ms(1, flags, &t))) return rc;
      if (t < tbest) {
        tbest = t;
        best = w;
      }
    }
    c->pair_waves = best;
    // the molecule-cluster form against the per-site one, where the box
    // qualifies (the rows are built with a forced list build)
    if (c->clu_ok) {
      const bool was = c->clu_enabled;
      c->clu_enabled = !was;
      double t;
      if ((rc = pmx_energy_control(c, flags | PMX_EVAL_FORCE_LIST, &e))) return rc;
      if ((rc = stage_ms(1, flags, &t))) return rc;
      if (!(t < tbest)) {
        c->clu_enabled = was;
        if ((rc = pmx_energy_control(c, flags | PMX_EVAL_FORCE_LIST, &e))) return rc;
      }
    }
  }
  // PME tiles (stage 2, k space)
  if (c->pme) {
    const int cand[][3] = {{16, 16, 8}, {8, 8, 8}, {16, 8, 8}, {16, 16, 4}, {32, 16, 8}, {32, 8, 8}, {8, 8, 4}};
    int best[3] = {c->tile_cap[0], c->tile_cap[1], c->tile_cap[2]};
    double tbest = 1e300;
    for (const auto &cap : cand) {
      for (int d = 0; d < 3; d++) c->tile_cap[d] = cap[d];
      if (setup_pme_tiles(c) != PMX_OK) continue;
      double t;
      if ((rc = stage_ms(2, flags, &t))) return rc;
      if (t < tbest) {
        tbest = t;
        for (int d = 0; d < 3; d++) best[d] = cap[d];
      }
    }
    for (int d = 0; d < 3; d++) c->tile_cap[d] = best[d];
    if ((rc = setup_pme_tiles(c))) return rc;
  }
  // list build (stage 0, with the list rebuilt every call)
  {
    int nc[3];
    c->list_mode = 2;
    const bool cells_ok = choose_cells(c, nc) > 0;
    double tb = 0.0, tc = 1e300;
    c->list_mode = 1;
    if ((rc = stage_ms(0, flags | PMX_EVAL_FORCE_LIST, &tb))) return rc;
    if (cells_ok) {
      c->list_mode = 2;
      if ((rc = stage_ms(0, flags | PMX_EVAL_FORCE_LIST, &tc))) return rc;
    }
    c->list_mode = tc < tb ? 2 : 1;
  }
  // PME on the second stream (stage 4, the whole call), with the pair
  // kernels' wavefronts again: beside PME fewer of them leave the k-space
  // kernels room on the CUs (4096 q-TIP4P/F at P = 32: 0.717 ms with 3072
  // wavefronts and PME beside, 0.745 ms with the 8192 best alone)
  if (c->pme) {
    double tbest;
    if ((rc = stage_ms(4, flags, &tbest))) return rc;
    const int base = c->pair_waves;
    int best_w = base, best_ov = 0;
    const int cand[] = {base, 2048, 3072, 4096, 6144, 8192};
    c->overlap = 1;
    for (int w : cand) {
      if (w > PAIR_MAX_WAVES || (w > c->s.N && w != base)) continue;
      c->pair_waves = w;
      double t;
      if ((rc = stage_ms(4, flags, &t))) return rc;
      if (t < tbest) {
        tbest = t;
        best_w = w;
        best_ov = 1;
      }
    }
    c->overlap = best_ov;
    c->pair_waves = best_w;
  }
  c->epoch++;
  if ((rc = pmx_energy_control(c, flags | PMX_EVAL_FORCE_LIST, &e))) return rc;
  if (choice) {
    choice[0] = c->pa.T[0];
    choice[1] = c->pa.T[1];
    choice[2] = c->pa.T[2];
    choice[3] = c->pair_waves;
    choice[4] = c->list_mode;
  }
  return PMX_OK;
}

// Launch shapes of a batch picked by timing whole batched calls: the boxes
// share the device, so the shape that is fastest for a box alone is not the
// one that is fastest for the batch.  Each candidate (the pair kernels'
// wavefront cap, then the PME tile caps) is applied to every context and
// timed over reps calls after one untimed call (which recaptures the graph).
int pmx_batch_autotune(pmx_batch *b, unsigned flags, int reps, int *choice) {
  if (!b) return PMX_ERR_ARG;
  if (reps < 1) reps = 1;
  const size_t n = b->ctx.size();
  std::vector<pmx_energies> out(n);
  auto timed = [&](double *sec) -> int {
    int rc = pmx_batch_energy_control(b, flags, out.data());
    if (rc) return rc;
    const auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < reps; r++)
      if ((rc = pmx_batch_energy_control(b, flags, out.data()))) return rc;
    *sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / reps;
    return PMX_OK;
  };
  int rc = PMX_OK;
  {
    const int cand[] = {512, 1024, 2048, 4096, 8192, 16384};
    int best = b->ctx[0]->pair_waves;
    double tbest = 1e300;
    for (int w : cand) {
      for (pmx_ctx *c : b->ctx) {
        c->pair_waves = w;
        c->epoch++;
      }
      double t;
      if ((rc = timed(&t))) return rc;
      if (t < tbest) {
        tbest = t;
        best = w;
      }
    }
    for (pmx_ctx *c : b->ctx) {
      c->pair_waves = best;
      c->epoch++;
    }
    if (choice) choice[0] = best;
  }
  {
    const int cand[][3] = {{16, 16, 8}, {8, 8, 8}, {16, 8, 8}, {8, 8, 4}, {32, 16, 8}, {16, 16, 16}};
    int best = 0;
    double tbest = 1e300;
    for (int k = 0; k < (int)(sizeof cand / sizeof cand[0]); k++) {
      bool ok = true;
      for (pmx_ctx *c : b->ctx) {
        for (int d = 0; d < 3; d++) c->tile_cap[d] = cand[k][d];
        if (c->pme && setup_pme_tiles(c) != PMX_OK) ok = false;
        c->epoch++;
      }
      if (!ok) continue;
      double t;
      if ((rc = timed(&t))) return rc;
      if (t < tbest) {
        tbest = t;
        best = k;
      }
    }
    for (pmx_ctx *c : b->ctx) {
      for (int d = 0; d < 3; d++) c->tile_cap[d] = cand[best][d];
      if (c->pme && (rc = setup_pme_tiles(c))) return rc;
      c->epoch++;
    }
    if (choice)
      for (int d = 0; d < 3; d++) choice[1 + d] = cand[best][d];
  }
  return pmx_batch_energy_control(b, flags, out.data());
}

int pmx_get_forces(pmx_ctx *c, double *fx, double *fy, double *fz) {
  if (!c || !fx || !fy || !fz) return PMX_ERR_ARG;
  Setup &s = c->s;
  const size_t NP = (size_t)s.N * s.P;
  HIPCHK(hipSetDevice(c->device));
  double *src[3] = {c->fx, c->fy, c->fz};
  double *dst[3] = {fx, fy, fz};
  for (int d = 0; d < 3; d++) {
    hipLaunchKernelGGL(k_atom_to_bead, dim3(blocks(NP)), dim3(256), 0, c->st, s.N, s.P, src[d], c->stage);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(dst[d], c->stage, NP * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  }
  return PMX_OK;
}

int pmx_get_forces_device(pmx_ctx *c, const double **fx, const double **fy, const double **fz) {
  if (!c) return PMX_ERR_ARG;
  *fx = c->fx;
  *fy = c->fy;
  *fz = c->fz;
  return PMX_OK;
}

static int get_bead_major(pmx_ctx *c, double *const src[3], double *fx, double *fy, double *fz) {
  Setup &s = c->s;
  const size_t NP = (size_t)s.N * s.P;
  HIPCHK(hipSetDevice(c->device));
  double *dst[3] = {fx, fy, fz};
  for (int d = 0; d < 3; d++) {
    hipLaunchKernelGGL(k_atom_to_bead, dim3(blocks(NP)), dim3(256), 0, c->st, s.N, s.P, src[d], c->stage);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(dst[d], c->stage, NP * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  }
  return PMX_OK;
}

int pmx_get_fxt(pmx_ctx *c, double *fx, double *fy, double *fz) {
  // With RESPA off every force term enters fxt with the weight it has in f
  // (force_npol.c:1222-1283, ewald3d_pme.c:568-573, bond.c:272-285), so the
  // virial-estimator forces are the bead forces; after a call that returned
  // mode forces they are the bead forces kept before the transform.
  if (!c || !fx || !fy || !fz) return PMX_ERR_ARG;
  if (c->last_weighted) {
    set_err(c, "pmx_get_fxt: the last call was a RESPA call whose unweighted forces were not formed (a "
               "short-range or intramolecular call, or a batched one)");
    return PMX_ERR_STATE;
  }
  if (c->last_modes || c->fxt_kept) {
    double *src[3] = {c->fxt, c->fyt, c->fzt};
    return get_bead_major(c, src, fx, fy, fz);
  }
  return pmx_get_forces(c, fx, fy, fz);
}

static int set_bead_major(pmx_ctx *c, double *const dst[3], const double *x, const double *y, const double *z) {
  Setup &s = c->s;
  const size_t NP = (size_t)s.N * s.P;
  HIPCHK(hipSetDevice(c->device));
  const double *src[3] = {x, y, z};
  for (int d = 0; d < 3; d++) {
    HIPCHK(hipMemcpyAsync(c->stage, src[d], NP * sizeof(double), hipMemcpyHostToDevice, c->st));
    hipLaunchKernelGGL(k_bead_to_atom, dim3(blocks(NP)), dim3(256), 0, c->st, s.N, s.P, c->stage, dst[d]);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(c->st));
  return PMX_OK;
}

int pmx_set_velocities(pmx_ctx *c, const double *vx, const double *vy, const double *vz) {
  if (!c || !vx || !vy || !vz) return PMX_ERR_ARG;
  double *dst[3] = {c->vx, c->vy, c->vz};
  return set_bead_major(c, dst, vx, vy, vz);
}

int pmx_get_velocities(pmx_ctx *c, double *vx, double *vy, double *vz) {
  if (!c || !vx || !vy || !vz) return PMX_ERR_ARG;
  double *src[3] = {c->vx, c->vy, c->vz};
  return get_bead_major(c, src, vx, vy, vz);
}

int pmx_get_positions(pmx_ctx *c, double *x, double *y, double *z) {
  if (!c || !x || !y || !z) return PMX_ERR_ARG;
  double *src[3] = {c->x, c->y, c->z};
  return get_bead_major(c, src, x, y, z);
}

// Velocity Verlet in the NVE ensemble, int_NVE (md/int_NVE.c:29-125):
// int_0_to_dt2_nve (int_0_to_dt2.c:35-124: v += f dt/2 / m, x += v dt, ghost
// positions), the force call, int_dt2_to_dt_nve (int_dt2_to_dt.c:33-100:
// v += f dt/2 / m) and the kinetic energy of get_tvten.  Classical (P = 1).
int pmx_md_nve(pmx_ctx *c, int nsteps, double dt, pmx_energies *out, double *kinetic) {
  if (!c || !out || nsteps < 0) return PMX_ERR_ARG;
  Setup &s = c->s;
  if (s.P != 1) {
    set_err(c, "pmx_md_nve: the device integrator is classical (P = 1); use the force call for beads");
    return PMX_ERR_NOTIMPL;
  }
  if (respa_on(c)) {
    set_err(c, "pmx_md_nve: int_NVE runs without RESPA; pmx_set_respa(ctx, NULL) first");
    return PMX_ERR_STATE;
  }
  if (!c->have_pos) {
    set_err(c, "pmx_md_nve: no positions set");
    return PMX_ERR_STATE;
  }
  const size_t N = (size_t)s.N;
  int rc = PMX_OK;
  if (!c->forces_valid || c->last_modes) {
    rc = pmx_energy_control(c, 0, out);
    if (rc) return rc;
  }
  const double dt2 = dt / 2.0;
  for (int step = 0; step < nsteps; step++) {
    hipLaunchKernelGGL(k_vv_kick, dim3(blocks(N)), dim3(256), 0, c->st, (int)N, dt2, c->mass, c->fx, c->fy, c->fz,
                       c->vx, c->vy, c->vz);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_vv_drift, dim3(blocks(N)), dim3(256), 0, c->st, (int)N, dt, c->vx, c->vy, c->vz, c->x,
                       c->y, c->z);
    HIPCHK(hipGetLastError());
    if (c->nghost) {
      hipLaunchKernelGGL(k_ghost_pos, dim3(blocks((size_t)c->nghost)), dim3(256), 0, c->st, c->nghost, 1, c->gatom,
                         c->gncomp, c->gcomp, c->gcoef, c->x, c->y, c->z);
      HIPCHK(hipGetLastError());
    }
    rc = pmx_energy_control(c, 0, out);
    if (rc) return rc;
    hipLaunchKernelGGL(k_vv_kick, dim3(blocks(N)), dim3(256), 0, c->st, (int)N, dt2, c->mass, c->fx, c->fy, c->fz,
                       c->vx, c->vy, c->vz);
    HIPCHK(hipGetLastError());
  }
  if (kinetic) {
    HIPCHK(hipMemsetAsync(c->ke, 0, sizeof(double), c->st));
    hipLaunchKernelGGL(k_kinetic, dim3(blocks(N)), dim3(256), 0, c->st, (int)N, c->mass, c->vx, c->vy, c->vz, c->ke);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(kinetic, c->ke, sizeof(double), hipMemcpyDeviceToHost, c->st));
  }
  HIPCHK(hipStreamSynchronize(c->st));
  return PMX_OK;
}

namespace {

// Allocates (zeroed) chain state for chains of length len.
int nhc_alloc(pmx_ctx *c, int len) {
  Setup &s = c->s;
  const size_t NP = (size_t)s.N * s.P;
  if (c->nhc_len == len) return PMX_OK;
  dfree(c, c->vnhc);
  dfree(c, c->xnhc);
  dfree(c, c->nhc_cls);
  c->nhc_len = 0;
  TRY_RC(dalloc(c, &c->vnhc, (size_t)len * 3 * NP));
  TRY_RC(dalloc(c, &c->xnhc, (size_t)len * 3 * NP));
  TRY_RC(dalloc(c, &c->nhc_cls, (size_t)2 * len + 1));
  HIPCHK(hipMemset(c->vnhc, 0, (size_t)len * 3 * NP * sizeof(double)));
  HIPCHK(hipMemset(c->xnhc, 0, (size_t)len * 3 * NP * sizeof(double)));
  HIPCHK(hipMemset(c->nhc_cls, 0, ((size_t)2 * len + 1) * sizeof(double)));
  if (!c->mfac) TRY_RC(dalloc(c, &c->mfac, 64));
  if (!c->nhc_part) TRY_RC(dalloc(c, &c->nhc_part, 2 * NHC_BLOCKS));
  if (!c->nhc_out) TRY_RC(dalloc(c, &c->nhc_out, 2));
  if (!c->ghostflag) {
    std::vector<int> g(s.N, 0);
    for (int k = 0; k < c->nghost; k++) g[c->ghost_h[k]] = 1;
    TRY_RC(upload(c, &c->ghostflag, g.data(), g.size()));
  }
  c->nhc_len = len;
  return PMX_OK;
}

// The integrator's constants: set_yosh (int_utils.c:244-318) for the chain
// step dt / nres, and the chains' masses and gkt as set_atm_NHC sets them
// (set_atm_NHC.c:355-370, 510-547 for the centroid chain on the global
// thermostat, 715-755 for the bead chains), with the mode mass factors of
// path_integral_init (transform_cnt.c:175-233).
int nhc_args(pmx_ctx *c, const pmx_nhc_params *np, NhcArgs &a, std::vector<double> &mf) {
  Setup &s = c->s;
  const int P = s.P;
  a = NhcArgs{};
  a.N = s.N;
  a.P = P;
  a.len = np->len_nhc;
  a.nres = np->nres_nhc;
  a.nyosh = np->nyosh_nhc;
  double w[10] = {0};
  if (a.nyosh == 1) {
    w[1] = 1.0;
  } else if (a.nyosh == 3) {
    const double onethird = 1.0 / 3.0, temp = pow(2.0, onethird);
    w[1] = 1.0 / (2.0 - temp);
    w[2] = -temp / (2.0 - temp);
    w[3] = 1.0 / (2.0 - temp);
  } else if (a.nyosh == 5) {
    const double onethird = 1.0 / 3.0, temp = pow(4.0, onethird), p2 = 1.0 / (4.0 - temp);
    w[1] = p2;
    w[2] = p2;
    w[3] = 1.0 - 4.0 * p2;
    w[4] = p2;
    w[5] = p2;
  } else {
    set_err(c, "pmx_pimd_nvt: yosh_steps_nhc %d (1, 3 or 5 are supported)", a.nyosh);
    return PMX_ERR_NOTIMPL;
  }
  const double dti = np->dt / (double)a.nres;
  for (int i = 1; i <= a.nyosh; i++) {
    const double wd = w[i] * dti;
    a.wdti2[i] = wd / 2.0;
    a.wdti4[i] = wd / 4.0;
    a.wdti8[i] = wd / 8.0;
  }
  const double T = c->par.temperature > 0 ? c->par.temperature : 300.0, tau = np->tau_nhc;
  const int nreal = s.N - c->nghost;
  const double m1 = tau * tau * T, g1 = 3.0 * (double)nreal;
  for (int j = 2; j <= a.len; j++) {
    a.cgkt[j] = T / BOLTZ;
    a.cmass[j] = m1 / BOLTZ;
  }
  a.cmass[1] = m1 * g1 / BOLTZ;
  a.cgkt[1] = g1 * T / BOLTZ;
  const double pre = (BOLTZ * BOLTZ * np->gamma_adb * np->gamma_adb) / ((double)P);
  const double bm1 = pre / (T * T);
  for (int j = 1; j <= a.len; j++) {
    a.bgkt[j] = T / BOLTZ;
    a.bmass[j] = bm1 * T / BOLTZ;
  }
  mf.assign(P, 1.0);
  const double dP = (double)P;
  for (int ip = 2; ip <= P; ip++) {
    double v;
    if (c->par.pimd_typ == PMX_PIMD_STAGING) {
      v = ((double)(ip)) / ((double)(ip - 1));
    } else if (ip == P) {
      v = 4.0 * dP;
    } else {
      const int k = ip / 2 + 1;  // veig[2k-2] = veig[2k-1] for k = 2..P/2
      v = (1.0 - cos(2.0 * M_PI / dP * (double)(k - 1))) * 4.0 * dP;
    }
    mf[ip - 1] = v * np->gamma_adb * np->gamma_adb;
  }
  a.dt = np->dt;
  a.dt2 = np->dt / 2.0;
  a.mass = c->mass;
  a.mfac = c->mfac;
  a.ghost = c->ghostflag;
  a.fx = c->fx;
  a.fy = c->fy;
  a.fz = c->fz;
  a.vx = c->vx;
  a.vy = c->vy;
  a.vz = c->vz;
  a.vnhc = c->vnhc;
  a.xnhc = c->xnhc;
  a.cls = c->nhc_cls;
  a.part = c->nhc_part;
  a.stag = c->par.pimd_typ == PMX_PIMD_STAGING ? 1 : 0;
  a.cs = c->cs;
  a.x = c->x;
  a.y = c->y;
  a.z = c->z;
  a.xm = c->xm;
  a.ym = c->ym;
  a.zm = c->zm;
  return PMX_OK;
}

int check_nhc_params(pmx_ctx *c, const pmx_nhc_params *np) {
  if (!np || np->len_nhc < 1 || np->len_nhc > 7 || np->nres_nhc < 1 || !(np->dt > 0.0) || !(np->tau_nhc > 0.0) ||
      !(np->gamma_adb > 0.0)) {
    set_err(c, "pmx_pimd_nvt: invalid integrator settings");
    return PMX_ERR_ARG;
  }
  return PMX_OK;
}

// One force call of energy_control_pimd with mode forces, launched on the
// context's stream without the readback of the energies (the list decision
// is the host's one wait).
int pimd_force(pmx_ctx *c, unsigned flags) {
  pmx_energies tmp;
  memset(&tmp, 0, sizeof tmp);
  c->grad_ready = false;
  if (int rc = list_stage_launch(c, c->st, false)) return rc;
  if (int rc = pre_sequence(c, flags, c->st, false, true)) return rc;
  HIPCHK(hipEventSynchronize(c->mxev));
  if (int rc = list_stage_finish(c, flags, &tmp)) return rc;
  return main_sequence(c, flags, c->st, false);
}

}  // namespace

int pmx_pimd_nvt(pmx_ctx *c, const pmx_nhc_params *np, int nsteps, pmx_energies *out, double *econs) {
  if (!c || !out || nsteps < 0) return PMX_ERR_ARG;
  Setup &s = c->s;
  const int N = s.N, P = s.P;
  const size_t NP = (size_t)N * P;
  if (P < 2) {
    set_err(c, "pmx_pimd_nvt: one bead: use pmx_md_nve");
    return PMX_ERR_ARG;
  }
  if (respa_on(c)) {
    set_err(c, "pmx_pimd_nvt: the device integrator runs int_NVT_pimd with RESPA off; pmx_set_respa(ctx, NULL) "
               "first");
    return PMX_ERR_STATE;
  }
  if (!c->have_pos) {
    set_err(c, "pmx_pimd_nvt: no positions set");
    return PMX_ERR_STATE;
  }
  if (int rc = check_nhc_params(c, np)) return rc;
  HIPCHK(hipSetDevice(c->device));
  if (int rc = nhc_alloc(c, np->len_nhc)) return rc;
  NhcArgs a;
  std::vector<double> mf;
  if (int rc = nhc_args(c, np, a, mf)) return rc;
  HIPCHK(hipMemcpyAsync(c->mfac, mf.data(), P * sizeof(double), hipMemcpyHostToDevice, c->st));
  const unsigned flags = PMX_EVAL_PIMD | PMX_EVAL_MODE_FORCES;
  memset(out, 0, sizeof(*out));
  // forces at the starting positions, xmod as the context has it
  if (c->clu_ok) hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->st, c->clu_flag + 1, 0, (size_t)1);
  if (int rc = pimd_force(c, flags)) return rc;
  const bool xmod_given = c->xmod_given;
  c->xmod_given = true;  // xmod is mode 1 of each drift from here on
  const unsigned gN = blocks(N), gNP = blocks(NP);
  int rc = PMX_OK;
  for (int step = 0; step < nsteps && !rc; step++) {
    hipLaunchKernelGGL(k_nhc_ckin, dim3(NHC_BLOCKS), dim3(256), 0, c->st, a);
    hipLaunchKernelGGL(k_nhc_class, dim3(1), dim3(64), 0, c->st, a, NHC_BLOCKS);
    hipLaunchKernelGGL(k_nhc_kick, dim3(gNP), dim3(256), 0, c->st, a, 0);
    hipLaunchKernelGGL(k_mode_drift, dim3(mode_blocks(N, P)), dim3(256), 0, c->st, a);
    if (c->nghost)
      hipLaunchKernelGGL(k_ghost_pos, dim3(blocks((size_t)c->nghost * P)), dim3(256), 0, c->st, c->nghost, P,
                         c->gatom, c->gncomp, c->gcomp, c->gcoef, c->x, c->y, c->z);
    if (hipGetLastError() != hipSuccess) {
      rc = PMX_ERR_DEVICE;
      break;
    }
    if ((rc = pimd_force(c, flags))) break;
    hipLaunchKernelGGL(k_nhc_kick, dim3(gNP), dim3(256), 0, c->st, a, 1);
    hipLaunchKernelGGL(k_nhc_ckin, dim3(NHC_BLOCKS), dim3(256), 0, c->st, a);
    hipLaunchKernelGGL(k_nhc_class, dim3(1), dim3(64), 0, c->st, a, NHC_BLOCKS);
    hipLaunchKernelGGL(k_nhc_scale0, dim3(gN), dim3(256), 0, c->st, a);
  }
  c->xmod_given = xmod_given;
  if (rc == PMX_OK && hipGetLastError() != hipSuccess) rc = PMX_ERR_DEVICE;
  if (rc) {
    if (rc == PMX_ERR_DEVICE) set_err(c, "pmx_pimd_nvt: kernel launch failed");
    return rc;
  }
  TRY_RC(readback_launch(c, c->st));
  double eo[2] = {0.0, 0.0};
  if (econs) {
    hipLaunchKernelGGL(k_pimd_energy, dim3(NHC_BLOCKS), dim3(256), 0, c->st, a);
    hipLaunchKernelGGL(k_pimd_energy_sum, dim3(1), dim3(64), 0, c->st, a, NHC_BLOCKS, c->nhc_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(eo, c->nhc_out, sizeof eo, hipMemcpyDeviceToHost, c->st));
  }
  HIPCHK(hipStreamSynchronize(c->st));
  assemble(c, flags, out, false);
  if (clu_recheck(c)) {
    set_err(c, "pmx_pimd_nvt: a listed site pair's periodic image differed from its molecule pair's; the "
               "molecule-cluster kernel is off until the next list build: repeat the call");
    return PMX_ERR_STATE;
  }
  if (econs) {
    econs[0] = eo[0];
    econs[1] = eo[1];
    econs[2] = out->vintert + out->vintrat + out->kin_harm;
    econs[3] = econs[0] + econs[1] + econs[2];
  }
  return PMX_OK;
}

int pmx_set_nhc_state(pmx_ctx *c, int len_nhc, const double *st) {
  if (!c || !st || len_nhc < 1 || len_nhc > 7) return PMX_ERR_ARG;
  Setup &s = c->s;
  const int N = s.N, P = s.P, L = len_nhc;
  if (P < 2) return PMX_ERR_ARG;
  HIPCHK(hipSetDevice(c->device));
  TRY_RC(nhc_alloc(c, L));
  const size_t NP = (size_t)N * P;
  std::vector<double> v((size_t)L * 3 * NP, 0.0), x((size_t)L * 3 * NP, 0.0), cls(2 * L + 1, 1.0);
  for (int j = 0; j < 2 * L; j++) cls[j] = st[j];
  for (int m = 1; m < P; m++)
    for (int i = 0; i < N; i++)
      for (int d = 0; d < 3; d++) {
        const double *src = st + 2 * L + ((size_t)(m - 1) * 3 * N + 3 * i + d) * 2 * L;
        for (int j = 0; j < L; j++) {
          const size_t o = ((size_t)(j * 3 + d)) * NP + (size_t)i * P + m;
          v[o] = src[j];
          x[o] = src[L + j];
        }
      }
  HIPCHK(hipMemcpy(c->vnhc, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->xnhc, x.data(), x.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->nhc_cls, cls.data(), cls.size() * sizeof(double), hipMemcpyHostToDevice));
  return PMX_OK;
}

int pmx_get_nhc_state(pmx_ctx *c, int len_nhc, double *st) {
  if (!c || !st) return PMX_ERR_ARG;
  Setup &s = c->s;
  const int N = s.N, P = s.P, L = len_nhc;
  if (P < 2 || L != c->nhc_len) {
    set_err(c, "pmx_get_nhc_state: no chain state of length %d", len_nhc);
    return PMX_ERR_STATE;
  }
  HIPCHK(hipSetDevice(c->device));
  const size_t NP = (size_t)N * P;
  std::vector<double> v((size_t)L * 3 * NP), x((size_t)L * 3 * NP), cls(2 * L + 1);
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipMemcpy(v.data(), c->vnhc, v.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(x.data(), c->xnhc, x.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(cls.data(), c->nhc_cls, cls.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int j = 0; j < 2 * L; j++) st[j] = cls[j];
  for (int m = 1; m < P; m++)
    for (int i = 0; i < N; i++)
      for (int d = 0; d < 3; d++) {
        double *dst = st + 2 * L + ((size_t)(m - 1) * 3 * N + 3 * i + d) * 2 * L;
        for (int j = 0; j < L; j++) {
          const size_t o = ((size_t)(j * 3 + d)) * NP + (size_t)i * P + m;
          dst[j] = v[o];
          dst[L + j] = x[o];
        }
      }
  return PMX_OK;
}

// New force-field parameters for the same topology: charges (the set of
// charged sites fixed), LJ sigma and eps, bond and bend coefficients.  The
// parameter-dependent setup (LJ tables, tail constant, charge sums) is
// rebuilt from the same input; the excluded-pair table keeps its range.
int pmx_set_force_field(pmx_ctx *c, const pmx_system *sys) {
  if (!c || !sys) return PMX_ERR_ARG;
  Setup &s = c->s;
  if (sys->natm != s.N || sys->ntype != s.ntype || sys->nbond != c->nbond || sys->nbend != c->nbend ||
      sys->nbond_type != c->nbond_type || sys->nbend_type != c->nbend_type || sys->nghost != c->nghost) {
    set_err(c, "pmx_set_force_field: the topology (sites, types, bonds, bends, ghosts) must not change");
    return PMX_ERR_ARG;
  }
  Setup s2;
  std::string m = build_setup(sys, &c->par, nullptr, nullptr, nullptr, s2);
  if (!m.empty()) {
    set_err(c, "pmx_set_force_field: %s", m.c_str());
    return PMX_ERR_ARG;
  }
  if (s2.ichrg != s.ichrg || s2.e1 != s.e1 || s2.e2 != s.e2 || s2.label != s.label) {
    set_err(c, "pmx_set_force_field: the charged sites and the LJ type pairs must not change");
    return PMX_ERR_ARG;
  }
  s.qsq = s2.qsq;
  s.qsum = s2.qsum;
  s.tab = s2.tab;
  s.lj_eps = s2.lj_eps;
  s.lj_sig6 = s2.lj_sig6;
  s.lj_sig = s2.lj_sig;
  s.vshift_lj = s2.vshift_lj;
  s.vshift_c = s2.vshift_c;
  s.clong = s2.clong;
  s.clong_res = s2.clong_res;
  s.clong_u = s2.clong_u;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->st));
  const size_t N = (size_t)s.N;
  HIPCHK(hipMemcpy(c->q, sys->q, N * sizeof(double), hipMemcpyHostToDevice));
  if (int rc = setup_molecules(c, sys)) return rc;  // the templates hold charge products
  {
    const size_t nt = (size_t)s.npair * s.nsplin;
    std::vector<double2> lj(nt), cc(nt);
    for (size_t k = 0; k < nt; k++) {
      lj[k] = make_double2(s.tab[4 * k + 0], s.tab[4 * k + 1]);
      cc[k] = make_double2(s.tab[4 * k + 2], s.tab[4 * k + 3]);
    }
    HIPCHK(hipMemcpy(c->tab_lj, lj.data(), nt * sizeof(double2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->tab_c, cc.data(), nt * sizeof(double2), hipMemcpyHostToDevice));
  }
  const std::vector<double> pp = pair_constants(c);
  HIPCHK(hipMemcpy(c->pair_pp, pp.data(), pp.size() * sizeof(double), hipMemcpyHostToDevice));
  const std::vector<double> ppr = pair_constants_res(c, pp);
  HIPCHK(hipMemcpy(c->pair_pp_res, ppr.data(), ppr.size() * sizeof(double), hipMemcpyHostToDevice));
  if (c->pair_mode == 2)
    if (int rc = build_pair_tables(c)) return rc;
  if (int rc = setup_cluster(c, sys)) return rc;  // same charged sites: the list and its rows stay
  c->bc_h.assign(sys->bond_c, sys->bond_c + 7 * (size_t)c->nbond_type);
  if (c->nbond_type) {
    HIPCHK(hipMemcpy(c->bc, sys->bond_c, 7 * (size_t)c->nbond_type * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->beq, sys->bond_eq, (size_t)c->nbond_type * sizeof(double), hipMemcpyHostToDevice));
  }
  if (c->nbend_type) {
    HIPCHK(hipMemcpy(c->ac, sys->bend_c, 7 * (size_t)c->nbend_type * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->as, sys->bend_s, 7 * (size_t)c->nbend_type * sizeof(double), hipMemcpyHostToDevice));
  }
  c->forces_valid = false;
  c->epoch++;  // captured sequences hold the molecule templates and counts
  c->grad_ready = false;
  return PMX_OK;
}

// Derivatives of every bead's potential energy with respect to the force
// field, at the state of the last force call (the backward of
// energy_control): site potentials from the real-space, k-space,
// excluded-pair, self and background terms; LJ sums incl. the tail; bond
// and bend moments.
int pmx_get_param_grad(pmx_ctx *c, pmx_param_grad *out) {
  if (!c || !out) return PMX_ERR_ARG;
  if (!c->grad_ready) {
    set_err(c, "pmx_get_param_grad: needs a force call (with the k-space term) at the current positions and "
               "parameters first");
    return PMX_ERR_STATE;
  }
  if (!c->full_list) {
    set_err(c, "pmx_get_param_grad: needs the full force list");
    return PMX_ERR_NOTIMPL;
  }
  if (c->ewald) {
    set_err(c, "pmx_get_param_grad: the charge gradients are formed for PME, not for standard Ewald");
    return PMX_ERR_NOTIMPL;
  }
  Setup &s = c->s;
  const int N = s.N, P = s.P;
  const size_t NP = (size_t)N * P;
  int nlj = 0;
  for (int t = 0; t < s.npair; t++) nlj += (s.label[t] == PMX_INTER_LJ);
  if (nlj > MAX_LJ_GRAD) {
    set_err(c, "pmx_get_param_grad: %d LJ type pairs (at most %d)", nlj, MAX_LJ_GRAD);
    return PMX_ERR_NOTIMPL;
  }
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = c->st;
  int rc = PMX_OK;
  const unsigned gx = pair_grid_x(N, c->pair_waves);
  const bool kernels = out->dq || out->deps || out->dsig || out->dbond_c || out->dbond_eq || out->dbend_c ||
                       out->dbend_s;
  if (kernels && !c->phi) {
    if ((rc = dalloc(c, &c->phi, NP))) return rc;
    if ((rc = dalloc(c, &c->gpart, (size_t)pair_grid_x(N, PAIR_MAX_WAVES) * P * PAIRG_NV))) return rc;
    if ((rc = dalloc(c, &c->gacc, (size_t)P * PAIRG_NV))) return rc;
    if ((rc = dalloc(c, &c->gbond, (size_t)P * std::max(c->nbond_type, 1) * 7))) return rc;
    if ((rc = dalloc(c, &c->gbend, (size_t)P * std::max(c->nbend_type, 1) * BEND_GRAD_NV))) return rc;
  }
  if (kernels) {  // real space: phi (stored for every site) and the LJ sums
    PairArgs a{};
    a.N = N;
    a.P = P;
    a.cap = c->cap;
    a.nsplin = s.nsplin;
    a.pimd = (P > 1);
    a.npair = s.npair;
    a.gx = pair_grid_x(N, c->pair_waves);
    a.alp = s.alp_ewd;
    a.p_alpha = kErfP * s.alp_ewd;
    a.two_alpha2 = 2.0 * s.alp_ewd * s.alp_ewd;
    a.spalp = 2.0 * s.alp_ewd / sqrt(M_PI);
    a.ctab = c->ctab;
    a.cswtab = c->cswtab;
    a.ct_r0 = c->ct_r0;
    a.ct_idr = c->ct_idr;
    a.ftab = c->ftab;
    a.nftab = c->nftab;
    for (size_t t = 0; t < c->ftab_off.size(); t++) a.ftoff[t] = c->ftab_off[t];
    a.rheal = c->par.rheal;
    a.irheal = c->par.rheal > 0.0 ? 1.0 / c->par.rheal : 0.0;
    a.swit = (c->par.ishift == PMX_SHIFT_SWIT);
    a.pp = c->pair_pp;
    a.x = c->x;
    a.y = c->y;
    a.z = c->z;
    a.site = c->site;
    a.type = c->type;
    a.jlist = c->jforce;
    a.cnt = c->nforce;
    a.tab_lj = c->tab_lj;
    a.tab_c = c->tab_c;
    a.cell = c->cell;
    a.part = c->gpart;
    a.phi = c->phi;
    HIPCHK(launch_pair_grad(a, c->pair_mode == 0 ? 0 : 1, st));
    RedArgs red{};
    red.P = P;
    red.nslot = PAIRG_NV;
    red.nstage = 1;
    red.st[0].off = 0;
    red.st[0].G = (int)gx;
    red.st[0].nv = PAIRG_NV;
    red.st[0].wv = 1.0;
    for (int v = 0; v < PAIRG_NV; v++) red.st[0].slot[v] = v;
    red.part = c->gpart;
    red.acc = c->gacc;
    hipLaunchKernelGGL(k_reduce_partials, dim3((unsigned)P * PAIRG_NV + 1), dim3(256), 0, st, red);
    HIPCHK(hipGetLastError());
  }
  const int nchrg = (int)s.ichrg.size();
  if (kernels && c->pme && nchrg > 0) {  // k space, on the grid the force call left
    PmeArgs a = c->pa;
    a.x = c->x;
    a.y = c->y;
    a.z = c->z;
    a.cell = c->cell;
    a.phi = c->phi;
    HIPCHK(launch_pme_phi(a, st));
  }
  IntraArgs ia = intra_args(c, P > 1);
  ia.nchrg = nchrg;
  ia.ichrg = c->ichrg;
  ia.ecs_off = c->ecs_off;
  ia.ecs_num = c->ecs_num;
  ia.ecs_j = c->ecs_j;
  ia.cself = -2.0 * s.alp_ewd * s.self_erf / sqrt(M_PI);
  ia.cbgr = -M_PI * s.qsum / (s.alp_ewd * s.alp_ewd * s.vol);
  ia.phi = c->phi;
  ia.nbond_type = c->nbond_type;
  ia.nbend_type = c->nbend_type;
  ia.gbond = c->gbond;
  ia.gbend = c->gbend;
  if (kernels && nchrg > 0) {
    hipLaunchKernelGGL(k_ecor_phi, dim3(blocks((size_t)nchrg * P)), dim3(256), 0, st, ia);
    HIPCHK(hipGetLastError());
  }
  if (kernels && c->nbond && c->nbond_type) {
    hipLaunchKernelGGL(k_bond_moments, dim3(c->nbond_type, P), dim3(256), 0, st, ia);
    HIPCHK(hipGetLastError());
  }
  if (kernels && c->nbend && c->nbend_type) {
    hipLaunchKernelGGL(k_bend_moments, dim3(c->nbend_type, P), dim3(256), 0, st, ia);
    HIPCHK(hipGetLastError());
  }
  std::vector<double> gacc((size_t)P * PAIRG_NV, 0.0), gb((size_t)P * c->nbond_type * 7, 0.0),
      ga((size_t)P * c->nbend_type * BEND_GRAD_NV, 0.0);
  if (kernels) {
    HIPCHK(hipMemcpyAsync(gacc.data(), c->gacc, gacc.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    if (!gb.empty())
      HIPCHK(hipMemcpyAsync(gb.data(), c->gbond, gb.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    if (!ga.empty())
      HIPCHK(hipMemcpyAsync(ga.data(), c->gbend, ga.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  if (out->dq) {
    hipLaunchKernelGGL(k_atom_to_bead, dim3(blocks(NP)), dim3(256), 0, st, N, P, c->phi, c->stage);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out->dq, c->stage, NP * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  const double vlong = s.clong / s.vol;
  double vself = 0.0, vbgr = 0.0;
  if (nchrg > 0) {
    vself = -s.qsq * s.alp_ewd * s.self_erf / sqrt(M_PI);
    vbgr = -0.5 * s.qsum * s.qsum * M_PI / (s.alp_ewd * s.alp_ewd * s.vol);
  }
  for (int b = 0; b < P; b++) {
    const double *e = &c->acc_h[(size_t)b * ACC_NUM];
    if (out->energy)
      out->energy[b] = e[ACC_VVDW] + e[ACC_VCOUL] + e[ACC_VRECIP] + e[ACC_VECOR] + e[ACC_VBOND] + e[ACC_VBEND] +
                       vlong + vself + vbgr;
    for (int t = 0; t < s.npair; t++) {
      double de = 0.0, ds = 0.0;
      const int g = c->lj_slot[t];
      if (g >= 0) {
        const double U = gacc[(size_t)b * PAIRG_NV + g], S = gacc[(size_t)b * PAIRG_NV + MAX_LJ_GRAD + g];
        const double eps = s.lj_eps[t], sig = s.lj_sig[t];
        const double tail = s.clong_u[t] / s.vol;
        de = U + tail;
        ds = (eps * S + 6.0 * eps * tail) / sig;
      }
      if (out->deps) out->deps[(size_t)b * s.npair + t] = de;
      if (out->dsig) out->dsig[(size_t)b * s.npair + t] = ds;
    }
    const std::vector<double> &bc = c->bc_h;
    for (int ty = 0; ty < c->nbond_type; ty++) {
      const double *M = &gb[((size_t)b * c->nbond_type + ty) * 7];
      if (out->dbond_c)
        for (int k = 0; k < 7; k++) out->dbond_c[((size_t)b * c->nbond_type + ty) * 7 + k] = M[k];
      if (out->dbond_eq) {
        double d = 0.0;
        for (int k = 1; k < 7; k++) d += (double)k * bc[7 * ty + k] * M[k - 1];
        out->dbond_eq[(size_t)b * c->nbond_type + ty] = -d;
      }
    }
    for (int ty = 0; ty < c->nbend_type; ty++) {
      const double *M = &ga[((size_t)b * c->nbend_type + ty) * BEND_GRAD_NV];
      double *dc = out->dbend_c ? out->dbend_c + ((size_t)b * c->nbend_type + ty) * 7 : nullptr;
      double *ds = out->dbend_s ? out->dbend_s + ((size_t)b * c->nbend_type + ty) * 7 : nullptr;
      if (dc)
        for (int k = 0; k < 7; k++) dc[k] = M[k];
      if (ds) {
        ds[0] = 0.0;
        ds[1] = M[7];
        for (int k = 2; k < 7; k++) ds[k] = M[6 + k];
      }
    }
  }
  return PMX_OK;
}

int pmx_get_bead_energies(pmx_ctx *c, double *out) {
  if (!c || !out) return PMX_ERR_ARG;
  for (int b = 0; b < c->s.P; b++) {
    const double *a = &c->acc_h[(size_t)b * ACC_NUM];
    out[4 * b + 0] = a[ACC_VVDW] + a[ACC_VCOUL];
    out[4 * b + 1] = a[ACC_VRECIP];
    out[4 * b + 2] = a[ACC_VECOR];
    out[4 * b + 3] = a[ACC_VBOND] + a[ACC_VBEND];
  }
  return PMX_OK;
}

int pmx_get_verlet_list(pmx_ctx *c, int *nter, int *jver_off, int *jter, int cap, int *nver) {
  if (!c || !nter || !jver_off || !nver) return PMX_ERR_ARG;
  if (!c->have_list) {
    set_err(c, "no Verlet list yet");
    return PMX_ERR_STATE;
  }
  const int N = c->s.N;
  HIPCHK(hipSetDevice(c->device));
  if (!c->ref_valid)
    if (int rc = build_ref_list(c)) return rc;
  HIPCHK(hipMemcpy(nter, c->nfull, N * sizeof(int), hipMemcpyDeviceToHost));
  long tot = 0;
  for (int i = 0; i < N; i++) {
    jver_off[i] = (int)tot;
    tot += nter[i];
  }
  *nver = (int)tot;
  if (cap < tot || !jter) return cap < tot ? PMX_ERR_RANGE : PMX_OK;
  std::vector<int> row(c->cap);
  for (int i = 0; i < N; i++) {
    if (nter[i] == 0) continue;
    HIPCHK(hipMemcpy(jter + jver_off[i], c->jfull + (size_t)i * c->cap, nter[i] * sizeof(int),
                     hipMemcpyDeviceToHost));
  }
  return PMX_OK;
}

int pmx_get_nktot(pmx_ctx *c, int *nktot, int *ngrid) {
  if (!c || !nktot || !ngrid) return PMX_ERR_ARG;
  *nktot = c->s.nktot;
  for (int d = 0; d < 3; d++) ngrid[d] = c->s.ngrid[d];
  return PMX_OK;
}

int pmx_get_kvectors(pmx_ctx *c, int *ka, int *kb, int *kc, double *bw, int cap) {
  if (!c) return PMX_ERR_ARG;
  if (cap < c->s.nktot) return PMX_ERR_RANGE;
  for (int k = 0; k < c->s.nktot; k++) {
    ka[k] = c->s.ka[k];
    kb[k] = c->s.kb[k];
    kc[k] = c->s.kc[k];
    bw[k] = c->s.bweight[k];
  }
  return PMX_OK;
}

int pmx_get_spline(pmx_ctx *c, int which, int ipair, double *out, int cap) {
  if (!c || !out) return PMX_ERR_ARG;
  const Setup &s = c->s;
  if (cap < s.nsplin) return PMX_ERR_RANGE;
  if (which >= 0 && which <= 3) {
    if (ipair < 0 || ipair >= s.npair) return PMX_ERR_ARG;
    for (int k = 0; k < s.nsplin; k++) out[k] = s.tab[((size_t)ipair * s.nsplin + k) * 4 + which];
    return PMX_OK;
  }
  if (which == 4 || which == 5) {
    for (int k = 0; k < s.nsplin; k++) out[k] = s.ecor_tab[2 * k + (which - 4)];
    return PMX_OK;
  }
  return PMX_ERR_ARG;
}

int pmx_get_setup_scalars(pmx_ctx *c, double *o) {
  if (!c || !o) return PMX_ERR_ARG;
  const Setup &s = c->s;
  o[0] = s.alp_ewd;
  o[1] = s.ecut;
  o[2] = s.vol;
  o[3] = s.clong;
  o[4] = s.clong_res;
  o[5] = s.self_erf;
  o[6] = s.e_rmin;
  o[7] = s.e_rmax;
  o[8] = s.e_dr;
  o[9] = s.nktot;
  for (int d = 0; d < 3; d++) {
    o[10 + d] = s.ngrid[d];
    o[13 + d] = s.kmaxv[d + 1];
  }
  o[16] = (double)s.e1.size();
  o[17] = s.nexcl;
  o[18] = s.cutoff_max;
  return PMX_OK;
}

int pmx_get_timings(pmx_ctx *c, double *ms5) {
  if (!c || !ms5) return PMX_ERR_ARG;
  for (int k = 0; k < 5; k++) ms5[k] = c->ms[k];
  return PMX_OK;
}

int pmx_get_pair_count(pmx_ctx *c, double *npair) {
  if (!c || !npair) return PMX_ERR_ARG;
  if (c->clu_used) {
    // the cluster kernel does not count: a counting pass over the same rows
    // and positions (the state of the last force call)
    HIPCHK(hipSetDevice(c->device));
    CluArgs a = clu_args(c);
    unsigned long long *cnt = reinterpret_cast<unsigned long long *>(c->clu_flag + 2);
    HIPCHK(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), c->st));
    a.count = cnt;
    HIPCHK(launch_pair_clu(a, pair_grid_x(c->nmol, c->pair_waves), c->st));
    unsigned long long n = 0;
    HIPCHK(hipMemcpyAsync(&n, cnt, sizeof n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    *npair = 0.5 * (double)n;  // the full list holds every pair from both ends
    return PMX_OK;
  }
  double s = 0.0;
  for (int b = 0; b < c->s.P; b++) s += c->acc_h[(size_t)b * ACC_NUM + ACC_NPAIR];
  *npair = s;
  return PMX_OK;
}

int pmx_bond_coef_harm(double fk, double *c7) {
  if (!c7) return PMX_ERR_ARG;
  bond_coef_harm(fk, c7);
  return PMX_OK;
}

int pmx_bond_coef_morse(double d0, double alpha, double *c7) {
  if (!c7) return PMX_ERR_ARG;
  bond_coef_morse(d0, alpha, c7);
  return PMX_OK;
}

int pmx_bend_coef_harm(double fk, double theta0, double *c7, double *s7) {
  if (!c7 || !s7) return PMX_ERR_ARG;
  bend_coef_harm(fk, theta0, c7, s7);
  return PMX_OK;
}

}  // extern "C"
