// channelflow.hip -- host side of the MI355X RK3 timestep: context, device
// memory, launch sequence of one stage, and the C-ABI of include/channelflow.h.
//
// State representation on the device.  The reference keeps every field in
// physical space and transforms on each use (expand/compress/xderiv per level,
// comp1.f90:116-168).  Here u, v, w and their RK registers ru, rv, rw live as
// band spectra (kx < nx/2, |ky| < ny/2), the exact set of modes the reference's
// fields ever hold after deal (les.F90:185-190): compress truncates to that band
// (filter_fftw.f90:300-303), xderiv zeroes the Nyquist terms
// (filter_fftw.f90:182-183), and the pressure keeps no Nyquist modes
// (press.F90:109-110).  The SGS TKE e is clamped pointwise in physical space
// (comp2.f90:38) and so stays physical, with its x-spectral rows kept beside
// it for the next stage's synthesis.  Pressure p is a band spectrum in FP64.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/channelflow.h"
#include "cf_kernels.hpp"
#include "cf_stats.hpp"

using cf::cx;

// ---------------------------------------------------------------------------
// Scratch of the allocation-free entry point cf_calc_spectra_noalloc: a
// file-scope device array of fixed size.  Each GPU the library runs on holds
// its own copy until the process ends; every call on a GPU shares that GPU's
// copy (calls that overlap on one GPU race).  Host code uses the address that
// hipGetSymbolAddress returns for the current device (scratch_ptr), never the
// array's name, which on the host is only a stand-in.
// Layout: k_spectra's ring sums [nlev][3][nb+3], then the normalised outputs
// hats[3][nlev][nb], var[3][nlev], lscales[6][nlev], then the level list.
// ---------------------------------------------------------------------------
constexpr int kSpecMaxLevels = 64;
constexpr int kSpecMaxBins = 512 / 2 + 1;                                      // nx/2 + 1, nx <= 512
constexpr size_t kSpecRing = (size_t)kSpecMaxLevels * 3 * (kSpecMaxBins + 3);
constexpr size_t kSpecOut = (size_t)kSpecMaxLevels * (3 * kSpecMaxBins + 3 + 6);
constexpr size_t kSpecScratch = kSpecRing + kSpecOut + kSpecMaxLevels / 2;   // + nlev int32 levels
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kSpecScratch];

namespace {

thread_local std::string g_create_error;

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) {                                                            \
      ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);                    \
      return CF_E_HIP;                                                                 \
    }                                                                                  \
  } while (0)

constexpr double PI = 3.14159265358979323846264338327950288;

double* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Looks up this device's copy of scratch once per device; the device must be current.
int scratch_setup(int device, std::string& err) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) {
    err = "device ordinal beyond the scratch table";
    return CF_E_ARG;
  }
  if (g_scratch_ptr[device]) return CF_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return CF_E_HIP;
  }
  g_scratch_ptr[device] = (double*)ptr;
  return CF_OK;
}

double* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
}

const char* kKernelNames[CF_NKERN] = {"k_wall",        "k_col_synth", "k_divz_band",    "k_rows",
                                      "k_col_anal",    "k_sgs_cols_in", "k_sgs_rows", "k_sgs_cols_out",
                                      "k_press",       "k_erows",     "k_reduce_part"};

}  // namespace

struct Engine;

// Levels per chunk of k_press_part: its registers hold 20 (FP32) or 32 (FP64)
// VGPRs per level, within 256 at two waves per SIMD.
template <typename T>
constexpr int PRESS_CHUNK = sizeof(T) == 4 ? 8 : 4;
// Modes per workgroup of k_press_part: runs of 256 bytes per level and field
// in a wave's loads and stores (32 FP32 or 16 FP64 modes), and at most 512
// threads, so that its two waves per SIMD fit 256 VGPRs each.
template <typename T>
constexpr int press_modes(int np) {
  return (int)(256 / sizeof(cx<T>)) * np <= 512 ? (int)(256 / sizeof(cx<T>)) : 512 / np;
}

struct cf_ctx {
  cf_params p{};
  int nx = 0, ny = 0, nz = 0, mx = 0, nxb = 0, nyb = 0, slots = 0, zchunk = 16, zchunk_mlp = 32, zchunk_sgs = 32,
      ncu = 256;
  bool fp32 = false;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<void*> allocs;
  int64_t bytes = 0;
  Engine* eng = nullptr;
  double dt = 0;
  // buffers (element type depends on precision; cx<T> / T)
  void *U = nullptr, *V = nullptr, *W = nullptr;
  void *Ruo = nullptr, *Rvo = nullptr, *Rwo = nullptr;
  void *Run = nullptr, *Rvn = nullptr, *Rwn = nullptr;
  void *Ex = nullptr, *Tre = nullptr, *e = nullptr, *re = nullptr;
  void *Bs = nullptr, *Ba = nullptr, *band_tmp = nullptr, *xs_tmp = nullptr;
  void *twNX = nullptr, *twNXB = nullptr, *kxp = nullptr, *kyp = nullptr;
  cx<double>* P = nullptr;
  cx<double>* xs64 = nullptr;
  cx<double>* tw64NX = nullptr;
  double* piv = nullptr;        // k_press reciprocal pivots [slots][modes] (set at create)
  cx<double>* sr = nullptr;     // k_press forward-eliminated rhs [slots][modes]
  int press_np = 0;             // chunks per mode of k_press_part; 0: the serial k_press
  double* M0 = nullptr;         // FP64 (0,0) coefficients of u, v, ru, rv [4][slots]
  cx<double>* band64 = nullptr; // FP32 contexts: FP64 band of the state crossing the boundary
  double* part = nullptr;
  double* red = nullptr;        // [7][slots]: divz and the 6 SGS means
  double* kxd = nullptr;
  double* kyd = nullptr;
  double* staging = nullptr;    // nz planes of nx*ny doubles
  cf::WallScalars* ws = nullptr;
  std::vector<double> sgs_host;  // 6*nz, valid after a mnout stage 1
  bool have_sgs = false;
  // learned closure (k_rows_mlp in place of k_rows when set)
  cf::mlp::Net* closure_net = nullptr;
  unsigned long long* closure_counts = nullptr;   // accepted, evaluated cells
  bool closure_on = false;
  int closure_grid = CF_CLOSURE_GRID_DEALIASED;
  void* sgsS = nullptr;         // small-grid closure: 6 x-spectral input rows (band-sized each)
  void* sgsO = nullptr;         // and its 6 output rows
  double* sred = nullptr;       // its SGS means [6][slots]
  hipStream_t stream2 = nullptr;    // the small-grid closure's columns-in and rows run here,
  hipEvent_t ev_fork = nullptr;     // concurrently with k_col_synth .. k_col_anal
  hipEvent_t ev_join = nullptr;
  bool slot_used[CF_NKERN] = {};    // per-kernel timing: which slots launched work
  double* wpart = nullptr;      // isurface 2..5: wall-plane sums [2][NYB][8]
  int* dirty = nullptr;         // faces whose expanded e was clamped this stage; [0]: any
  void* eclDX = nullptr;        // clamp residual derivatives [nz+1][NYB][NXB]
  void* eclDY = nullptr;
  void* X = nullptr;            // imodel 4: K at faces 0..nnz [nz+1][NYB][NXB]
  bool have_sred = false;       // sred holds the last mnout stage's means
  // box-filter work buffers (allocated on first use)
  double* filt_in = nullptr;    // u, v, w, e physical, 4 * nz*ny*nx
  double* filt_out = nullptr;
  size_t filt_out_n = 0;
  double* stats = nullptr;      // compmn level statistics (cf::NMEAN per level)
};

namespace {

template <typename X>
int dalloc(cf_ctx* ctx, X** ptr, size_t count) {
  void* p = nullptr;
  const size_t nb = count * sizeof(X);
  hipError_t e = hipMalloc(&p, nb);
  if (e != hipSuccess) {
    ctx->err = "hipMalloc(" + std::to_string(nb) + " bytes): " + hipGetErrorString(e);
    return CF_E_HIP;
  }
  e = hipMemset(p, 0, nb);
  if (e != hipSuccess) {
    ctx->err = std::string("hipMemset: ") + hipGetErrorString(e);
    return CF_E_HIP;
  }
  ctx->allocs.push_back(p);
  ctx->bytes += (int64_t)nb;
  *ptr = (X*)p;
  return CF_OK;
}

// Chunks per mode of k_press_part: the smallest of 4, 8, 16, 32 that holds
// ceil(nz / PRESS_CHUNK) chunks.  0 selects the serial k_press: more levels
// than 32 chunks hold, or a wavenumber so small (k^2 < 0.05) that the
// interface system's diagonal dominance no longer bounds its rounding.
template <typename T>
int press_chunks(const cf_ctx* ctx) {
  const int nc = (ctx->nz + PRESS_CHUNK<T> - 1) / PRESS_CHUNK<T>;
  const double kx1 = 2.0 * PI / ctx->p.xl, ky1 = 2.0 * PI / ctx->p.yl;
  if (std::min(kx1 * kx1, ky1 * ky1) < 0.05) return 0;
  for (int np : {4, 8, 16, 32})
    if (nc <= np) return np;
  return 0;
}

template <typename T>
std::vector<cx<T>> twiddles(int n) {
  std::vector<cx<T>> t(n);
  for (int m = 0; m < n; ++m) {
    const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
    t[m] = cx<T>{(T)cosl(a), (T)sinl(a)};
  }
  return t;
}

}  // namespace

// ---------------------------------------------------------------------------
// Engine: precision/size-specialised launch sequences
// ---------------------------------------------------------------------------
struct Engine {
  virtual ~Engine() {}
  virtual int alloc(cf_ctx* ctx) = 0;
  virtual int stage(cf_ctx* ctx, int istage, int mnout, hipEvent_t* ev) = 0;
  virtual int set_state(cf_ctx* ctx, const double* u, const double* v, const double* w,
                        const double* e) = 0;
  // which: 0 u, 1 v, 2 w, 3 e, 4 p, 5 ru, 6 rv, 7 rw, 8 re, 9 run, 10 rvn, 11 rwn
  virtual int get_field(cf_ctx* ctx, int which, double* out) = 0;
  // the same field, physical, left in ctx->staging (enqueued on the stream)
  virtual int physical(cf_ctx* ctx, int which) = 0;
  // d/dx (dir 0) or d/dy (dir 1) of u, v or w (fld 0..2), physical, in ctx->staging
  virtual int physical_deriv(cf_ctx* ctx, int fld, int dir) = 0;
  // compmn level statistics (cf::NMEAN per level) into the device buffer out
  virtual int means(cf_ctx* ctx, double* out) = 0;
  // calc_spectra ring sums for the device level list zl (nlev levels) into out
  virtual int spectra(cf_ctx* ctx, const int* zl, int nlev, double* out) = 0;
};

template <typename T, int NX>
struct EngineT : Engine {
  static constexpr int MX = NX / 2, NY = NX, NXB = 3 * NX / 2, NYB = 3 * NX / 2;
  using C = cx<T>;
  using CS = cf::ColShape<T, NX, NY>;
  using RS = cf::RowShape<T, NX>;

  static C* cptr(void* p) { return (C*)p; }
  static T* rptr(void* p) { return (T*)p; }

  int alloc(cf_ctx* ctx) override {
    const size_t S = ctx->slots;
    const size_t band = S * NY * MX, big = S * NYB * MX, phys = S * NY * NX;
    int rc;
    C* c;
    T* r;
#define AL(field, ty, n)                                      \
  if ((rc = dalloc(ctx, (ty**)&ctx->field, (n))) != CF_OK) return rc;
    AL(U, C, band) AL(V, C, band) AL(W, C, band)
    AL(Ruo, C, band) AL(Rvo, C, band) AL(Rwo, C, band)
    AL(Run, C, band) AL(Rvn, C, band) AL(Rwn, C, band)
    AL(Ex, C, band) AL(Tre, C, band) AL(band_tmp, C, band) AL(xs_tmp, C, band)
    AL(e, T, phys) AL(re, T, phys)
    AL(Bs, C, 8 * big) AL(Ba, C, 8 * big)
    AL(twNX, C, NX) AL(twNXB, C, NXB) AL(kxp, T, MX) AL(kyp, T, NY)
#undef AL
    (void)c;
    (void)r;
    if ((rc = dalloc(ctx, &ctx->P, band)) != CF_OK) return rc;
    if ((rc = dalloc(ctx, &ctx->xs64, band)) != CF_OK) return rc;
    if ((rc = dalloc(ctx, &ctx->tw64NX, (size_t)NX)) != CF_OK) return rc;
    ctx->press_np = press_chunks<T>(ctx);
    if (!ctx->press_np) {                                   // the serial sweep's tables
      if ((rc = dalloc(ctx, &ctx->piv, S * NY * MX)) != CF_OK) return rc;
      if ((rc = dalloc(ctx, &ctx->sr, S * NY * MX)) != CF_OK) return rc;
    }
    if ((rc = dalloc(ctx, &ctx->M0, 4 * S)) != CF_OK) return rc;
    if ((rc = dalloc(ctx, &ctx->part, 7 * S * NYB)) != CF_OK) return rc;
    if ((rc = dalloc(ctx, &ctx->red, 7 * S)) != CF_OK) return rc;
    if ((rc = dalloc(ctx, &ctx->kxd, (size_t)MX)) != CF_OK) return rc;
    if ((rc = dalloc(ctx, &ctx->kyd, (size_t)NY)) != CF_OK) return rc;
    if ((rc = dalloc(ctx, &ctx->staging, (size_t)ctx->nz * NY * NX)) != CF_OK) return rc;
    if ((rc = dalloc(ctx, &ctx->ws, (size_t)1)) != CF_OK) return rc;
    if (ctx->p.isurface >= 2 && ctx->p.isurface <= 5)      // wall-plane statistics
      if ((rc = dalloc(ctx, &ctx->wpart, (size_t)2 * NYB * 8)) != CF_OK) return rc;
    if ((rc = dalloc(ctx, &ctx->dirty, (size_t)ctx->nz + 3)) != CF_OK) return rc;
    if ((rc = dalloc(ctx, (T**)&ctx->eclDX, (size_t)(ctx->nz + 1) * NYB * NXB)) != CF_OK) return rc;
    if ((rc = dalloc(ctx, (T**)&ctx->eclDY, (size_t)(ctx->nz + 1) * NYB * NXB)) != CF_OK) return rc;
    if (ctx->p.imodel == 4)                                 // K at faces 0..nnz on the 3/2 grid
      if ((rc = dalloc(ctx, (T**)&ctx->X, (size_t)(ctx->nz + 1) * NYB * NXB)) != CF_OK) return rc;
    // tables
    auto t1 = twiddles<T>(NX), t2 = twiddles<T>(NXB);
    auto t3 = twiddles<double>(NX);
    HIPCHK(hipMemcpy(ctx->twNX, t1.data(), NX * sizeof(C), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->twNXB, t2.data(), NXB * sizeof(C), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->tw64NX, t3.data(), NX * sizeof(cx<double>), hipMemcpyHostToDevice));
    std::vector<double> kx(MX), ky(NY);
    for (int k = 0; k < MX; ++k) kx[k] = (double)k * 2.0 * PI / ctx->p.xl;   // init.f90:24-27
    for (int j = 0; j < NY; ++j) {
      const int kk = (j < NY / 2) ? j : (j == NY / 2 ? 0 : j - NY);
      ky[j] = (double)kk * 2.0 * PI / ctx->p.yl;                            // init.f90:28-31
    }
    std::vector<T> kxt(kx.begin(), kx.end()), kyt(ky.begin(), ky.end());
    HIPCHK(hipMemcpy(ctx->kxp, kxt.data(), MX * sizeof(T), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->kyp, kyt.data(), NY * sizeof(T), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->kxd, kx.data(), MX * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->kyd, ky.data(), NY * sizeof(double), hipMemcpyHostToDevice));
    // the serial sweep's reciprocal pivots depend on k^2 and dz only
    if (!ctx->press_np) {
      hipLaunchKernelGGL(cf::k_press_pivots, dim3((MX * NY + 63) / 64), dim3(64), 0, ctx->stream, ctx->piv,
                         ctx->kxd, ctx->kyd, MX, NY, ctx->nz, ctx->p.zl / ctx->nz);
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CF_OK;
  }

  cf::Consts consts(cf_ctx* ctx, int istage, int mnout) {
    static const double gama[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};   // les.F90:27
    static const double zetas[3] = {0.0, 17.0 / 60.0, 5.0 / 12.0};      // les.F90:26
    const cf_params& p = ctx->p;
    cf::Consts c{};
    const double dx = p.xl / ctx->nx, dy = p.yl / ctx->ny, dz = p.zl / ctx->nz;
    c.dz = dz;
    c.dzm = 1.0 / dz;
    c.dzm2 = c.dzm * c.dzm;
    c.todzm = 0.5 / dz;
    const double dsl = std::pow(dx * dy * dz, 1.0 / 3.0);                // les.F90:118
    c.alk = p.alk > 0 ? p.alk : dsl;
    c.dsl = dsl;
    c.sgs_momentum = (ctx->closure_on && ctx->closure_grid == CF_CLOSURE_GRID_SMALL) ? 0 : 1;
    c.isurface = ctx->p.isurface;
    c.imodel = ctx->p.imodel;
    c.xkmax = dz * dz / ctx->dt / 5.0;                                  // les.F90:216
    c.ceps = 0.93 / dsl;                                                // les.F90:120
    c.emin = 1.0e-6;                                                    // les.F90:102
    c.ugal = p.ugal;
    c.vgal = p.vgal;
    c.dtab1 = ctx->dt * gama[istage - 1];
    c.dtab2 = ctx->dt * zetas[istage - 1];
    c.dtabm = 1.0 / c.dtab1;
    c.xymb = 1.0 / (double)NXB / (double)NYB;
    c.zody = std::log(dz / 2.0 / p.z0);                                 // les.F90:110-111
    c.vk = 0.4;
    c.zl = p.zl;
    c.nz = ctx->nz;
    c.istage = istage;
    c.mnout = mnout;
    c.zchunk = ctx->zchunk;
    return c;
  }

  int stage(cf_ctx* ctx, int istage, int mnout, hipEvent_t* ev) override {
    hipStream_t s = ctx->stream;
    const int nz = ctx->nz;
    const bool sgs = mnout && istage == 1;
    const cf::Consts c = consts(ctx, istage, sgs ? 1 : 0);
    const size_t bstride = (size_t)ctx->slots * NYB * MX;
    int ek = 0;
    // records the event that closes timing slot ek-1 (used: whether that slot
    // launched work on this path) and opens slot ek
    auto mark = [&](bool used) {
      if (!ev) return;
      if (ek > 0) ctx->slot_used[ek - 1] = used;
      (void)hipEventRecord(ev[ek++], s);
    };
    // The small-grid closure's divz, columns-in and row kernel depend only on
    // the stage-start state and write only their own buffers: they run on the
    // second stream beside k_rows .. k_col_anal (serially when per-kernel
    // events are being recorded), and k_sgs_cols_out joins after k_col_anal.
    bool small = false;
    if constexpr (NX % 64 == 0) small = ctx->closure_on && ctx->closure_grid == CF_CLOSURE_GRID_SMALL;
    const bool overlap = small && !ev;
    mark(false);
    hipLaunchKernelGGL(cf::k_wall, dim3(1), dim3(64), 0, s, ctx->M0, ctx->slots, c, ctx->ws, ctx->dirty,
                       nz + 3, ctx->red, ctx->slots);
    mark(true);
    {
      cf::SynthArgs<T> a{};
      a.S[0] = cptr(ctx->U);
      a.S[1] = cptr(ctx->V);
      a.S[2] = cptr(ctx->W);
      a.Ex = cptr(ctx->Ex);
      a.B = cptr(ctx->Bs);
      a.bstride = bstride;
      a.kyp = rptr(ctx->kyp);
      a.twNY = cptr(ctx->twNX);
      a.twNYB = cptr(ctx->twNXB);
      a.nz = nz;
      a.lean = small ? 1 : 0;
      hipLaunchKernelGGL((cf::k_col_synth<T, NX, NY>), dim3(MX / CS::TKC, nz, 4), dim3(CS::NT), 0,
                         s, a);
    }
    if (overlap) {                 // forked after k_col_synth: the closure chain overlaps k_rows
      HIPCHK(hipEventRecord(ctx->ev_fork, s));
      HIPCHK(hipStreamWaitEvent(ctx->stream2, ctx->ev_fork, 0));
      if constexpr (NX % 64 == 0) {
        int rc = divz_band(ctx, c.dzm, ctx->stream2);
        if (rc == CF_OK) rc = sgs_cols_in(ctx, ctx->stream2);
        if (rc == CF_OK) rc = sgs_rows(ctx, c, sgs, ctx->stream2);
        if (rc != CF_OK) return rc;
      }
      HIPCHK(hipEventRecord(ctx->ev_join, ctx->stream2));
    }
    if (ctx->p.isurface >= 2 && ctx->p.isurface <= 5) {    // wall-plane statistics, then the wall coefficients
      hipLaunchKernelGGL((cf::k_wall_plane<T, NX>), dim3(NYB, 2), dim3(RS::NT), 0, s, cptr(ctx->Bs), bstride,
                         rptr(ctx->kxp), cptr(ctx->twNXB), nz, ctx->ws, ctx->wpart);
      hipLaunchKernelGGL(cf::k_wall_final, dim3(1), dim3(256), 0, s, ctx->wpart, NYB, (double)NYB * NXB, c,
                         ctx->ws);
    }
    if (ctx->p.imodel == 4) {                               // K from the face shear
      cf::Consts cs = c;
      cs.zchunk = ctx->zchunk;
      const int nchs = (nz + 1 + ctx->zchunk - 1) / ctx->zchunk;
      hipLaunchKernelGGL((cf::k_shear<T, NX>), dim3(NYB, nchs), dim3(RS::NT), 0, s, cptr(ctx->Bs), bstride,
                         rptr(ctx->kxp), cptr(ctx->twNXB), cs, rptr(ctx->X));
    }
    mark(true);
    if (!overlap) {               // divz (comp1.f90:145) by Parseval from the band spectra
      int rc = divz_band(ctx, c.dzm, s);
      if (rc != CF_OK) return rc;
    }
    mark(true);
    {
      cf::RowArgs<T> a{};
      a.B = cptr(ctx->Bs);
      a.bstride = bstride;
      a.A = cptr(ctx->Ba);
      a.astride = bstride;
      a.kxp = rptr(ctx->kxp);
      a.twNXB = cptr(ctx->twNXB);
      a.ws = ctx->ws;
      a.part = ctx->part;
      a.pstride = (size_t)ctx->slots * NYB;
      a.flags = &ctx->ws->flags;
      a.X = rptr(ctx->X);
      a.dirty = ctx->dirty;
      a.c = c;
      bool learned = false;
      if constexpr ((3 * NX / 2) % 64 == 0) {
        if (ctx->closure_on && ctx->closure_grid == CF_CLOSURE_GRID_DEALIASED) {
          // longer chunks: each chunk re-evaluates the network at two extra levels
          a.c.zchunk = ctx->zchunk_mlp;
          const int nchm = (nz + ctx->zchunk_mlp - 1) / ctx->zchunk_mlp;
          hipLaunchKernelGGL((cf::k_rows_mlp<T, NX>), dim3(NYB, nchm), dim3(RS::NT), 0, s, a,
                             ctx->closure_net, ctx->closure_counts);
          learned = true;
        }
      }
      // 64-level chunks for the nnx = 512 sweeps from nnz = 128: each chunk
      // first synthesises its window's pairs, 5 % of a 16-level chunk's
      // transforms (512^2 x 256: the lean sweep 5.48 -> 5.33 ms per launch,
      // learned step 52.8 -> 51.7 ms; FP32 TKE 32.2 -> 31.7 ms, FP64 67.4 ->
      // 66.1 ms).  At nnx 256 16 levels stay (32 measured slower).
      const int zc = (NX == 512 && nz >= 128) ? 64 : ctx->zchunk;
      a.c.zchunk = zc;
      const int nch = (nz + zc - 1) / zc;
      if (!learned) {
        if (small) {
          if constexpr (sizeof(T) == 4 && NX == 512)   // two points per thread, wave-owned transforms
            hipLaunchKernelGGL((cf::k_rows_lean2<T, NX>), dim3(NYB, nch), dim3(RS::NT / 2), 0, s, a);
          else
            hipLaunchKernelGGL((cf::k_rows<T, NX, true>), dim3(NYB, nch), dim3(RS::NT), 0, s, a);
        } else {
          hipLaunchKernelGGL((cf::k_rows<T, NX, false>), dim3(NYB, nch), dim3(RS::NT), 0, s, a);
        }
      }
      // the comp1 clamp of the expanded e, where it fired (no work otherwise)
      cf::EclampArgs<T> ea{};
      ea.B = cptr(ctx->Bs);
      ea.bstride = bstride;
      ea.A = cptr(ctx->Ba);
      ea.astride = bstride;
      ea.dirty = ctx->dirty;
      ea.DX = rptr(ctx->eclDX);
      ea.DY = rptr(ctx->eclDY);
      ea.X = rptr(ctx->X);
      ea.kxp = rptr(ctx->kxp);
      ea.twNXB = cptr(ctx->twNXB);
      ea.kxb = 2.0 * PI / ctx->p.xl;
      ea.kyb = 2.0 * PI / ctx->p.yl;
      ea.c = c;
      hipLaunchKernelGGL((cf::k_eclamp_rows<T, NX>), dim3(NYB), dim3(RS::NT), 0, s, ea);
      hipLaunchKernelGGL((cf::k_eclamp_cols<T, NX>), dim3(NXB / 16), dim3(256), 0, s, ea);
      hipLaunchKernelGGL((cf::k_eclamp_flux<T, NX>), dim3(NYB), dim3(RS::NT), 0, s, ea);
    }
    mark(true);
    {
      cf::AnalArgs<T> a{};
      a.A = cptr(ctx->Ba);
      a.astride = bstride;
      a.R[0] = cptr(ctx->Run);
      a.R[1] = cptr(ctx->Rvn);
      a.R[2] = cptr(ctx->Rwn);
      a.Tre = cptr(ctx->Tre);
      a.kyp = rptr(ctx->kyp);
      a.twNY = cptr(ctx->twNX);
      a.twNYB = cptr(ctx->twNXB);
      a.nz = nz;
      const int ntiles = (MX / CS::TKC) * nz * 4;     // tiles: kx fastest, then level, then component
      hipLaunchKernelGGL((cf::k_col_anal<T, NX, NY>), dim3(ntiles), dim3(CS::NT), 0, s, a);
    }
    mark(true);
    if constexpr (NX % 64 == 0) {
      if (small && !overlap) {
        int rc = sgs_cols_in(ctx, s);
        if (rc != CF_OK) return rc;
      }
    }
    mark(small);
    if constexpr (NX % 64 == 0) {
      if (small) {
        if (overlap) {
          HIPCHK(hipStreamWaitEvent(s, ctx->ev_join, 0));
        } else {
          int rc = sgs_rows(ctx, c, sgs, s);
          if (rc != CF_OK) return rc;
        }
      }
    }
    mark(small);
    if constexpr (NX % 64 == 0) {
      if (small) {
        int rc = sgs_small_out(ctx, s);
        if (rc != CF_OK) return rc;
      }
    }
    mark(small);
    {
      cf::PressArgs<T> a{};
      a.U[0] = cptr(ctx->U);
      a.U[1] = cptr(ctx->V);
      a.U[2] = cptr(ctx->W);
      a.Ro[0] = cptr(ctx->Ruo);
      a.Ro[1] = cptr(ctx->Rvo);
      a.Ro[2] = cptr(ctx->Rwo);
      a.Rn[0] = cptr(ctx->Run);
      a.Rn[1] = cptr(ctx->Rvn);
      a.Rn[2] = cptr(ctx->Rwn);
      a.P = ctx->P;
      a.piv = ctx->piv;
      a.sr = ctx->sr;
      a.M0 = ctx->M0;
      a.slots = ctx->slots;
      a.kxd = ctx->kxd;
      a.kyd = ctx->kyd;
      a.mx = MX;
      a.ny = NY;
      a.nz = nz;
      a.istage = istage;
      a.dz = c.dz;
      a.dzm = c.dzm;
      a.dtab1 = c.dtab1;
      a.dtab2 = c.dtab2;
      a.dtabm = c.dtabm;
      const int modes = MX * NY;
      if (ctx->press_np) {
        // 16 modes x press_np chunks per workgroup, and one more for mode (0,0)
        constexpr int CH = PRESS_CHUNK<T>;
        const int np = ctx->press_np, mw = press_modes<T>(np);
        const dim3 g((modes + mw - 1) / mw + 1), b(mw * np);
        switch (np) {
          case 4: hipLaunchKernelGGL((cf::k_press_part<T, CH, 4, press_modes<T>(4)>), g, b, 0, s, a); break;
          case 8: hipLaunchKernelGGL((cf::k_press_part<T, CH, 8, press_modes<T>(8)>), g, b, 0, s, a); break;
          case 16: hipLaunchKernelGGL((cf::k_press_part<T, CH, 16, press_modes<T>(16)>), g, b, 0, s, a); break;
          default: hipLaunchKernelGGL((cf::k_press_part<T, CH, 32, press_modes<T>(32)>), g, b, 0, s, a); break;
        }
      } else {
        // one lane per mode, and one more workgroup for mode (0,0)
        hipLaunchKernelGGL(cf::k_press<T>, dim3((modes + 63) / 64 + 1), dim3(64), 0, s, a);
      }
    }
    mark(true);
    {
      cf::ERowArgs<T> a{};
      a.e = rptr(ctx->e);
      a.re = rptr(ctx->re);
      a.Tre = cptr(ctx->Tre);
      a.Ex = cptr(ctx->Ex);
      a.twNX = cptr(ctx->twNX);
      a.flags = &ctx->ws->flags;
      a.nz = nz;
      a.istage = istage;
      a.mode = 0;
      a.dtab1 = c.dtab1;
      a.dtab2 = c.dtab2;
      a.emin = c.emin;
      hipLaunchKernelGGL((cf::k_erows<T, NX>), dim3(NY / 2, nz), dim3(MX), 0, s, a);
    }
    mark(true);
    {
      const int q0 = 1, nq = sgs ? 7 : 1;        // the SGS means (slot 0, divz: k_divz_band)
      if (nq > q0)
        hipLaunchKernelGGL(cf::k_reduce_part, dim3(ctx->slots), dim3(256), 0, s, ctx->part,
                           (size_t)ctx->slots * NYB, NYB, q0, nq, c.xymb, ctx->red, ctx->slots);
    }
    mark(sgs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      ctx->err = std::string("stage launch: ") + hipGetErrorString(e);
      return CF_E_HIP;
    }
    if (sgs) {
      ctx->have_sgs = true;
      ctx->have_sred = ctx->closure_on && ctx->closure_grid == CF_CLOSURE_GRID_SMALL;
    }
    return CF_OK;
  }

  // FP32 contexts move the state across the boundary through FP64
  // transforms: the FP32 transforms' rounding is relative to a level's
  // largest coefficient (the mean of u), which would add white noise to the
  // much smaller fluctuations.  band64 is allocated on first use.
  int need_band64(cf_ctx* ctx) {
    if (ctx->band64) return CF_OK;
    return dalloc(ctx, &ctx->band64, (size_t)ctx->slots * NY * MX);
  }

  // physical (host, nz planes) -> band spectrum slots 1..nz of dst
  int to_band(cf_ctx* ctx, const double* host, C* dst) {
    const int nz = ctx->nz;
    HIPCHK(hipMemcpyAsync(ctx->staging, host, (size_t)nz * NY * NX * sizeof(double),
                          hipMemcpyHostToDevice, ctx->stream));
    if constexpr (sizeof(T) == 4) {
      int rc = need_band64(ctx);
      if (rc != CF_OK) return rc;
      constexpr int TK64 = cf::ColShape<double, NX, NY>::TK;
      const size_t plane = (size_t)NY * MX;
      hipLaunchKernelGGL((cf::k_rows_fwd<double, NX>), dim3(NY / 2, nz), dim3(MX), 0, ctx->stream,
                         ctx->staging, ctx->xs64, ctx->tw64NX, 1);
      hipLaunchKernelGGL((cf::k_cols_small<double, NX, NY, false>), dim3(MX / TK64, nz), dim3(256), 0,
                         ctx->stream, ctx->xs64, ctx->band64, ctx->tw64NX, 1);
      const size_t nd = 2 * (size_t)nz * plane;
      hipLaunchKernelGGL((cf::k_cast<double, T>), dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, ctx->stream,
                         (const double*)(ctx->band64 + plane), (T*)(dst + plane), nd);
      HIPCHK(hipGetLastError());
      return CF_OK;
    }
    hipLaunchKernelGGL((cf::k_rows_fwd<T, NX>), dim3(NY / 2, nz), dim3(MX), 0, ctx->stream,
                       ctx->staging, cptr(ctx->xs_tmp), cptr(ctx->twNX), 1);
    hipLaunchKernelGGL((cf::k_cols_small<T, NX, NY, false>), dim3(MX / CS::TK, nz), dim3(256), 0,
                       ctx->stream, cptr(ctx->xs_tmp), dst, cptr(ctx->twNX), 1);
    HIPCHK(hipGetLastError());
    return CF_OK;
  }

  // band spectrum slots 1..nz -> staging (physical doubles)
  template <typename Q>
  int from_band(cf_ctx* ctx, const cx<Q>* src, cx<Q>* xs, const cx<Q>* tw) {
    const int nz = ctx->nz;
    constexpr int TKQ = cf::ColShape<Q, NX, NY>::TK;
    hipLaunchKernelGGL((cf::k_cols_small<Q, NX, NY, true>), dim3(MX / TKQ, nz), dim3(256), 0,
                       ctx->stream, src, xs, tw, 1);
    hipLaunchKernelGGL((cf::k_rows_inv<Q, NX>), dim3(NY / 2, nz), dim3(MX), 0, ctx->stream, xs,
                       ctx->staging, tw, 1);
    HIPCHK(hipGetLastError());
    return CF_OK;
  }

  // divz (comp1.f90:145) on every path: the plane mean of (ux + vy + wz)^2 on
  // the 3/2 grid, by Parseval from the stage-start band spectra (the square of
  // a band-limited field is resolved there, so the two agree to rounding);
  // into the q = 0 slots of red.
  int divz_band(cf_ctx* ctx, double dzm, hipStream_t s) {
    // the q = 0 slots of red were cleared by the stage's k_wall
    const int parts = std::max(1, std::min(16, (MX * NY) / 2048));
    hipLaunchKernelGGL(cf::k_divz_band<T>, dim3(ctx->nz, parts), dim3(256), 0, s, cptr(ctx->U), cptr(ctx->V),
                       cptr(ctx->W), rptr(ctx->kxp), rptr(ctx->kyp), MX, NY, ctx->nz, dzm, ctx->red);
    HIPCHK(hipGetLastError());
    return CF_OK;
  }

  // The learned closure on the nnx x nny grid: sgs_cols_in (the y-inverse of
  // U, V, W and their y-derivatives to x-spectral small rows), sgs_rows (the
  // row kernel: network, TKE fallback, divergence rows) and sgs_small_out
  // (the y-forward of the divergence rows, added to the comp1 tendencies).
  int sgs_cols_in(cf_ctx* ctx, hipStream_t s) {
    constexpr int TK = CS::TKS;
    cf::SgsColsInArgs<T> a{};
    a.U[0] = cptr(ctx->U);
    a.U[1] = cptr(ctx->V);
    a.U[2] = cptr(ctx->W);
    a.S = cptr(ctx->sgsS);
    a.sstride = (size_t)ctx->slots * NY * MX;
    a.kyp = rptr(ctx->kyp);
    a.twNY = cptr(ctx->twNX);
    a.nz = ctx->nz;
    hipLaunchKernelGGL((cf::k_sgs_cols_in<T, NX, NY>), dim3(MX / TK, ctx->nz, 3), dim3(512), 0, s, a);
    HIPCHK(hipGetLastError());
    return CF_OK;
  }

  int sgs_rows(cf_ctx* ctx, const cf::Consts& c, bool sgs, hipStream_t s) {
    const int nz = ctx->nz;
    const size_t S = ctx->slots;
    const size_t fstride = S * NY * MX;
    if (sgs) HIPCHK(hipMemsetAsync(ctx->sred, 0, 6 * S * sizeof(double), s));
    cf::SgsRowArgs<T> a{};
    C* in = cptr(ctx->sgsS);
    a.S = in;
    a.sstride = fstride;
    a.O = cptr(ctx->sgsO);
    a.e = rptr(ctx->e);
    a.kxp = rptr(ctx->kxp);
    a.twNX = cptr(ctx->twNX);
    a.sred = ctx->sred;
    a.slots = (int)S;
    a.counts = ctx->closure_counts;
    a.c = c;
    a.c.zchunk = ctx->zchunk_sgs;
    const int nch = (nz + ctx->zchunk_sgs - 1) / ctx->zchunk_sgs;
    hipLaunchKernelGGL((cf::k_sgs_rows<T, NX>), dim3(NY, nch), dim3(NX), 0, s, a, ctx->closure_net);
    HIPCHK(hipGetLastError());
    return CF_OK;
  }

  // Second part: the y-forward of the divergence rows added to the comp1
  // tendencies (after k_col_anal has written them).
  int sgs_small_out(cf_ctx* ctx, hipStream_t s) {
    const int nz = ctx->nz;
    constexpr int TK = CS::TKS;
    cf::SgsColsOutArgs<T> b{};
    b.O = cptr(ctx->sgsO);
    b.ostride = (size_t)ctx->slots * NY * MX;
    b.R[0] = cptr(ctx->Run);
    b.R[1] = cptr(ctx->Rvn);
    b.R[2] = cptr(ctx->Rwn);
    b.kyp = rptr(ctx->kyp);
    b.twNY = cptr(ctx->twNX);
    b.nz = nz;
    hipLaunchKernelGGL((cf::k_sgs_cols_out<T, NX, NY>), dim3(MX / TK, nz, 3), dim3(512), 0, s, b);
    HIPCHK(hipGetLastError());
    return CF_OK;
  }

  int set_state(cf_ctx* ctx, const double* u, const double* v, const double* w,
                const double* e) override {
    const int nz = ctx->nz;
    const size_t S = ctx->slots;
    const size_t bandb = S * NY * MX * sizeof(C);
    hipStream_t s = ctx->stream;
    int rc;
    if ((rc = to_band(ctx, u, cptr(ctx->U))) != CF_OK) return rc;
    if ((rc = to_band(ctx, v, cptr(ctx->V))) != CF_OK) return rc;
    if ((rc = to_band(ctx, w, cptr(ctx->W))) != CF_OK) return rc;
    // w(nnz) is the top wall (upperbc.f90:26)
    HIPCHK(hipMemsetAsync(cptr(ctx->W) + (size_t)nz * NY * MX, 0, NY * MX * sizeof(C), s));
    // the plane means of u and v in FP64 (the (0,0) coefficients; deal keeps
    // them), the RK registers' means zero
    std::vector<double> m0(4 * S, 0.0);
    const size_t nxy = (size_t)NY * NX;
    for (int z = 1; z <= nz; ++z) {
      long double su = 0.0L, sv = 0.0L;
      const double* uz = u + (size_t)(z - 1) * nxy;
      const double* vz = v + (size_t)(z - 1) * nxy;
      for (size_t i = 0; i < nxy; ++i) {
        su += uz[i];
        sv += vz[i];
      }
      m0[cf::M0_U * S + z] = (double)(su / (long double)nxy);
      m0[cf::M0_V * S + z] = (double)(sv / (long double)nxy);
    }
    HIPCHK(hipMemcpyAsync(ctx->M0, m0.data(), m0.size() * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((cf::k_set_mean<T>), dim3((nz + 63) / 64), dim3(64), 0, s, cptr(ctx->U), cptr(ctx->V),
                       ctx->M0, (int)S, (size_t)NY * MX, nz);
    // e: deal in spectral space, back to physical, keep its x-spectral rows
    if ((rc = to_band(ctx, e, cptr(ctx->band_tmp))) != CF_OK) return rc;
    hipLaunchKernelGGL((cf::k_cols_small<T, NX, NY, true>), dim3(MX / CS::TK, nz), dim3(256), 0, s,
                       cptr(ctx->band_tmp), cptr(ctx->Ex), cptr(ctx->twNX), 1);
    hipLaunchKernelGGL((cf::k_rows_inv<T, NX>), dim3(NY / 2, nz), dim3(MX), 0, s, cptr(ctx->Ex),
                       ctx->staging, cptr(ctx->twNX), 1);
    const size_t n = (size_t)nz * NY * NX;
    hipLaunchKernelGGL((cf::k_cast<double, T>), dim3((n + 255) / 256), dim3(256), 0, s,
                       ctx->staging, rptr(ctx->e) + (size_t)NY * NX, n);
    // registers and pressure start at zero
    for (void* b : {ctx->Ruo, ctx->Rvo, ctx->Rwo, ctx->Run, ctx->Rvn, ctx->Rwn, ctx->Tre})
      HIPCHK(hipMemsetAsync(b, 0, bandb, s));
    HIPCHK(hipMemsetAsync(ctx->re, 0, S * NY * NX * sizeof(T), s));
    HIPCHK(hipMemsetAsync(ctx->P, 0, S * NY * MX * sizeof(cx<double>), s));
    HIPCHK(hipMemsetAsync(ctx->red, 0, 7 * S * sizeof(double), s));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    ctx->have_sgs = false;
    ctx->have_sred = false;
    return CF_OK;
  }

  int get_field(cf_ctx* ctx, int which, double* out) override {
    const size_t n = (size_t)ctx->nz * NY * NX;
    int rc = physical(ctx, which);
    if (rc != CF_OK) return rc;
    HIPCHK(hipMemcpyAsync(out, ctx->staging, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CF_OK;
  }

  int means(cf_ctx* ctx, double* out) override {
    hipLaunchKernelGGL((cf::k_compmn<T>), dim3(ctx->nz), dim3(256), 0, ctx->stream, cptr(ctx->U),
                       cptr(ctx->V), cptr(ctx->W), rptr(ctx->e), ctx->M0 + cf::M0_U * ctx->slots,
                       ctx->M0 + cf::M0_V * ctx->slots, MX, NY, ctx->nz, out);
    HIPCHK(hipGetLastError());
    return CF_OK;
  }

  int spectra(cf_ctx* ctx, const int* zl, int nlev, double* out) override {
    const size_t lds = (size_t)(NX / 2 + 4) * sizeof(double);
    hipLaunchKernelGGL((cf::k_spectra<T>), dim3(nlev, 3), dim3(256), lds, ctx->stream, cptr(ctx->U),
                       cptr(ctx->V), cptr(ctx->W), zl, MX, NY, NX, out);
    HIPCHK(hipGetLastError());
    return CF_OK;
  }

  int physical_deriv(cf_ctx* ctx, int fld, int dir) override {
    const size_t n = (size_t)ctx->nz * NY * MX;
    void* src = fld == 0 ? ctx->U : (fld == 1 ? ctx->V : ctx->W);
    hipLaunchKernelGGL((cf::k_band_deriv<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       cptr(src), cptr(ctx->band_tmp), rptr(ctx->kxp), rptr(ctx->kyp), MX, NY, ctx->nz, dir);
    HIPCHK(hipGetLastError());
    return from_band<T>(ctx, cptr(ctx->band_tmp), cptr(ctx->xs_tmp), cptr(ctx->twNX));
  }

  int physical(cf_ctx* ctx, int which) override {
    const int nz = ctx->nz;
    hipStream_t s = ctx->stream;
    const size_t n = (size_t)nz * NY * NX;
    int rc = CF_OK;
    void* bands[12] = {ctx->U,   ctx->V,   ctx->W,   nullptr, nullptr, ctx->Ruo,
                       ctx->Rvo, ctx->Rwo, nullptr,  ctx->Run, ctx->Rvn, ctx->Rwn};
    if (which == 3 || which == 8) {
      T* src = rptr(which == 3 ? ctx->e : ctx->re) + (size_t)NY * NX;
      hipLaunchKernelGGL((cf::k_cast<T, double>), dim3((n + 255) / 256), dim3(256), 0, s, src,
                         ctx->staging, n);
      HIPCHK(hipGetLastError());
    } else if (which == 4) {
      rc = from_band<double>(ctx, ctx->P, ctx->xs64, ctx->tw64NX);
    } else {
      // u, v, ru, rv: the plane means from their FP64 copies; the comp1
      // tendency of u with the imposed gradient, which k_press adds (comp1.f90:152)
      const int q = which == 0 ? cf::M0_U : which == 1 ? cf::M0_V : which == 5 ? cf::M0_RU
                                                                   : which == 6 ? cf::M0_RV : -1;
      const double add = which == 9 ? cf::FORCING : 0.0;
      const double* m0 = q >= 0 ? ctx->M0 + (size_t)q * ctx->slots : nullptr;
      const size_t plane = (size_t)NY * MX;
      if constexpr (sizeof(T) == 4) {           // FP64 synthesis of the FP32 band
        if ((rc = need_band64(ctx)) != CF_OK) return rc;
        const size_t nd = 2 * (size_t)ctx->slots * plane;
        hipLaunchKernelGGL((cf::k_cast<T, double>), dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, s,
                           (const T*)bands[which], (double*)ctx->band64, nd);
        rc = from_band<double>(ctx, ctx->band64, ctx->xs64, ctx->tw64NX);
        if (rc == CF_OK && (m0 || add != 0.0))
          hipLaunchKernelGGL((cf::k_shift_levels<double>), dim3(nz), dim3(256), 0, s, ctx->staging, m0,
                             ctx->band64, plane, (size_t)NY * NX, add);
      } else {
        rc = from_band<T>(ctx, cptr(bands[which]), cptr(ctx->xs_tmp), cptr(ctx->twNX));
        if (rc == CF_OK && (m0 || add != 0.0))
          hipLaunchKernelGGL((cf::k_shift_levels<T>), dim3(nz), dim3(256), 0, s, ctx->staging, m0,
                             cptr(bands[which]), plane, (size_t)NY * NX, add);
      }
      HIPCHK(hipGetLastError());
    }
    return rc;
  }
};

template <typename T>
Engine* make_engine_t(int nx) {
  switch (nx) {
    case 32: return new EngineT<T, 32>();
    case 64: return new EngineT<T, 64>();
    case 128: return new EngineT<T, 128>();
    case 256: return new EngineT<T, 256>();
    case 512: return new EngineT<T, 512>();
    default: return nullptr;
  }
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
namespace {

int check_flags(cf_ctx* ctx) {
  int flags = 0;
  HIPCHK(hipMemcpy(&flags, &ctx->ws->flags, sizeof(int), hipMemcpyDeviceToHost));
  if (flags & cf::FLAG_UTAU) {
    ctx->err = "utau > 10: wall-model guard (lowerbc.f90:77-80)";
    return CF_E_UTAU;
  }
  if (flags & cf::FLAG_NONFINITE) {
    ctx->err = "non-finite SGS TKE";
    return CF_E_NONFINITE;
  }
  return CF_OK;
}

}  // namespace

extern "C" {

int cf_create(const cf_params* p, cf_ctx** out) {
  if (!p || !out) {
    g_create_error = "null argument";
    return CF_E_ARG;
  }
  *out = nullptr;
  const int nx = p->nnx, ny = p->nny, nz = p->nnz;
  if (nx != ny || !(nx == 32 || nx == 64 || nx == 128 || nx == 256 || nx == 512)) {
    g_create_error = "nnx == nny in {32,64,128,256,512} required";
    return CF_E_ARG;
  }
  if (nz < 4) {
    g_create_error = "nnz >= 4 required";
    return CF_E_ARG;
  }
  if ((p->imodel != 3 && p->imodel != 4) || p->isurface < 1 || p->isurface > 6) {
    g_create_error = "imodel must be 3 or 4 and isurface 1..6";
    return CF_E_ARG;
  }
  if (!(p->dt > 0) || !(p->xl > 0) || !(p->yl > 0) || !(p->zl > 0) || !(p->z0 > 0)) {
    g_create_error = "dt, xl, yl, zl, z0 must be positive";
    return CF_E_ARG;
  }
  if (p->precision != CF_PREC_FP64 && p->precision != CF_PREC_FP32) {
    g_create_error = "precision must be CF_PREC_FP64 or CF_PREC_FP32";
    return CF_E_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= p->device || p->device < 0) {
    g_create_error = "no HIP device";
    return CF_E_NODEV;
  }
  cf_ctx* ctx = new cf_ctx();
  ctx->p = *p;
  ctx->nx = nx;
  ctx->ny = ny;
  ctx->nz = nz;
  ctx->mx = nx / 2;
  ctx->nxb = 3 * nx / 2;
  ctx->nyb = 3 * ny / 2;
  ctx->slots = nz + 2;
  ctx->fp32 = p->precision == CF_PREC_FP32;
  ctx->dt = p->dt;
  ctx->zchunk = nz >= 64 ? 16 : 8;
  ctx->zchunk_mlp = nz >= 64 ? 32 : ctx->zchunk;
  // k_sgs_rows: each chunk evaluates the network at one extra level; 64-level
  // chunks halve that (measured 0.826 -> 0.810 ms per stage at 256^2 x 128)
  ctx->zchunk_sgs = nz >= 128 ? 64 : ctx->zchunk_mlp;
  if (hipSetDevice(p->device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
    g_create_error = "hipSetDevice/hipStreamCreate failed";
    delete ctx;
    return CF_E_HIP;
  }
  if (hipDeviceGetAttribute(&ctx->ncu, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess ||
      ctx->ncu <= 0)
    ctx->ncu = 256;
  ctx->eng = ctx->fp32 ? make_engine_t<float>(nx) : make_engine_t<double>(nx);
  int rc = scratch_setup(p->device, ctx->err);
  if (rc != CF_OK) {
    g_create_error = ctx->err;
    cf_destroy(ctx);
    return rc;
  }
  rc = ctx->eng->alloc(ctx);
  if (rc != CF_OK) {
    g_create_error = ctx->err;
    cf_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return CF_OK;
}

int cf_destroy(cf_ctx* ctx) {
  if (!ctx) return CF_OK;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
  for (void* p : ctx->allocs) (void)hipFree(p);
  if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
  if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
  if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx->eng;
  delete ctx;
  return CF_OK;
}

int cf_set_state(cf_ctx* ctx, const double* u, const double* v, const double* w,
                 const double* e) {
  if (!ctx || !u || !v || !w || !e) return CF_E_ARG;
  HIPCHK(hipSetDevice(ctx->p.device));
  HIPCHK(hipMemset(&ctx->ws->flags, 0, sizeof(int)));
  return ctx->eng->set_state(ctx, u, v, w, e);
}

static uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

int cf_seeded_init(cf_ctx* ctx, uint64_t seed, double amp) {
  if (!ctx) return CF_E_ARG;
  const int nx = ctx->nx, ny = ctx->ny, nz = ctx->nz;
  const size_t pl = (size_t)nx * ny;
  std::vector<double> f[4];
  for (auto& a : f) a.resize((size_t)nz * pl);
  const double zl = ctx->p.zl, dz = zl / nz;
  const double ut = std::sqrt(0.016 * zl / 2.0);
  const uint64_t key0 = seed * 0x100000001B3ull;
  for (int fld = 0; fld < 3; ++fld)
    for (int k = 1; k <= nz; ++k) {
      const double zm = (k - 0.5) * dz;
      const double zw = zm < zl - zm ? zm : zl - zm;
      const double prof = (fld == 0) ? ut / 0.4 * std::log(zw / ctx->p.z0) : 0.0;
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const uint64_t key = key0 ^ ((uint64_t)fld << 56) ^ ((uint64_t)k << 40) ^
                               ((uint64_t)j << 20) ^ (uint64_t)i;
          const double xi = (double)(splitmix64(key) >> 11) * 0x1.0p-53 * 2.0 - 1.0;
          double val = prof + amp * ut * xi;
          if (fld == 2 && k == nz) val = 0.0;
          f[fld][(size_t)(k - 1) * pl + (size_t)j * nx + i] = val;
        }
    }
  std::fill(f[3].begin(), f[3].end(), 0.01 * ut * ut);
  return cf_set_state(ctx, f[0].data(), f[1].data(), f[2].data(), f[3].data());
}

static int get_fields(cf_ctx* ctx, const int* which, double* const* outs, int n) {
  if (!ctx) return CF_E_ARG;
  HIPCHK(hipSetDevice(ctx->p.device));
  int rc = cf_sync(ctx);
  for (int i = 0; i < n; ++i) {
    if (!outs[i]) continue;
    int r = ctx->eng->get_field(ctx, which[i], outs[i]);
    if (r != CF_OK) return r;
  }
  return rc;
}

int cf_get_state(cf_ctx* ctx, double* u, double* v, double* w, double* e, double* p) {
  const int which[5] = {0, 1, 2, 3, 4};
  double* outs[5] = {u, v, w, e, p};
  return get_fields(ctx, which, outs, 5);
}

int cf_get_tendency(cf_ctx* ctx, double* ru, double* rv, double* rw, double* re) {
  const int which[4] = {5, 6, 7, 8};
  double* outs[4] = {ru, rv, rw, re};
  return get_fields(ctx, which, outs, 4);
}

int cf_get_comp1_tendency(cf_ctx* ctx, double* ru, double* rv, double* rw) {
  const int which[3] = {9, 10, 11};
  double* outs[3] = {ru, rv, rw};
  return get_fields(ctx, which, outs, 3);
}

int cf_stage(cf_ctx* ctx, int istage, int mnout) {
  if (!ctx || istage < 1 || istage > 3) return CF_E_ARG;
  return ctx->eng->stage(ctx, istage, mnout, nullptr);
}

int cf_step(cf_ctx* ctx, int nsteps) {
  if (!ctx || nsteps < 0) return CF_E_ARG;
  for (int it = 0; it < nsteps; ++it)
    for (int s = 1; s <= 3; ++s) {
      int rc = ctx->eng->stage(ctx, s, 0, nullptr);
      if (rc != CF_OK) return rc;
    }
  return CF_OK;
}

int cf_set_dt(cf_ctx* ctx, double dt) {
  if (!ctx || !(dt > 0)) return CF_E_ARG;
  ctx->dt = dt;
  return CF_OK;
}

int cf_sync(cf_ctx* ctx) {
  if (!ctx) return CF_E_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return check_flags(ctx);
}

int cf_get_divz(cf_ctx* ctx, double* divz) {
  if (!ctx || !divz) return CF_E_ARG;
  int rc = cf_sync(ctx);
  HIPCHK(hipMemcpy(divz, ctx->red + 1, ctx->nz * sizeof(double), hipMemcpyDeviceToHost));
  return rc;
}

int cf_get_scalars(cf_ctx* ctx, double* s4) {
  if (!ctx || !s4) return CF_E_ARG;
  int rc = cf_sync(ctx);
  cf::WallScalars ws{};
  HIPCHK(hipMemcpy(&ws, ctx->ws, sizeof(ws), hipMemcpyDeviceToHost));
  s4[0] = ws.utau;
  s4[1] = ws.utau_s;
  s4[2] = ws.dudt;
  s4[3] = ws.dvdt;
  return rc;
}

int cf_compmn(cf_ctx* ctx, double* meanbuf) {
  if (!ctx || !meanbuf) return CF_E_ARG;
  const int nz = ctx->nz;
  HIPCHK(hipSetDevice(ctx->p.device));
  if (!ctx->stats) {
    int rc = dalloc(ctx, &ctx->stats, (size_t)cf::NMEAN * nz);
    if (rc != CF_OK) return rc;
  }
  int rc = ctx->eng->means(ctx, ctx->stats);
  if (rc != CF_OK) return rc;
  std::vector<double> lv((size_t)cf::NMEAN * nz), red(7 * (size_t)ctx->slots);
  HIPCHK(hipMemcpyAsync(lv.data(), ctx->stats, lv.size() * sizeof(double), hipMemcpyDeviceToHost,
                        ctx->stream));
  HIPCHK(hipMemcpyAsync(red.data(), ctx->red, red.size() * sizeof(double), hipMemcpyDeviceToHost,
                        ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  rc = check_flags(ctx);
  if (rc != CF_OK) return rc;
  std::memset(meanbuf, 0, sizeof(double) * 19 * nz);
  for (int i = 0; i < nz; ++i) {                                     // compmn.F90:18-62
    double* mb = meanbuf + 19 * (size_t)i;
    const double* l = &lv[(size_t)cf::NMEAN * i];
    mb[3] = l[0];                                                    // uxym
    mb[4] = l[1];                                                    // vxym
    mb[5] = l[2];                                                    // wxym
    mb[6] = l[3];                                                    // uule
    mb[7] = l[4];                                                    // vvle
    mb[8] = l[5];                                                    // wwle
    mb[9] = l[6];                                                    // uwle
    mb[10] = l[7];                                                   // vwle
    mb[11] = l[8];                                                   // uvle
    mb[0] = 0.5 * (l[3] + l[4] + l[5]);                              // englez
    mb[1] = l[9];                                                    // engsbz
    mb[2] = red[0 * ctx->slots + i + 1];                             // divz
    mb[12] = 0.0;                                                    // pple is never computed
    // SGS means uusb vvsb uvsb wwsb uwsb vwsb (allocate.f90:106-111)
    for (int r = 0; r < 6; ++r) mb[13 + r] = ctx->have_sgs ? red[(size_t)(r + 1) * ctx->slots + i + 1] : 0.0;
  }
  if (ctx->have_sgs && ctx->have_sred) {
    std::vector<double> sr(6 * (size_t)ctx->slots);
    HIPCHK(hipMemcpy(sr.data(), ctx->sred, sr.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < nz; ++i)
      for (int r = 0; r < 6; ++r) meanbuf[19 * (size_t)i + 13 + r] += sr[(size_t)r * ctx->slots + i + 1];
  }
  return CF_OK;
}

int cf_calc_spectra(cf_ctx* ctx, int32_t nlev, const int32_t* zlevels, double* hats, double* var,
                    double* lscales) {
  if (!ctx || nlev <= 0 || nlev > ctx->nz || !zlevels || !hats || !var || !lscales) return CF_E_ARG;
  for (int i = 0; i < nlev; ++i)
    if (zlevels[i] < 1 || zlevels[i] > ctx->nz - 1) {
      ctx->err = "calc_spectra levels must lie in 1..nnz-1";
      return CF_E_ARG;
    }
  HIPCHK(hipSetDevice(ctx->p.device));
  const int nx = ctx->nx, ny = ctx->ny, nb = nx / 2 + 1;
  const size_t nout = (size_t)nlev * 3 * (nb + 3);
  int* dz = nullptr;
  double* dout = nullptr;
  HIPCHK(hipMallocAsync((void**)&dz, nlev * sizeof(int), ctx->stream));
  HIPCHK(hipMallocAsync((void**)&dout, nout * sizeof(double), ctx->stream));
  HIPCHK(hipMemcpyAsync(dz, zlevels, nlev * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  int rc = ctx->eng->spectra(ctx, dz, nlev, dout);
  std::vector<double> o(nout);
  if (rc == CF_OK)
    HIPCHK(hipMemcpyAsync(o.data(), dout, nout * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipFreeAsync(dz, ctx->stream));
  HIPCHK(hipFreeAsync(dout, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (rc != CF_OK) return rc;
  // ring counts of one (field, level): calc_spectra.f90:150-166 over the r2c
  // half plane kx = 0..nx/2 (weight 2 for 0 < kx < nx/2), every ky
  std::vector<double> cnt(nb, 0.0);
  for (int kx = 0; kx <= nx / 2; ++kx)
    for (int j = 0; j < ny; ++j) {
      const int ky = j <= nx / 2 ? j : j - nx;
      const int r = (int)std::lrint(std::sqrt((double)(kx * kx + ky * ky)));
      if (r <= nx / 2) cnt[r] += (kx >= 1 && kx <= nx / 2 - 1) ? 2.0 : 1.0;
    }
  for (int f = 0; f < 3; ++f)
    for (int l = 0; l < nlev; ++l) {
      const double* b = &o[((size_t)l * 3 + f) * (nb + 3)];
      double* h = hats + ((size_t)f * nlev + l) * nb;
      h[0] = b[0];                                                   // the r = 0 bin is not normalised
      for (int r = 1; r < nb; ++r)                                   // calc_spectra.f90:238-243
        h[r] = b[r] * PI * ((r + 0.5) * (r + 0.5) - (r - 0.5) * (r - 0.5)) / cnt[r];
      const double v = b[nb + 2];
      var[(size_t)f * nlev + l] = v;
      lscales[(size_t)(2 * f) * nlev + l] = b[nb] * PI / v;          // l_f1
      lscales[(size_t)(2 * f + 1) * nlev + l] = b[nb + 1] * PI / v;  // l_f2
    }
  return CF_OK;
}

int cf_calc_spectra_noalloc(cf_ctx* ctx, int32_t nlev, const int32_t* zlevels, double* hats, double* var,
                            double* lscales) {
  if (!ctx || nlev <= 0 || nlev > ctx->nz || !zlevels || !hats || !var || !lscales) return CF_E_ARG;
  if (nlev > kSpecMaxLevels) {
    ctx->err = "cf_calc_spectra_noalloc takes at most " + std::to_string(kSpecMaxLevels) + " levels";
    return CF_E_ARG;
  }
  for (int i = 0; i < nlev; ++i)
    if (zlevels[i] < 1 || zlevels[i] > ctx->nz - 1) {
      ctx->err = "calc_spectra levels must lie in 1..nnz-1";
      return CF_E_ARG;
    }
  HIPCHK(hipSetDevice(ctx->p.device));
  double* base = scratch_ptr(ctx->p.device);
  if (!base) {
    ctx->err = "no scratch address for this device";
    return CF_E_HIP;
  }
  const int nb = ctx->nx / 2 + 1;
  if (nb > kSpecMaxBins) {
    ctx->err = "cf_calc_spectra_noalloc's scratch holds nnx <= " + std::to_string(2 * (kSpecMaxBins - 1));
    return CF_E_ARG;
  }
  double* ring = base;
  double* outp = base + kSpecRing;
  int32_t* dz = (int32_t*)(base + kSpecRing + kSpecOut);
  HIPCHK(hipMemcpyAsync(dz, zlevels, nlev * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
  int rc = ctx->eng->spectra(ctx, dz, nlev, ring);            // writes every ring sum it reads back
  if (rc != CF_OK) return rc;
  hipLaunchKernelGGL(cf::k_spectra_finish, dim3(1), dim3(256), 0, ctx->stream, ring, nlev, ctx->nx, ctx->ny,
                     PI, outp);
  HIPCHK(hipGetLastError());
  const size_t nh = (size_t)3 * nlev * nb, nv = (size_t)3 * nlev, nl = (size_t)6 * nlev;
  HIPCHK(hipMemcpyAsync(hats, outp, nh * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(var, outp + nh, nv * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(lscales, outp + nh + nv, nl * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return CF_OK;
}

int cf_time_steps(cf_ctx* ctx, int nsteps, double* total_ms, double* kernel_ms) {
  if (!ctx || nsteps < 1 || !total_ms) return CF_E_ARG;
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipEventRecord(a, ctx->stream));
  int rc = CF_OK;
  for (int it = 0; it < nsteps && rc == CF_OK; ++it)
    for (int s = 1; s <= 3 && rc == CF_OK; ++s) rc = ctx->eng->stage(ctx, s, 0, nullptr);
  float ms = 0;
  if (rc == CF_OK) {
    HIPCHK(hipEventRecord(b, ctx->stream));
    HIPCHK(hipEventSynchronize(b));
    HIPCHK(hipEventElapsedTime(&ms, a, b));
  }
  *total_ms = ms;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc != CF_OK) return rc;
  if (kernel_ms) {
    constexpr int NE = CF_NKERN + 1;
    for (int k = 0; k < CF_NKERN; ++k) kernel_ms[k] = 0;
    std::vector<hipEvent_t> ev(NE, nullptr);
    struct Destroy {
      std::vector<hipEvent_t>& v;
      ~Destroy() {
        for (auto& x : v)
          if (x) (void)hipEventDestroy(x);
      }
    } destroy{ev};
    for (auto& x : ev) HIPCHK(hipEventCreate(&x));
    for (int it = 0; it < nsteps; ++it)
      for (int s = 1; s <= 3; ++s) {
        rc = ctx->eng->stage(ctx, s, 0, ev.data());
        if (rc != CF_OK) return rc;
        HIPCHK(hipEventSynchronize(ev[NE - 1]));
        for (int k = 0; k < NE - 1; ++k) {
          if (!ctx->slot_used[k]) continue;      // no kernel of this slot on this path
          float t = 0;
          HIPCHK(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
          kernel_ms[k] += t / (3.0 * nsteps);
        }
      }
  }
  return check_flags(ctx);
}

const char* cf_kernel_name(int k) {
  if (k < 0 || k >= CF_NKERN) return "";
  return kKernelNames[k];
}

int64_t cf_device_bytes(const cf_ctx* ctx) { return ctx ? ctx->bytes : 0; }

const char* cf_last_error(const cf_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Learned SGS closure: the network object (trainer and published tables),
// the closure switch of a context, and the training-data box filter.  No
// counterpart in the reference.
// ---------------------------------------------------------------------------
namespace mlpns = cf::mlp;

struct cf_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<void*> allocs;
  float in_mu[mlpns::NIN] = {}, in_sd[mlpns::NIN] = {}, out_mu[mlpns::NOUT] = {},
        out_sd[mlpns::NOUT] = {};
  mlpns::Net* net = nullptr;      // published forward tables + normalisation
  mlpns::bf16x8* bw = nullptr;    // published backward tables
  float* w = nullptr;             // FP32 master weights
  float* m = nullptr;             // Adam moments
  float* v = nullptr;
  float* grad = nullptr;
  float* partial = nullptr;       // [kMaxBlk][NPARAM+1]
  float* losses = nullptr;        // [kMaxSteps] per-step loss of one train call
  float* x = nullptr;             // samples [11][cap], normalised
  float* t = nullptr;             // targets [6][cap], normalised
  int64_t cap = 0, n = 0;
  int64_t step = 0;               // Adam step count
  float nrm_host[48] = {};        // staging of the Net normalisation block
};

namespace {

constexpr int kMaxBlk = 1024;
constexpr int kMaxSteps = 1 << 16;

#define MHIPCHK(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) {                                                            \
      m->err = std::string(#expr) + ": " + hipGetErrorString(e_);                      \
      return CF_E_HIP;                                                                 \
    }                                                                                  \
  } while (0)

template <typename X>
int malloc_mlp(cf_mlp* m, X** p, size_t count) {
  void* q = nullptr;
  MHIPCHK(hipMalloc(&q, count * sizeof(X)));
  MHIPCHK(hipMemset(q, 0, count * sizeof(X)));
  m->allocs.push_back(q);
  *p = (X*)q;
  return CF_OK;
}

// Upload the normalisation into the published Net and rebuild its tables.
int publish(cf_mlp* m) {
  // Net::in_mu[16], in_rs[16], out_sd[8], out_mu[8] are contiguous
  static_assert(offsetof(mlpns::Net, out_mu) - offsetof(mlpns::Net, in_mu) == 40 * sizeof(float),
                "normalisation block layout");
  float* nrm = m->nrm_host;
  std::memset(nrm, 0, sizeof(m->nrm_host));
  for (int k = 0; k < mlpns::NIN; ++k) {
    nrm[k] = m->in_mu[k];
    nrm[16 + k] = 1.0f / m->in_sd[k];
  }
  for (int k = 0; k < mlpns::NOUT; ++k) {
    nrm[32 + k] = m->out_sd[k];
    nrm[40 + k] = m->out_mu[k];
  }
  MHIPCHK(hipMemcpyAsync(m->net->in_mu, nrm, sizeof(m->nrm_host), hipMemcpyHostToDevice, m->stream));
  const int tot = mlpns::FTN * 512 + mlpns::BTN * 512 + mlpns::FBN;
  hipLaunchKernelGGL(mlpns::k_mlp_publish, dim3((tot + 255) / 256), dim3(256), 0, m->stream, m->w,
                     m->net, m->bw);
  MHIPCHK(hipGetLastError());
  return CF_OK;
}

int check_norm(cf_mlp* m) {
  for (int k = 0; k < mlpns::NIN; ++k)
    if (!(m->in_sd[k] > 0) || !std::isfinite(m->in_mu[k])) {
      m->err = "input normalisation: sd must be > 0 (set with cf_mlp_set_norm)";
      return CF_E_ARG;
    }
  for (int k = 0; k < mlpns::NOUT; ++k)
    if (!(m->out_sd[k] > 0) || !std::isfinite(m->out_mu[k])) {
      m->err = "output normalisation: sd must be > 0";
      return CF_E_ARG;
    }
  return CF_OK;
}

// Gradient over samples [off, off+cnt) in order (cnt > 0) or over batch
// random samples (cnt == 0); then Adam when update != 0.
int grad_step(cf_mlp* m, int64_t off, int64_t cnt, int64_t batch, uint64_t seed, float lr,
              int update, float* loss_dev) {
  const int64_t total = cnt > 0 ? cnt : batch;
  const int64_t nchunk = (total + 63) / 64;
  const int nblk = (int)std::min<int64_t>(kMaxBlk, (nchunk + mlpns::TW - 1) / mlpns::TW);
  mlpns::TrainArgs ta{};
  ta.net = m->net;
  ta.bw = m->bw;
  ta.x = m->x;
  ta.t = m->t;
  ta.cap = m->cap;
  ta.n = m->n;
  ta.off = off;
  ta.cnt = cnt;
  ta.batch = batch;
  ta.seed = seed;
  ta.gscale = (float)(2.0 / ((double)total * mlpns::NOUT));
  ta.partial = m->partial;
  hipLaunchKernelGGL(mlpns::k_mlp_grad, dim3(nblk), dim3(64 * mlpns::TW), 0, m->stream, ta);
  mlpns::AdamArgs aa{};
  aa.partial = m->partial;
  aa.nblk = nblk;
  aa.w = m->w;
  aa.m = m->m;
  aa.v = m->v;
  aa.grad = m->grad;
  aa.loss = loss_dev;
  aa.lr = lr;
  aa.b1 = 0.9f;
  aa.b2 = 0.999f;
  aa.eps = 1e-8f;
  if (update) {
    m->step += 1;
    aa.bc1 = (float)(1.0 - std::pow(0.9, (double)m->step));
    aa.bc2 = (float)(1.0 - std::pow(0.999, (double)m->step));
  } else {
    aa.bc1 = aa.bc2 = 1.f;
  }
  aa.inv_n6 = (float)(1.0 / ((double)total * mlpns::NOUT));
  aa.update = update;
  hipLaunchKernelGGL(mlpns::k_mlp_adam, dim3((mlpns::NPARAM + 1 + 255) / 256), dim3(256), 0,
                     m->stream, aa);
  MHIPCHK(hipGetLastError());
  if (update) return publish(m);
  return CF_OK;
}

}  // namespace

extern "C" {

int cf_mlp_create(int32_t device, cf_mlp** out) {
  if (!out) return CF_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    g_create_error = "no HIP device";
    return CF_E_NODEV;
  }
  cf_mlp* m = new cf_mlp();
  m->device = device;
  int rc = CF_OK;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&m->stream) != hipSuccess) {
    g_create_error = "hipSetDevice/hipStreamCreate failed";
    delete m;
    return CF_E_HIP;
  }
  if ((rc = malloc_mlp(m, &m->net, 1)) != CF_OK || (rc = malloc_mlp(m, &m->bw, mlpns::BTN * 64)) ||
      (rc = malloc_mlp(m, &m->w, mlpns::NPARAM)) || (rc = malloc_mlp(m, &m->m, mlpns::NPARAM)) ||
      (rc = malloc_mlp(m, &m->v, mlpns::NPARAM)) || (rc = malloc_mlp(m, &m->grad, mlpns::NPARAM)) ||
      (rc = malloc_mlp(m, &m->partial, (size_t)kMaxBlk * (mlpns::NPARAM + 1))) ||
      (rc = malloc_mlp(m, &m->losses, kMaxSteps))) {
    g_create_error = m->err;
    cf_mlp_destroy(m);
    return rc;
  }
  for (int k = 0; k < mlpns::NIN; ++k) m->in_sd[k] = 1.f;
  for (int k = 0; k < mlpns::NOUT; ++k) m->out_sd[k] = 1.f;
  if ((rc = publish(m)) != CF_OK || hipStreamSynchronize(m->stream) != hipSuccess) {
    g_create_error = m->err;
    cf_mlp_destroy(m);
    return rc != CF_OK ? rc : CF_E_HIP;
  }
  *out = m;
  return CF_OK;
}

int cf_mlp_destroy(cf_mlp* m) {
  if (!m) return CF_OK;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (void* p : m->allocs) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return CF_OK;
}

int cf_mlp_set_params(cf_mlp* m, const float* params) {
  if (!m || !params) return CF_E_ARG;
  for (int i = 0; i < mlpns::NPARAM; ++i)
    if (!std::isfinite(params[i])) {
      m->err = "non-finite parameter";
      return CF_E_ARG;
    }
  MHIPCHK(hipSetDevice(m->device));
  MHIPCHK(hipMemcpyAsync(m->w, params, mlpns::NPARAM * sizeof(float), hipMemcpyHostToDevice, m->stream));
  int rc = publish(m);
  if (rc != CF_OK) return rc;
  MHIPCHK(hipStreamSynchronize(m->stream));
  return CF_OK;
}

int cf_mlp_get_params(cf_mlp* m, float* params) {
  if (!m || !params) return CF_E_ARG;
  MHIPCHK(hipSetDevice(m->device));
  MHIPCHK(hipMemcpyAsync(params, m->w, mlpns::NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MHIPCHK(hipStreamSynchronize(m->stream));
  return CF_OK;
}

int cf_mlp_set_norm(cf_mlp* m, const float* in_mu, const float* in_sd, const float* out_mu,
                    const float* out_sd) {
  if (!m || !in_mu || !in_sd || !out_mu || !out_sd) return CF_E_ARG;
  std::memcpy(m->in_mu, in_mu, sizeof(m->in_mu));
  std::memcpy(m->in_sd, in_sd, sizeof(m->in_sd));
  std::memcpy(m->out_mu, out_mu, sizeof(m->out_mu));
  std::memcpy(m->out_sd, out_sd, sizeof(m->out_sd));
  int rc = check_norm(m);
  if (rc != CF_OK) return rc;
  MHIPCHK(hipSetDevice(m->device));
  rc = publish(m);
  if (rc != CF_OK) return rc;
  MHIPCHK(hipStreamSynchronize(m->stream));
  return CF_OK;
}

int cf_mlp_get_norm(cf_mlp* m, float* in_mu, float* in_sd, float* out_mu, float* out_sd) {
  if (!m) return CF_E_ARG;
  if (in_mu) std::memcpy(in_mu, m->in_mu, sizeof(m->in_mu));
  if (in_sd) std::memcpy(in_sd, m->in_sd, sizeof(m->in_sd));
  if (out_mu) std::memcpy(out_mu, m->out_mu, sizeof(m->out_mu));
  if (out_sd) std::memcpy(out_sd, m->out_sd, sizeof(m->out_sd));
  return CF_OK;
}

// File: "CFMLP001", int32 {11, 64, 6, 9478}, float in_mu[11], in_sd[11],
// out_mu[6], out_sd[6], params[9478], little endian.
int cf_mlp_save(cf_mlp* m, const char* path) {
  if (!m || !path) return CF_E_ARG;
  std::vector<float> p(mlpns::NPARAM);
  int rc = cf_mlp_get_params(m, p.data());
  if (rc != CF_OK) return rc;
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    m->err = std::string("cannot open ") + path;
    return CF_E_ARG;
  }
  const int32_t hdr[4] = {mlpns::NIN, mlpns::NH, mlpns::NOUT, mlpns::NPARAM};
  bool ok = std::fwrite("CFMLP001", 1, 8, f) == 8 && std::fwrite(hdr, 4, 4, f) == 4 &&
            std::fwrite(m->in_mu, 4, mlpns::NIN, f) == (size_t)mlpns::NIN &&
            std::fwrite(m->in_sd, 4, mlpns::NIN, f) == (size_t)mlpns::NIN &&
            std::fwrite(m->out_mu, 4, mlpns::NOUT, f) == (size_t)mlpns::NOUT &&
            std::fwrite(m->out_sd, 4, mlpns::NOUT, f) == (size_t)mlpns::NOUT &&
            std::fwrite(p.data(), 4, p.size(), f) == p.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) {
    m->err = std::string("write failed: ") + path;
    return CF_E_ARG;
  }
  return CF_OK;
}

int cf_mlp_load(cf_mlp* m, const char* path) {
  if (!m || !path) return CF_E_ARG;
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    m->err = std::string("cannot open ") + path;
    return CF_E_ARG;
  }
  char magic[8];
  int32_t hdr[4];
  float nrm[2 * mlpns::NIN + 2 * mlpns::NOUT];
  std::vector<float> p(mlpns::NPARAM);
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "CFMLP001", 8) == 0 &&
            std::fread(hdr, 4, 4, f) == 4 && hdr[0] == mlpns::NIN && hdr[1] == mlpns::NH &&
            hdr[2] == mlpns::NOUT && hdr[3] == mlpns::NPARAM &&
            std::fread(nrm, 4, 34, f) == 34 && std::fread(p.data(), 4, p.size(), f) == p.size();
  std::fclose(f);
  if (!ok) {
    m->err = std::string("not a CFMLP001 network file: ") + path;
    return CF_E_ARG;
  }
  int rc = cf_mlp_set_norm(m, nrm, nrm + mlpns::NIN, nrm + 2 * mlpns::NIN,
                           nrm + 2 * mlpns::NIN + mlpns::NOUT);
  if (rc != CF_OK) return rc;
  return cf_mlp_set_params(m, p.data());
}

int cf_mlp_forward(cf_mlp* m, int64_t n, const float* feats, float* stresses) {
  if (!m || n < 0 || (n > 0 && (!feats || !stresses))) return CF_E_ARG;
  if (n == 0) return CF_OK;
  MHIPCHK(hipSetDevice(m->device));
  float *df = nullptr, *dt = nullptr;
  MHIPCHK(hipMallocAsync((void**)&df, (size_t)n * mlpns::NIN * sizeof(float), m->stream));
  MHIPCHK(hipMallocAsync((void**)&dt, (size_t)n * mlpns::NOUT * sizeof(float), m->stream));
  MHIPCHK(hipMemcpyAsync(df, feats, (size_t)n * mlpns::NIN * sizeof(float), hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(mlpns::k_mlp_forward, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, m->stream,
                     m->net, df, dt, n);
  MHIPCHK(hipGetLastError());
  MHIPCHK(hipMemcpyAsync(stresses, dt, (size_t)n * mlpns::NOUT * sizeof(float), hipMemcpyDeviceToHost,
                         m->stream));
  MHIPCHK(hipFreeAsync(df, m->stream));
  MHIPCHK(hipFreeAsync(dt, m->stream));
  MHIPCHK(hipStreamSynchronize(m->stream));
  return CF_OK;
}

int cf_mlp_reserve(cf_mlp* m, int64_t capacity) {
  if (!m || capacity <= 0) return CF_E_ARG;
  MHIPCHK(hipSetDevice(m->device));
  MHIPCHK(hipStreamSynchronize(m->stream));
  for (float** p : {&m->x, &m->t}) {
    if (*p) {
      auto it = std::find(m->allocs.begin(), m->allocs.end(), (void*)*p);
      if (it != m->allocs.end()) m->allocs.erase(it);
      (void)hipFree(*p);
      *p = nullptr;
    }
  }
  m->cap = 0;
  m->n = 0;
  int rc;
  if ((rc = malloc_mlp(m, &m->x, (size_t)capacity * mlpns::NIN)) != CF_OK) return rc;
  if ((rc = malloc_mlp(m, &m->t, (size_t)capacity * mlpns::NOUT)) != CF_OK) return rc;
  m->cap = capacity;
  return CF_OK;
}

int cf_mlp_put_samples(cf_mlp* m, int64_t offset, int64_t n, const float* feats,
                       const float* targets) {
  if (!m || offset < 0 || n <= 0 || !feats || !targets || offset + n > m->cap) {
    if (m) m->err = "samples outside the reserved capacity";
    return CF_E_ARG;
  }
  int rc = check_norm(m);
  if (rc != CF_OK) return rc;
  MHIPCHK(hipSetDevice(m->device));
  std::vector<float> col((size_t)n);
  for (int k = 0; k < mlpns::NIN; ++k) {
    const float mu = m->in_mu[k], rs = 1.0f / m->in_sd[k];
    for (int64_t i = 0; i < n; ++i) col[(size_t)i] = (feats[i * mlpns::NIN + k] - mu) * rs;
    MHIPCHK(hipMemcpy(m->x + (size_t)k * m->cap + offset, col.data(), (size_t)n * sizeof(float),
                      hipMemcpyHostToDevice));
  }
  for (int k = 0; k < mlpns::NOUT; ++k) {
    const float mu = m->out_mu[k], rs = 1.0f / m->out_sd[k];
    for (int64_t i = 0; i < n; ++i) col[(size_t)i] = (targets[i * mlpns::NOUT + k] - mu) * rs;
    MHIPCHK(hipMemcpy(m->t + (size_t)k * m->cap + offset, col.data(), (size_t)n * sizeof(float),
                      hipMemcpyHostToDevice));
  }
  m->n = std::max(m->n, offset + n);
  return CF_OK;
}

int cf_mlp_loss_grad(cf_mlp* m, int64_t offset, int64_t n, float* loss, float* grad) {
  if (!m || offset < 0 || n <= 0 || offset + n > m->n) {
    if (m) m->err = "sample range outside the stored samples";
    return CF_E_ARG;
  }
  MHIPCHK(hipSetDevice(m->device));
  int rc = grad_step(m, offset, n, 0, 0, 0.f, 0, m->losses);
  if (rc != CF_OK) return rc;
  if (grad)
    MHIPCHK(hipMemcpyAsync(grad, m->grad, mlpns::NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  if (loss) MHIPCHK(hipMemcpyAsync(loss, m->losses, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MHIPCHK(hipStreamSynchronize(m->stream));
  return CF_OK;
}

int cf_mlp_train(cf_mlp* m, int32_t nsteps, int64_t batch, float lr, uint64_t seed, float* losses) {
  if (!m || nsteps <= 0 || nsteps > kMaxSteps || batch < 0 || !(lr > 0) || m->n <= 0) {
    if (m) m->err = "cf_mlp_train: bad arguments or no samples";
    return CF_E_ARG;
  }
  MHIPCHK(hipSetDevice(m->device));
  for (int s = 0; s < nsteps; ++s) {
    const uint64_t sd = seed * 0x9E3779B97F4A7C15ull + (uint64_t)(m->step + 1) * 0xD1B54A32D192ED03ull;
    int rc = batch > 0 ? grad_step(m, 0, 0, batch, sd, lr, 1, m->losses + s)
                       : grad_step(m, 0, m->n, 0, 0, lr, 1, m->losses + s);
    if (rc != CF_OK) return rc;
  }
  if (losses)
    MHIPCHK(hipMemcpyAsync(losses, m->losses, (size_t)nsteps * sizeof(float), hipMemcpyDeviceToHost,
                           m->stream));
  MHIPCHK(hipStreamSynchronize(m->stream));
  return CF_OK;
}

int cf_mlp_reset_optimizer(cf_mlp* m) {
  if (!m) return CF_E_ARG;
  MHIPCHK(hipSetDevice(m->device));
  MHIPCHK(hipMemsetAsync(m->m, 0, mlpns::NPARAM * sizeof(float), m->stream));
  MHIPCHK(hipMemsetAsync(m->v, 0, mlpns::NPARAM * sizeof(float), m->stream));
  MHIPCHK(hipStreamSynchronize(m->stream));
  m->step = 0;
  return CF_OK;
}

const char* cf_mlp_last_error(const cf_mlp* m) { return m ? m->err.c_str() : g_create_error.c_str(); }

int cf_set_closure(cf_ctx* ctx, const cf_mlp* m) {
  if (!ctx) return CF_E_ARG;
  HIPCHK(hipSetDevice(ctx->p.device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (!m) {
    ctx->closure_on = false;
    return CF_OK;
  }
  if ((ctx->nxb % 64) != 0) {
    ctx->err = "the learned closure needs nnx in {128, 256, 512} (3/2-grid rows of whole wavefronts)";
    return CF_E_ARG;
  }
  if (ctx->p.imodel != 3) {
    ctx->err = "the learned closure falls back to the TKE closure: imodel 3 required";
    return CF_E_ARG;
  }
  if (m->device != ctx->p.device) {
    ctx->err = "network and context are on different devices";
    return CF_E_ARG;
  }
  if (!ctx->closure_net) {
    int rc = dalloc(ctx, &ctx->closure_net, 1);
    if (rc != CF_OK) return rc;
    rc = dalloc(ctx, &ctx->closure_counts, 2);
    if (rc != CF_OK) return rc;
  }
  HIPCHK(hipStreamSynchronize(m->stream));
  HIPCHK(hipMemcpyAsync(ctx->closure_net, m->net, sizeof(mlpns::Net), hipMemcpyDeviceToDevice, ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->closure_counts, 0, 2 * sizeof(unsigned long long), ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->closure_on = true;
  return CF_OK;
}

int cf_set_closure_grid(cf_ctx* ctx, int32_t grid) {
  if (!ctx) return CF_E_ARG;
  if (grid != CF_CLOSURE_GRID_DEALIASED && grid != CF_CLOSURE_GRID_SMALL) {
    ctx->err = "closure grid must be CF_CLOSURE_GRID_DEALIASED or CF_CLOSURE_GRID_SMALL";
    return CF_E_ARG;
  }
  HIPCHK(hipSetDevice(ctx->p.device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (grid == CF_CLOSURE_GRID_SMALL && !ctx->stream2) {
    HIPCHK(hipStreamCreate(&ctx->stream2));                   // default priority
    HIPCHK(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
  }
  if (grid == CF_CLOSURE_GRID_SMALL && !ctx->sgsS) {
    const size_t elem = ctx->fp32 ? 2 * sizeof(float) : 2 * sizeof(double);
    const size_t band = (size_t)ctx->slots * ctx->ny * ctx->mx * elem;
    void* p = nullptr;
    int rc;
    if ((rc = dalloc(ctx, (char**)&p, 6 * band)) != CF_OK) return rc;
    ctx->sgsS = p;
    if ((rc = dalloc(ctx, (char**)&p, 6 * band)) != CF_OK) return rc;
    ctx->sgsO = p;
    if ((rc = dalloc(ctx, &ctx->sred, 6 * (size_t)ctx->slots)) != CF_OK) return rc;
  }
  ctx->closure_grid = grid;
  return CF_OK;
}

int cf_get_closure_counts(cf_ctx* ctx, int64_t* counts) {
  if (!ctx || !counts) return CF_E_ARG;
  counts[0] = counts[1] = 0;
  if (!ctx->closure_counts) return CF_OK;
  unsigned long long h[2];
  HIPCHK(hipMemcpyAsync(h, ctx->closure_counts, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->closure_counts, 0, sizeof(h), ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  counts[0] = (int64_t)h[0];
  counts[1] = (int64_t)h[1];
  return CF_OK;
}

int cf_filter_snapshot(cf_ctx* ctx, int32_t fx, int32_t fy, int32_t fz, double* out) {
  if (!ctx || !out) return CF_E_ARG;
  auto okf = [](int f) { return f == 1 || f == 2 || f == 4 || f == 8; };
  if (!okf(fx) || !okf(fy) || !(fz == 1 || fz == 2 || fz == 4) || ctx->nx % fx || ctx->ny % fy ||
      ctx->nz % fz) {
    ctx->err = "filter widths: fx, fy in {1,2,4,8}, fz in {1,2,4}, dividing the grid";
    return CF_E_ARG;
  }
  HIPCHK(hipSetDevice(ctx->p.device));
  const size_t n = (size_t)ctx->nz * ctx->ny * ctx->nx;
  const size_t nc = n / ((size_t)fx * fy * fz);
  constexpr int NF = mlpns::NFILT;
  if (!ctx->filt_in) {
    int rc = dalloc(ctx, &ctx->filt_in, 10 * n);
    if (rc != CF_OK) return rc;
  }
  if (ctx->filt_out_n < NF * nc) {
    int rc = dalloc(ctx, &ctx->filt_out, NF * nc);
    if (rc != CF_OK) return rc;
    ctx->filt_out_n = NF * nc;
  }
  // u, v, w, e, then ux, uy, vx, vy, wx, wy
  for (int f = 0; f < 10; ++f) {
    int rc = f < 4 ? ctx->eng->physical(ctx, f) : ctx->eng->physical_deriv(ctx, (f - 4) / 2, (f - 4) % 2);
    if (rc != CF_OK) return rc;
    HIPCHK(hipMemcpyAsync(ctx->filt_in + f * n, ctx->staging, n * sizeof(double),
                          hipMemcpyDeviceToDevice, ctx->stream));
  }
  const double dz = ctx->p.zl / ctx->nz;
  const double dsl = std::pow(ctx->p.xl / ctx->nx * ctx->p.yl / ctx->ny * dz, 1.0 / 3.0);
  mlpns::FilterArgs fa{};
  const double* fi = ctx->filt_in;
  fa.u = fi; fa.v = fi + n; fa.w = fi + 2 * n; fa.e = fi + 3 * n;
  fa.ux = fi + 4 * n; fa.uy = fi + 5 * n; fa.vx = fi + 6 * n; fa.vy = fi + 7 * n;
  fa.wx = fi + 8 * n; fa.wy = fi + 9 * n;
  fa.out = ctx->filt_out;
  fa.nx = ctx->nx; fa.ny = ctx->ny; fa.nz = ctx->nz; fa.fx = fx; fa.fy = fy; fa.fz = fz;
  fa.dz = dz;
  fa.alk = ctx->p.alk > 0 ? ctx->p.alk : dsl;
  fa.xkmax = dz * dz / ctx->dt / 5.0;                       // les.F90:216
  hipLaunchKernelGGL(mlpns::k_box_filter, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, ctx->stream, fa);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, ctx->filt_out, NF * nc * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return CF_OK;
}

}  // extern "C"
