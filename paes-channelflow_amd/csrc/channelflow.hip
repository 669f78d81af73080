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
// (press.F90:153-154).  The SGS TKE e is clamped pointwise in physical space
// (comp2.f90:261) and so stays physical, with its x-spectral rows kept beside
// it for the next stage's synthesis.  Pressure p is a band spectrum in FP64.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/channelflow.h"
#include "cf_kernels.hpp"

using cf::cx;

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

const char* kKernelNames[CF_NKERN] = {"k_wall",  "k_col_synth", "k_rows",        "k_col_anal",
                                      "k_press", "k_erows",     "k_reduce_part", "(unused)"};

}  // namespace

struct Engine;

struct cf_ctx {
  cf_params p{};
  int nx = 0, ny = 0, nz = 0, mx = 0, nxb = 0, nyb = 0, slots = 0, zchunk = 16, ncu = 256;
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
  double* sd = nullptr;
  cx<double>* sr = nullptr;
  double* part = nullptr;
  double* red = nullptr;        // [7][slots]: divz and the 6 SGS means
  double* kxd = nullptr;
  double* kyd = nullptr;
  double* staging = nullptr;    // nz planes of nx*ny doubles
  cf::WallScalars* ws = nullptr;
  std::vector<double> sgs_host;  // 6*nz, valid after a mnout stage 1
  bool have_sgs = false;
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
    if ((rc = dalloc(ctx, &ctx->sd, S * NY * MX)) != CF_OK) return rc;
    if ((rc = dalloc(ctx, &ctx->sr, S * NY * MX)) != CF_OK) return rc;
    if ((rc = dalloc(ctx, &ctx->part, 7 * S * NYB)) != CF_OK) return rc;
    if ((rc = dalloc(ctx, &ctx->red, 7 * S)) != CF_OK) return rc;
    if ((rc = dalloc(ctx, &ctx->kxd, (size_t)MX)) != CF_OK) return rc;
    if ((rc = dalloc(ctx, &ctx->kyd, (size_t)NY)) != CF_OK) return rc;
    if ((rc = dalloc(ctx, &ctx->staging, (size_t)ctx->nz * NY * NX)) != CF_OK) return rc;
    if ((rc = dalloc(ctx, &ctx->ws, (size_t)1)) != CF_OK) return rc;
    // tables
    auto t1 = twiddles<T>(NX), t2 = twiddles<T>(NXB);
    auto t3 = twiddles<double>(NX);
    HIPCHK(hipMemcpy(ctx->twNX, t1.data(), NX * sizeof(C), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->twNXB, t2.data(), NXB * sizeof(C), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->tw64NX, t3.data(), NX * sizeof(cx<double>), hipMemcpyHostToDevice));
    std::vector<double> kx(MX), ky(NY);
    for (int k = 0; k < MX; ++k) kx[k] = (double)k * 2.0 * PI / ctx->p.xl;   // init.f90:437-440
    for (int j = 0; j < NY; ++j) {
      const int kk = (j < NY / 2) ? j : (j == NY / 2 ? 0 : j - NY);
      ky[j] = (double)kk * 2.0 * PI / ctx->p.yl;                            // init.f90:441-444
    }
    std::vector<T> kxt(kx.begin(), kx.end()), kyt(ky.begin(), ky.end());
    HIPCHK(hipMemcpy(ctx->kxp, kxt.data(), MX * sizeof(T), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->kyp, kyt.data(), NY * sizeof(T), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->kxd, kx.data(), MX * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->kyd, ky.data(), NY * sizeof(double), hipMemcpyHostToDevice));
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
    const size_t band_plane = (size_t)NY * MX;
    const size_t bstride = (size_t)ctx->slots * NYB * MX;
    int ek = 0;
    auto mark = [&]() {
      if (ev) (void)hipEventRecord(ev[ek++], s);
    };
    mark();
    hipLaunchKernelGGL(cf::k_wall<T>, dim3(1), dim3(64), 0, s, cptr(ctx->U), cptr(ctx->V),
                       cptr(ctx->Ruo), cptr(ctx->Rvo), band_plane, c, ctx->ws);
    mark();
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
      hipLaunchKernelGGL((cf::k_col_synth<T, NX, NY>), dim3(MX / CS::TK, nz, 4), dim3(CS::NT), 0,
                         s, a);
    }
    mark();
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
      a.c = c;
      const int nch = (nz + ctx->zchunk - 1) / ctx->zchunk;
      hipLaunchKernelGGL((cf::k_rows<T, NX>), dim3(NYB, nch), dim3(RS::NT), 0, s, a);
    }
    mark();
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
      const int ntiles = (MX / CS::TK) * nz * 4;
      const int grid = std::min(ntiles, 2 * ctx->ncu);
      hipLaunchKernelGGL((cf::k_col_anal<T, NX, NY>), dim3(grid), dim3(CS::NT), 0, s, a);
    }
    mark();
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
      a.sd = ctx->sd;
      a.sr = ctx->sr;
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
      hipLaunchKernelGGL(cf::k_press<T>, dim3((modes + 63) / 64), dim3(64), 0, s, a);
    }
    mark();
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
    mark();
    hipLaunchKernelGGL(cf::k_reduce_part, dim3(ctx->slots), dim3(256), 0, s, ctx->part,
                       (size_t)ctx->slots * NYB, NYB, sgs ? 7 : 1, c.xymb, ctx->red, ctx->slots);
    mark();
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      ctx->err = std::string("stage launch: ") + hipGetErrorString(e);
      return CF_E_HIP;
    }
    if (sgs) ctx->have_sgs = true;
    return CF_OK;
  }

  // physical (host, nz planes) -> band spectrum slots 1..nz of dst
  int to_band(cf_ctx* ctx, const double* host, C* dst) {
    const int nz = ctx->nz;
    HIPCHK(hipMemcpyAsync(ctx->staging, host, (size_t)nz * NY * NX * sizeof(double),
                          hipMemcpyHostToDevice, ctx->stream));
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
    // w(nnz) is the top wall (upperbc.f90:203)
    HIPCHK(hipMemsetAsync(cptr(ctx->W) + (size_t)nz * NY * MX, 0, NY * MX * sizeof(C), s));
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
    return CF_OK;
  }

  int get_field(cf_ctx* ctx, int which, double* out) override {
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
      rc = from_band<T>(ctx, cptr(bands[which]), cptr(ctx->xs_tmp), cptr(ctx->twNX));
    }
    if (rc != CF_OK) return rc;
    HIPCHK(hipMemcpyAsync(out, ctx->staging, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return CF_OK;
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
  if (flags & cf::FLAG_ECLAMP) {
    ctx->err = "3/2-grid SGS TKE below emin: the comp1 clamp (comp1.f90:109,121) is not "
               "supported on the device path";
    return CF_E_ECLAMP;
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
  if (p->imodel != 3 || p->isurface != 1) {
    g_create_error = "device path implements imodel=3 and isurface=1";
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
  if (hipSetDevice(p->device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
    g_create_error = "hipSetDevice/hipStreamCreate failed";
    delete ctx;
    return CF_E_HIP;
  }
  if (hipDeviceGetAttribute(&ctx->ncu, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess ||
      ctx->ncu <= 0)
    ctx->ncu = 256;
  ctx->eng = ctx->fp32 ? make_engine_t<float>(nx) : make_engine_t<double>(nx);
  int rc = ctx->eng->alloc(ctx);
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
  for (void* p : ctx->allocs) (void)hipFree(p);
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
  const int nx = ctx->nx, ny = ctx->ny, nz = ctx->nz;
  const size_t pl = (size_t)nx * ny;
  std::vector<double> u((size_t)nz * pl), v(u.size()), w(u.size()), e(u.size());
  int rc = cf_get_state(ctx, u.data(), v.data(), w.data(), e.data(), nullptr);
  if (rc != CF_OK) return rc;
  std::vector<double> red(7 * (size_t)ctx->slots);
  HIPCHK(hipMemcpy(red.data(), ctx->red, red.size() * sizeof(double), hipMemcpyDeviceToHost));
  const double xym = 1.0 / (double)pl;
  auto avg = [&](const double* a) {
    double s = 0;
    for (size_t i = 0; i < pl; ++i) s += a[i];
    return s * xym;
  };
  // compmn.F90:18-26: plane means
  std::vector<double> ux(nz + 1, 0.0), vx(nz + 1, 0.0);
  for (int i = 0; i < nz; ++i) {
    ux[i] = avg(&u[i * pl]);
    vx[i] = avg(&v[i * pl]);
  }
  std::memset(meanbuf, 0, sizeof(double) * 19 * nz);
  for (int i = 0; i < nz; ++i) {
    double* mb = meanbuf + 19 * (size_t)i;
    const double* ui = &u[i * pl];
    const double* vi = &v[i * pl];
    const double* wi = &w[i * pl];
    const double* up = (i + 1 < nz) ? &u[(i + 1) * pl] : nullptr;   // u(nnz+1) = 0 (upperbc.f90:189-191)
    const double* vp = (i + 1 < nz) ? &v[(i + 1) * pl] : nullptr;
    double uu = 0, vv = 0, ww = 0, uw = 0, vw = 0, uv = 0;
    for (size_t k = 0; k < pl; ++k) {                               // compmn.F90:29-57
      const double ul = 0.5 * ((ui[k] - ux[i]) + (up ? up[k] : 0.0) - ux[i + 1]);
      const double vl = 0.5 * ((vi[k] - vx[i]) + (vp ? vp[k] : 0.0) - vx[i + 1]);
      uu += ul * ul;
      vv += vl * vl;
      ww += wi[k] * wi[k];
      uw += ul * wi[k];
      vw += vl * wi[k];
      uv += ul * vl;
    }
    mb[6] = uu * xym;
    mb[7] = vv * xym;
    mb[8] = ww * xym;
    mb[9] = uw * xym;
    mb[10] = vw * xym;
    mb[11] = uv * xym;
    mb[0] = (mb[6] + mb[7] + mb[8]) * 0.5;                           // englez
    mb[1] = avg(&e[i * pl]);                                         // engsbz
    mb[2] = red[0 * ctx->slots + i + 1];                             // divz
    mb[3] = ux[i];
    mb[4] = vx[i];
    mb[5] = avg(wi);
    mb[12] = 0.0;                                                    // pple is never computed
    // SGS means uusb vvsb uvsb wwsb uwsb vwsb (allocate.f90:289-294)
    const int q_of_row[6] = {1, 2, 3, 4, 5, 6};  // red rows: uu vv uv ww uw vw
    for (int r = 0; r < 6; ++r)
      mb[13 + r] = ctx->have_sgs ? red[(size_t)q_of_row[r] * ctx->slots + i + 1] : 0.0;
  }
  return CF_OK;
}

int cf_time_steps(cf_ctx* ctx, int nsteps, double* total_ms, double* kernel_ms) {
  if (!ctx || nsteps < 1 || !total_ms) return CF_E_ARG;
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipEventRecord(a, ctx->stream));
  for (int it = 0; it < nsteps; ++it)
    for (int s = 1; s <= 3; ++s) {
      int rc = ctx->eng->stage(ctx, s, 0, nullptr);
      if (rc != CF_OK) return rc;
    }
  HIPCHK(hipEventRecord(b, ctx->stream));
  HIPCHK(hipEventSynchronize(b));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, a, b));
  *total_ms = ms;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (kernel_ms) {
    constexpr int NE = 8;
    for (int k = 0; k < CF_NKERN; ++k) kernel_ms[k] = 0;
    std::vector<hipEvent_t> ev(NE);
    for (auto& x : ev) HIPCHK(hipEventCreate(&x));
    for (int it = 0; it < nsteps; ++it)
      for (int s = 1; s <= 3; ++s) {
        int rc = ctx->eng->stage(ctx, s, 0, ev.data());
        if (rc != CF_OK) return rc;
        HIPCHK(hipEventSynchronize(ev[NE - 1]));
        for (int k = 0; k < NE - 1; ++k) {
          float t = 0;
          HIPCHK(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
          kernel_ms[k] += t / (3.0 * nsteps);
        }
      }
    for (auto& x : ev) (void)hipEventDestroy(x);
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
