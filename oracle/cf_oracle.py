"""FP64 CPU oracle for the RK3 timestep of the channel-flow LES (fxie2/paes-channelflow).

TEST INFRASTRUCTURE ONLY.  This module is the checker: it may be imported by
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` and by nothing else.  The product path (the HIP library behind
``include/channelflow.h``) never calls it and fails loudly when its own build
is missing.

It restates, in vectorised numpy (one array axis per grid direction, all
levels at once), the algorithm of the reference's Fortran stage:

    lower (lowerbc.f90:1-20) -> upper (upperbc.f90:1-30)
    -> comp1 (comp1.f90:1-188: derivsbot/derivs, compadv1, complen, compsgs1,
       sufsim/sufsim_upper, compsgs3, compress)
    -> comprhs (comprhs.f90:1-37) -> press (press.F90:1-115, tridv.f90)
    -> comp2 (comp2.f90:1-45)

with the spectral layer of filter_fftw.f90 (ffthrz, xderiv, yderiv, expand,
compress, deal) written on numpy's pocketfft.  The reference links FFTW 2.x
(version unpinned, configure.user:10); its arithmetic here is the DFT with the
same sign, normalisation, Nyquist handling and c2r conventions.

Array convention: numpy arrays are indexed [level, y, x], which is the memory
order of the Fortran arrays (x, y, z).  Every 3-D field is stored with level
L at index L + 1 (so level -1 is index 0); ``lev(a, L)`` reads it.

Pinning (what checks each function):
  * spectral layer: closed forms (derivatives and 3/2 interpolation of band-
    limited trigonometric polynomials), tests/test_oracle_spectral.py;
  * tridv / press: manufactured solutions and scipy.linalg.solve_banded,
    tests/test_oracle_press.py;
  * complen/compsgs1/sufsim/lower/upper/tridv: the reference's own Fortran
    sources compiled by oracle/build_ref.sh (oracle/_ref), when built,
    tests/test_oracle_ref.py;
  * comp1 composition and the full stage: invariants (discrete divergence
    after a stage, momentum budget, z-mirror equivariance, x/y translation
    equivariance, zero-mean of w) and closed-form single-mode tendencies,
    tests/test_oracle_stage.py.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# Constants of the reference (les.F90 and parameters.f90).
GAMA = (8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)      # les.F90:27
ZETAS = (0.0, 17.0 / 60.0, 5.0 / 12.0)          # les.F90:26
VK = 0.4                                        # parameters.f90:7 (vk)
BETAA = 0.35                                    # parameters.f90:8
EMIN = 1.0e-6                                   # les.F90:102
FORCING = 0.016                                 # comp1.f90:152 imposed dp/dx
OFF = 1                                         # storage offset: level L -> index L+1
OF = OFF


def lev(a, L):
    """Plane of level ``L`` of a level-offset field."""
    return a[L + OFF]


# ----------------------------------------------------------------------------
# Parameters and derived constants (les.F90:92-121, init.f90:4-44)
# ----------------------------------------------------------------------------
@dataclass
class Params:
    nnx: int
    nny: int
    nnz: int
    xl: float = 2.0 * math.pi
    yl: float = math.pi
    zl: float = 2.0
    z0: float = 7.519e-4
    dt: float = 0.01
    imodel: int = 3           # 3 = Moeng84 TKE length, 4 = Smagorinsky (complen.f90:19-24)
    isurface: int = 1         # wall model 1..6 (lowerbc.f90:126-161); 1 = Schumann
    ufree: float = 0.0        # the uninitialised ufree of lowerbc.f90:38,63 (see wall_stress)
    alk: float | None = None  # SGS length scale (allocate.f90:61, never set; None -> dsl)
    ugal: float = 0.0         # storage module globals, zero in the reference run
    vgal: float = 0.0
    closure: object = None    # learned-closure hook (sgs_mlp.LearnedClosure); None = TKE

    def __post_init__(self):
        nnx, nny, nnz = self.nnx, self.nny, self.nnz
        self.nnxb = 3 * nnx // 2                    # les.F90:99-100
        self.nnyb = 3 * nny // 2
        self.xymb = 1.0 / self.nnxb / self.nnyb     # les.F90:101
        self.xym = 1.0 / (nnx * nny)                # les.F90:106
        self.dx = self.xl / nnx
        self.dy = self.yl / nny
        self.dz = self.zl / nnz
        self.z1 = self.dz / 2.0
        self.zody = math.log(self.z1 / self.z0)     # les.F90:111
        self.dzm = 1.0 / self.dz
        self.dzm2 = self.dzm ** 2
        self.todzm = 0.5 / self.dz
        if nny == 1:
            self.dsl = math.sqrt(self.dx * self.dz)
        else:
            self.dsl = (self.dx * self.dy * self.dz) ** (1.0 / 3.0)   # les.F90:118
        self.ceps = 0.93 / self.dsl                 # les.F90:120
        if self.alk is None:
            self.alk = self.dsl
        self.xkmax = self.dz * self.dz / self.dt / 5.0   # les.F90:216
        # wavenumbers (init.f90:15-31)
        self.xk = _waves(nnx, self.xl, nnx // 2 + 1)
        self.yk = _waves(nny, self.yl, nny // 2 + 1)
        self.xkb = _waves(self.nnxb, self.xl, self.nnxb // 2 + 1)
        self.ykb = _waves(self.nnyb, self.yl, self.nnyb // 2 + 1)


def _waves(n, length, ncut):
    """xk(ix) = (ix-1) 2pi/L, negative above ncut (init.f90:15-31)."""
    k = np.arange(n, dtype=np.float64) * 2.0 * math.pi / length
    ix = np.arange(1, n + 1)
    k[ix > ncut] = -(n - ix[ix > ncut] + 1) * 2.0 * math.pi / length
    return k


# ----------------------------------------------------------------------------
# Spectral layer (filter_fftw.f90)
# ----------------------------------------------------------------------------
# The DFTs go through numpy's pocketfft.  FFT_WORKERS > 1 hands them to
# scipy.fft's multi-threaded pocketfft instead (the same DFT); only the CPU
# baseline timing sets it.
FFT_WORKERS = 1


class _Fft:
    def __getattr__(self, name):
        if FFT_WORKERS > 1:
            import functools
            import scipy.fft
            return functools.partial(getattr(scipy.fft, name), workers=FFT_WORKERS)
        return getattr(np.fft, name)


_fft = _Fft()


def ffthrz_fwd(a):
    """ffthrz(fa,n1,n2,-1) (filter_fftw.f90:107-130): 2-D r2c scaled by 1/(n1 n2).
    Returns the (n2, n1/2+1) complex half spectrum."""
    return _fft.rfft2(a, norm="forward")


def ffthrz_inv(c, n1):
    """ffthrz(fa,n1,n2,+1): unscaled 2-D c2r (y inverse first, then x c2r)."""
    return _fft.irfft2(c, s=(c.shape[-2], n1), norm="forward")


def xderiv(a, length):
    """xderiv (filter_fftw.f90:138-187): 1-D spectral d/dx along rows, k=0 and
    the Nyquist coefficient zeroed, /n after the unscaled inverse."""
    n = a.shape[-1]
    c = _fft.rfft(a, axis=-1)
    k = np.arange(n // 2 + 1) * 2.0 * math.pi / length
    c = c * (1j * k)
    c[..., 0] = 0.0
    c[..., n // 2] = 0.0
    return _fft.irfft(c, n=n, axis=-1)


def yderiv(a, length):
    """yderiv (filter_fftw.f90:194-248): same along columns (axis -2)."""
    n = a.shape[-2]
    if n == 1:
        return np.zeros_like(a)
    c = _fft.rfft(a, axis=-2)
    k = np.arange(n // 2 + 1) * 2.0 * math.pi / length
    c = c * (1j * k)[:, None]
    c[..., 0, :] = 0.0
    c[..., n // 2, :] = 0.0
    return _fft.irfft(c, n=n, axis=-2)


def expand(a, nxb, nyb):
    """expand (filter_fftw.f90:254-278): 3/2 zero-pad interpolation.  Keeps
    kx < nx/2 and ky in [0, ny/2) U (-ny/2, 0); Nyquist modes are dropped."""
    ny, nx = a.shape[-2], a.shape[-1]
    s = ffthrz_fwd(a)
    b = np.zeros(a.shape[:-2] + (nyb, nxb // 2 + 1), dtype=complex)
    h = ny // 2
    b[..., :h, : nx // 2] = s[..., :h, : nx // 2]
    if h > 1:
        b[..., nyb - (h - 1):, : nx // 2] = s[..., h + 1:, : nx // 2]
    return ffthrz_inv(b, nxb)


def compress(b, nx, ny):
    """compress (filter_fftw.f90:285-308): truncate a 3/2-grid field to the
    small band (kx < nx/2, |ky| < ny/2) and return it on the small grid."""
    nyb = b.shape[-2]
    sb = ffthrz_fwd(b)
    s = np.zeros(b.shape[:-2] + (ny, nx // 2 + 1), dtype=complex)
    h = ny // 2
    s[..., :h, : nx // 2] = sb[..., :h, : nx // 2]
    if h > 1:
        s[..., h + 1:, : nx // 2] = sb[..., nyb - (h - 1):, : nx // 2]
    return ffthrz_inv(s, nx)


def deal(a):
    """deal (filter_fftw.f90:316-329): remove the kx and ky Nyquist modes."""
    ny, nx = a.shape[-2], a.shape[-1]
    s = ffthrz_fwd(a)
    s[..., nx // 2] = 0.0
    if ny != 1:
        s[..., ny // 2, :] = 0.0
    return ffthrz_inv(s, nx)


# ----------------------------------------------------------------------------
# State
# ----------------------------------------------------------------------------
@dataclass
class State:
    """Prognostic fields, their RK registers and diagnostics (modules.f90:1-61).
    All 3-D arrays have shape (nnz+4, nny, nnx); level L at index L+1."""
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    e: np.ndarray
    ru: np.ndarray
    rv: np.ndarray
    rw: np.ndarray
    re: np.ndarray
    p: np.ndarray
    utau: float = 0.0
    utau_s: float = 0.0
    dudt: float = 0.0
    dvdt: float = 0.0
    divz: np.ndarray = None
    sgs: dict = field(default_factory=dict)   # uusb, vvsb, uvsb, wwsb, uwsb, vwsb

    def copy(self):
        return State(*(getattr(self, n).copy() for n in FIELDS), utau=self.utau,
                     utau_s=self.utau_s, dudt=self.dudt, dvdt=self.dvdt,
                     divz=None if self.divz is None else self.divz.copy(),
                     sgs={k: v.copy() for k, v in self.sgs.items()})


FIELDS = ("u", "v", "w", "e", "ru", "rv", "rw", "re", "p")


def zeros_state(pr: Params) -> State:
    shp = (pr.nnz + 4, pr.nny, pr.nnx)
    return State(*(np.zeros(shp) for _ in FIELDS))


# ----------------------------------------------------------------------------
# Seeded initial condition (counter-based, identical on CPU and GPU)
# ----------------------------------------------------------------------------
_M64 = (1 << 64) - 1


def splitmix64(x):
    """SplitMix64 finaliser on uint64 numpy arrays (Steele, Lea, Flood 2014)."""
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(_M64)
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(_M64)
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(_M64)
    return x ^ (x >> np.uint64(31))


def noise(seed, fieldno, nnx, nny, nnz):
    """xi ~ U[-1,1) per point, keyed by (seed, field, i, j, k) with k = 1..nnz.
    Returns shape (nnz, nny, nnx)."""
    k = np.arange(1, nnz + 1, dtype=np.uint64)[:, None, None]
    j = np.arange(nny, dtype=np.uint64)[None, :, None]
    i = np.arange(nnx, dtype=np.uint64)[None, None, :]
    key = (np.uint64(seed) * np.uint64(0x100000001B3)) ^ (np.uint64(fieldno) << np.uint64(56))
    key = key ^ (k << np.uint64(40)) ^ (j << np.uint64(20)) ^ i
    with np.errstate(over="ignore"):
        r = splitmix64(key.astype(np.uint64))
    return (r >> np.uint64(11)).astype(np.float64) * (2.0 ** -53) * 2.0 - 1.0


def utau_forcing(zl):
    """u_tau of the imposed forcing 0.016 over half height h = zl/2."""
    return math.sqrt(FORCING * zl / 2.0)


def initial_fields(pr: Params, seed: int = 10972, amp: float = 0.1):
    """Seeded log-law + noise initial condition on levels 1..nnz.

    u = u_tau/kappa ln(z_w/z0) + amp u_tau xi, v = amp u_tau xi,
    w = amp u_tau xi at faces 1..nnz-1 (w(nnz) = 0, the top wall),
    e = 0.01 u_tau^2, with z_w = min(zm, zl - zm), zm = (iz - 1/2) dz.
    The log profile follows init.f90:71-78.  Returns arrays (nnz, nny, nnx)
    *before* the dealiasing that set_state applies (les.F90:185-190)."""
    ut = utau_forcing(pr.zl)
    zm = (np.arange(1, pr.nnz + 1) - 0.5) * pr.dz
    zw = np.minimum(zm, pr.zl - zm)
    prof = ut / VK * np.log(zw / pr.z0)
    u = prof[:, None, None] + amp * ut * noise(seed, 0, pr.nnx, pr.nny, pr.nnz)
    v = amp * ut * noise(seed, 1, pr.nnx, pr.nny, pr.nnz)
    w = amp * ut * noise(seed, 2, pr.nnx, pr.nny, pr.nnz)
    w[-1] = 0.0
    e = np.full_like(u, 0.01 * ut * ut)
    return u, v, w, e


def set_state(pr: Params, u, v, w, e) -> State:
    """Load interior levels 1..nnz and apply deal to u, v, w, e (les.F90:185-190).
    RK registers and p start at zero (ru/rv/rw/re/p are never initialised by the
    reference before their first use; zero is the declared choice)."""
    st = zeros_state(pr)
    for name, a in (("u", u), ("v", v), ("w", w), ("e", e)):
        getattr(st, name)[1 + OFF: pr.nnz + 1 + OFF] = deal(np.asarray(a, dtype=np.float64))
    return st


# ----------------------------------------------------------------------------
# Boundary fills (lowerbc.f90:1-20, upperbc.f90:1-30)
# ----------------------------------------------------------------------------
def lower(st: State, pr: Params):
    """u,v(0)=0; w(-1:0)=0; e(0)=e(1)  (np=(2,2,2,1), nm=(1,1,2,1))."""
    lev(st.u, 0)[:] = 0.0
    lev(st.v, 0)[:] = 0.0
    lev(st.w, -1)[:] = 0.0
    lev(st.w, 0)[:] = 0.0
    lev(st.e, 0)[:] = lev(st.e, 1)


def upper(st: State, pr: Params):
    """u,v(nnz+1:nnz+2)=0; e(nnz:nnz+1)=e(nnz-1); w(nnz:nnz+2)=0; rw(nnz)=0."""
    n = pr.nnz
    for L in (n + 1, n + 2):
        lev(st.u, L)[:] = 0.0
        lev(st.v, L)[:] = 0.0
    for L in (n, n + 1):
        lev(st.e, L)[:] = lev(st.e, n - 1)
    for L in (n, n + 1, n + 2):
        lev(st.w, L)[:] = 0.0
    lev(st.rw, n)[:] = 0.0


# ----------------------------------------------------------------------------
# Wall models (lowerbc.f90:25-165 sufsim, upperbc.f90:33-158 sufsim_upper)
# ----------------------------------------------------------------------------
BETAA = 0.35                 # parameters.f90:8


def _fmax(a, b):
    """Fortran MAX(a, b) as the reference's compiler evaluates it, a > b ? a : b
    (so a NaN second argument wins)."""
    return a if a > b else b


def wall_stress(pr: Params, ub1, vb1, upper_wall: bool, st: State):
    """Wall stress on the 3/2 grid from the level-1 (or nnz) plane.

    Returns (tauuw, tauvw, utau, abort_flag).  Scalars: lowerbc.f90:56-91,
    upperbc.f90:63-88 (the upper wall adds ugal to the plane mean).  The
    fluctuation statistics lowerbc.f90:95-123 / upperbc.f90:93-121 and the six
    models lowerbc.f90:126-161 / upperbc.f90:124-156.  Decisions where the
    reference reads undefined values:
      - ufree (uninitialised local, lowerbc.f90:38,63) is Params.ufree, 0 by
        default, so windm = max(<wind>, 0.01).  A NaN ufree reproduces what
        the compiled reference does (windm = 0.01);
      - case 2 reads ufluc, vfluc, which the reference computes only for
        isurface > 2; they are computed as defined (lowerbc.f90:97-98) for
        every case;
      - moengv2m is computed from moengu2, as the reference does
        (lowerbc.f90:121, upperbc.f90:119).
    Case 6 uses st.utau_s, st.dudt, st.dvdt of the lower wall's call."""
    N = ub1.size
    u1xy = float(np.sum(ub1)) / N
    v1xy = float(np.sum(vb1)) / N
    if upper_wall:
        u1xy += pr.ugal
        v1xy += pr.vgal
    wind = np.sqrt(ub1 ** 2 + vb1 ** 2)
    windm = float(np.sum(wind)) / N
    windm = _fmax(windm, pr.ufree)
    windm = _fmax(windm, 0.01)
    vsfc = math.sqrt(u1xy * u1xy + v1xy * v1xy)
    vsfc = max(vsfc, 0.01)
    utau = vsfc * VK / pr.zody
    tep = min(max(u1xy / vsfc, -1.0), 1.0)
    thta = math.acos(tep)
    utau2 = utau * utau
    sgn_v = 1.0 if v1xy >= 0 else -1.0               # sign(1., v1xy)
    if upper_wall:
        auuwm = utau2 * math.cos(thta)
        auvwm = utau2 * math.sin(thta) * sgn_v
    else:
        auuwm = -utau2 * math.cos(thta)
        auvwm = -utau2 * math.sin(thta) * sgn_v
    sfluc = wind - windm
    ufluc = ub1 + pr.ugal - u1xy
    vfluc = vb1 + pr.vgal - v1xy
    ufluc2mroot = math.sqrt(float(np.sum(ufluc ** 2)) / N)
    vfluc2mroot = math.sqrt(float(np.sum(vfluc ** 2)) / N)
    suflucm = float(np.sum(ufluc * sfluc)) / N
    svflucm = float(np.sum(vfluc * sfluc)) / N
    moengu = sfluc * u1xy + windm * ufluc
    moengv = sfluc * v1xy + windm * vfluc
    moengu2mroot = math.sqrt(float(np.sum(moengu ** 2)) / N)
    moengv2mroot = moengu2mroot                       # moengv2m = <moengu2>
    sg = 1.0 if upper_wall else -1.0
    c = pr.isurface
    with np.errstate(divide="ignore", invalid="ignore"):
        if c == 1:                                    # Schumann
            tuw = sg * utau2 / vsfc * (ub1 + pr.ugal)
            tvw = sg * utau2 / vsfc * (vb1 + pr.vgal)
        elif c == 2:                                  # Moeng
            tuw = sg * utau2 / (windm * vsfc) * (windm * ufluc + wind * u1xy)
            tvw = sg * utau2 / (windm * vsfc) * (windm * vfluc + wind * v1xy)
        elif c == 3:                                  # Wei, Schumann fluctuation
            tuw = auuwm - sg * BETAA * ufluc * utau2 / ufluc2mroot
            tvw = auvwm - sg * BETAA * vfluc * utau2 / vfluc2mroot
        elif c == 4:                                  # Moeng without linearisation
            denom = math.sqrt((windm * u1xy + suflucm) ** 2 + (windm * v1xy + svflucm) ** 2)
            tuw = sg * utau2 * (wind * (ub1 + pr.ugal)) / denom
            tvw = sg * utau2 * (wind * (vb1 + pr.vgal)) / denom
        elif c == 5:                                  # Wei, Moeng fluctuation
            tuw = auuwm - sg * BETAA * utau2 * moengu / moengu2mroot
            tvw = auvwm - sg * BETAA * utau2 * moengv / moengv2mroot
        elif c == 6:                                  # mean-forcing balance + Schumann
            a6u = -pr.zl / 2 * (FORCING - st.dudt)
            a6v = -pr.zl / 2 * (-st.dvdt)
            tuw = -sg * a6u + sg * utau * st.utau_s * ufluc / vsfc
            tvw = -sg * a6v + sg * utau * st.utau_s * vfluc / vsfc
        else:
            raise NotImplementedError("isurface must be 1..6 (lowerbc.f90:126-161)")
    return tuw, tvw, utau, utau > 10.0


class UtauAbort(RuntimeError):
    """lesabort after the utau > 10 guard (lowerbc.f90:77-80)."""


# ----------------------------------------------------------------------------
# comp1: tendencies on the 3/2 grid, all levels at once (comp1.f90:1-188)
# ----------------------------------------------------------------------------
def sgs_fields(st: State, pr: Params):
    """The 3/2-grid intermediates of comp1 for every level at once: expanded
    buffers, gradients (derivsbot/derivs), eddy diffusivity (complen),
    stresses (compsgs1) and wall stresses (sufsim/sufsim_upper).  Arrays are
    level-offset like the state (index L+1).  Also sets st.utau/utau_s/dudt/dvdt."""
    n = pr.nnz
    nxb, nyb = pr.nnxb, pr.nnyb
    xl, yl = pr.xl, pr.yl

    # --- expanded buffers of the stage-start fields (comp1.f90:91-121).  Every
    # expand happens before the RK substep touches that level (comp1.f90:116
    # precedes :126), so all of them see the stage-start state.
    ub = np.zeros((n + 4, nyb, nxb))
    vb = np.zeros_like(ub)
    wb = np.zeros_like(ub)
    eb = np.zeros_like(ub)
    ub[0 + OF: n + 2 + OF] = expand(st.u[0 + OF: n + 2 + OF], nxb, nyb)      # levels 0..nnz+1
    vb[0 + OF: n + 2 + OF] = expand(st.v[0 + OF: n + 2 + OF], nxb, nyb)
    wb[0 + OF: n + 2 + OF] = expand(st.w[0 + OF: n + 2 + OF], nxb, nyb)      # faces 0..nnz+1
    eb[0 + OF: n + 2 + OF] = np.maximum(expand(st.e[0 + OF: n + 2 + OF], nxb, nyb), EMIN)

    # --- gradients (derivs1.f90:1-66, derivs2.f90:1-42)
    ux = np.zeros_like(ub); uy = np.zeros_like(ub); vx = np.zeros_like(ub); vy = np.zeros_like(ub)
    wz = np.zeros_like(ub); uz = np.zeros_like(ub); vz = np.zeros_like(ub)
    wx = np.zeros_like(ub); wy = np.zeros_like(ub); ex = np.zeros_like(ub); ey = np.zeros_like(ub)
    sl = slice(1 + OF, n + 1 + OF)                                  # levels 1..nnz
    ux[sl] = xderiv(ub[sl], xl); uy[sl] = yderiv(ub[sl], yl)
    vx[sl] = xderiv(vb[sl], xl); vy[sl] = yderiv(vb[sl], yl)
    wz[sl] = (wb[sl] - wb[0 + OF: n + OF]) * pr.dzm                 # wz(k) = (w(k)-w(k-1))/dz
    fs = slice(1 + OF, n + OF)                                      # faces 1..nnz-1
    uz[fs] = (ub[2 + OF: n + 1 + OF] - ub[fs]) * pr.dzm
    vz[fs] = (vb[2 + OF: n + 1 + OF] - vb[fs]) * pr.dzm
    uz[0 + OF] = uz[1 + OF]; vz[0 + OF] = vz[1 + OF]                # derivs1.f90:37-38
    uz[n + OF] = uz[n - 1 + OF]; vz[n + OF] = vz[n - 1 + OF]        # derivs2.f90:27-28
    wx[fs] = xderiv(wb[fs], xl); wy[fs] = yderiv(wb[fs], yl)        # faces 0, nnz stay 0
    es = slice(1 + OF, n - 1 + OF)                                  # faces 1..nnz-2
    ex[es] = xderiv(eb[es], xl); ey[es] = yderiv(eb[es], yl)
    ex[n - 1 + OF] = ex[n - 2 + OF]; ey[n - 1 + OF] = ey[n - 2 + OF]  # stale at nnz-1 (derivs2.f90:9,20-21)

    # --- eddy diffusivity at faces 0..nnz (complen.f90:1-28)
    xkk = np.zeros_like(ub)
    fa = slice(0 + OF, n + 1 + OF)
    if pr.imodel == 3:
        xkk[fa] = np.minimum(0.1 * pr.alk * np.sqrt(eb[fa]), pr.xkmax)
    elif pr.imodel == 4:
        lo = slice(0 + OF, n + 1 + OF); hi = slice(1 + OF, n + 2 + OF)
        shear = (ux[lo] ** 2 + ux[hi] ** 2 + vy[lo] ** 2 + vy[hi] ** 2 + wz[lo] ** 2 + wz[hi] ** 2
                 + 0.5 * (uy[lo] + vx[lo]) ** 2 + 0.5 * (uy[hi] + vx[hi]) ** 2
                 + (uz[fa] + wx[fa]) ** 2 + (vz[fa] + wy[fa]) ** 2)
        xkk[fa] = np.minimum((0.1 * pr.dsl) ** 2 * np.sqrt(shear), pr.xkmax)
    else:
        raise NotImplementedError("imodel must be 3 or 4 (complen.f90:19-24)")

    # --- stresses (compsgs1.f90:1-53)
    km = slice(0 + OF, n + OF)        # face k-1 for levels 1..nnz
    k0 = slice(1 + OF, n + 1 + OF)    # face k
    xs = xkk[km] + xkk[k0]
    tauuu = -xs * ux[sl]
    tauvv = -xs * vy[sl]
    tauuv = -xs * (uy[sl] + vx[sl]) * 0.5
    tauww = np.zeros_like(ub)
    tauww[sl] = -xs * wz[sl]                                        # level nnz+1 stays 0
    tauuw = np.zeros_like(ub); tauvw = np.zeros_like(ub)
    tauuw[fs] = -xkk[fs] * (uz[fs] + wx[fs])
    tauvw[fs] = -xkk[fs] * (vz[fs] + wy[fs])

    # wall stresses; sufsim reads the previous-stage R for its diagnostics
    st.dudt = float(np.sum(st.ru[1 + OF: n + 1 + OF])) / (pr.nnx * pr.nny * n)   # lowerbc.f90:70-72
    st.dvdt = float(np.sum(st.rv[1 + OF: n + 1 + OF])) / (pr.nnx * pr.nny * n)
    st.utau_s = math.sqrt(pr.zl / 2 * math.sqrt((-st.dvdt) ** 2 + (FORCING - st.dudt) ** 2))
    t0u, t0v, ut_lo, ab_lo = wall_stress(pr, ub[1 + OF], vb[1 + OF], False, st)
    if ab_lo:
        raise UtauAbort(f"utau={ut_lo}")
    tauuw[0 + OF] = t0u; tauvw[0 + OF] = t0v
    tnu, tnv, ut_hi, ab_hi = wall_stress(pr, ub[n + OF], vb[n + OF], True, st)
    if ab_hi:
        raise UtauAbort(f"utau={ut_hi}")
    tauuw[n + OF] = tnu; tauvw[n + OF] = tnv
    st.utau = ut_hi                                                 # sufsim_upper runs last

    return dict(ub=ub, vb=vb, wb=wb, eb=eb, ux=ux, uy=uy, vx=vx, vy=vy, wz=wz, uz=uz, vz=vz,
                wx=wx, wy=wy, ex=ex, ey=ey, xkk=xkk, tauuu=tauuu, tauvv=tauvv, tauuv=tauuv,
                tauww=tauww, tauuw=tauuw, tauvw=tauvw)


def comp1(st: State, pr: Params, istage: int, mnout: bool, dtab2: float):
    n = pr.nnz
    xl, yl = pr.xl, pr.yl
    F = sgs_fields(st, pr)
    if pr.closure is not None:
        pr.closure.apply(F, pr, st)          # learned stresses (no reference counterpart)
    ub, vb, wb, eb = F["ub"], F["vb"], F["wb"], F["eb"]
    ux, uy, vx, vy, wz = F["ux"], F["uy"], F["vx"], F["vy"], F["wz"]
    ex, ey, xkk = F["ex"], F["ey"], F["xkk"]
    tauuu, tauvv, tauuv = F["tauuu"], F["tauvv"], F["tauuv"]
    tauww, tauuw, tauvw = F["tauww"], F["tauuw"], F["tauvw"]
    sl = slice(1 + OF, n + 1 + OF)
    km = slice(0 + OF, n + OF)
    k0 = slice(1 + OF, n + 1 + OF)

    # --- advection (compadv1.f90:3-66); levels k=1..nnz
    up = ub[2 + OF: n + 2 + OF]; u0 = ub[sl]; um = ub[0 + OF: n + OF]
    vp = vb[2 + OF: n + 2 + OF]; v0 = vb[sl]; vm = vb[0 + OF: n + OF]
    wp = wb[2 + OF: n + 2 + OF]; w0 = wb[sl]; wm = wb[0 + OF: n + OF]
    ep = eb[2 + OF: n + 2 + OF]; e0 = eb[sl]; em = eb[0 + OF: n + OF]
    rub = -xderiv(u0 * u0, xl) - yderiv(u0 * v0, yl) - ((up + u0) * w0 - (um + u0) * wm) * pr.todzm
    rvb = -xderiv(u0 * v0, xl) - yderiv(v0 * v0, yl) - ((vp + v0) * w0 - (vm + v0) * wm) * pr.todzm
    rwb = (-0.5 * xderiv((u0 + up) * w0, xl) - 0.5 * yderiv((v0 + vp) * w0, yl)
           - ((0.5 * (wp + w0)) ** 2 - (0.5 * (w0 + wm)) ** 2) * pr.dzm)
    reb = (-0.5 * (xderiv(e0 * (u0 + up), xl) + yderiv(e0 * (v0 + vp), yl))
           - (wp * ep - wm * em) * pr.todzm)
    rub = rub + FORCING                                             # comp1.f90:152

    # --- SGS terms (compsgs3.f90:1-80)
    xk0 = xkk[k0]; xkp = xkk[2 + OF: n + 2 + OF]; xkm = xkk[km]
    temp = (xderiv(2.0 * xk0 * ex[sl], xl) + yderiv(2.0 * xk0 * ey[sl], yl)
            + ((xkp + xk0) * (ep - e0) - (xk0 + xkm) * (e0 - em)) * pr.dzm2)
    reb = reb + temp
    reb = reb + (-pr.ceps * e0 * np.sqrt(e0))
    rub = rub - xderiv(tauuu, xl) - yderiv(tauuv, yl) - (tauuw[k0] - tauuw[km]) * pr.dzm
    rvb = rvb - xderiv(tauuv, xl) - yderiv(tauvv, yl) - (tauvw[k0] - tauvw[km]) * pr.dzm
    rwb = rwb - xderiv(tauuw[k0], xl) - yderiv(tauvw[k0], yl) - (tauww[2 + OF: n + 2 + OF] - tauww[sl]) * pr.dzm
    rwb[n - 1] = 0.0                                                # level nnz: compsgs3.f90:60-64
    if mnout and istage == 1:
        mean = lambda a: np.sum(a, axis=(-2, -1)) * pr.xymb  # noqa: E731
        st.sgs = dict(uusb=mean(tauuu), vvsb=mean(tauvv), uvsb=mean(tauuv),
                      wwsb=mean(tauww[sl]), uwsb=mean(tauuw[k0]), vwsb=mean(tauvw[k0]))

    # --- divergence of the stage-start field on the 3/2 grid (comp1.f90:145)
    st.divz = np.sum((ux[sl] + vy[sl] + wz[sl]) ** 2, axis=(-2, -1)) * pr.xymb

    # --- RK substep phi -= dtab2 R_old (comp1.f90:126-133), then dealias
    if istage != 1:
        st.u[sl] -= dtab2 * st.ru[sl]
        st.v[sl] -= dtab2 * st.rv[sl]
        ns = slice(1 + OF, n + OF)
        st.w[ns] -= dtab2 * st.rw[ns]
        st.e[ns] -= dtab2 * st.re[ns]
    st.ru[sl] = compress(rub, pr.nnx, pr.nny)
    st.rv[sl] = compress(rvb, pr.nnx, pr.nny)
    ns = slice(1 + OF, n + OF)
    rw = compress(rwb[: n - 1], pr.nnx, pr.nny)
    rw -= (np.sum(rw, axis=(-2, -1)) * pr.xym)[:, None, None]      # comp1.f90:172
    st.rw[ns] = rw
    st.re[ns] = compress(reb[: n - 1], pr.nnx, pr.nny)
    st.rw[n + OF] = 0.0                                             # comp1.f90:174-177
    st.re[n + OF] = 0.0
    if pr.closure is not None and hasattr(pr.closure, "post"):
        pr.closure.post(st, pr, mnout and istage == 1)              # small-grid learned stresses



# ----------------------------------------------------------------------------
# comprhs, press, tridv, comp2
# ----------------------------------------------------------------------------
def comprhs(st: State, pr: Params, dtabm: float):
    """Poisson right-hand side per level (comprhs.f90:1-37)."""
    n = pr.nnz
    sl = slice(1 + OF, n + 1 + OF)
    rhs = xderiv(st.ru[sl] + st.u[sl] * dtabm, pr.xl)
    rhs = rhs + (yderiv(st.rv[sl] + st.v[sl] * dtabm, pr.yl) if pr.nny != 1 else 0.0)
    w = st.w; rw = st.rw
    zt = np.empty_like(rhs)
    zt[0] = pr.dzm * (lev(w, 1) * dtabm + lev(rw, 1))
    zt[n - 1] = -pr.dzm * (lev(w, n - 1) * dtabm + lev(rw, n - 1))
    if n > 2:
        zt[1: n - 1] = ((w[2 + OF: n + OF] - w[1 + OF: n - 1 + OF]) * dtabm
                        + rw[2 + OF: n + OF] - rw[1 + OF: n - 1 + OF]) * pr.dzm
    return rhs + zt


def tridv(sub, diag, sup, r):
    """Thomas algorithm with reciprocal pivots (tridv.f90:1-55).

    Arrays have the z index first: (n, ...).  b(1) and a(n) are not used."""
    d = diag.astype(np.float64).copy()
    r = r.copy()
    n = d.shape[0]
    if n <= 1:
        return r / d
    d[0] = 1.0 / d[0]
    for i in range(1, n):
        fac = sub[i] * d[i - 1]
        d[i] = 1.0 / (d[i] - fac * sup[i - 1])
        r[i] = r[i] - fac * r[i - 1]
    r[n - 1] = r[n - 1] * d[n - 1]
    for i in range(n - 2, -1, -1):
        r[i] = d[i] * (r[i] - sup[i] * r[i + 1])
    return r


def k2_table(pr: Params):
    """xk2 over the half spectrum (init.f90:33-42): xk(ix)^2 + yk(iy)^2."""
    ncx = pr.nnx // 2 + 1
    return pr.xk[:ncx][None, :] ** 2 + pr.yk[:, None] ** 2       # (nny, ncx)


def press(rhs, pr: Params):
    """press.F90:1-115 (serial): FFT, tridiagonal z-solve per (kx,ky) with
    Neumann rows at 1 and nnz, (0,0) mode pinned to 0, Nyquist rows zeroed."""
    n = pr.nnz
    P = ffthrz_fwd(rhs)                                # (n, ny, ncx)
    dzdz = pr.dz * pr.dz
    k2 = k2_table(pr)
    diag = np.broadcast_to(-dzdz * k2 - 2.0, P.shape).copy()
    diag[0] = -dzdz * k2 - 1.0
    diag[n - 1] = -dzdz * k2 - 1.0
    sub = np.ones(P.shape)
    sup = np.ones(P.shape)
    r = dzdz * P
    # kx=ky=0 mode: identity with zero right-hand side (press.F90:81-88)
    sub[:, 0, 0] = 0.0; diag[:, 0, 0] = 1.0; sup[:, 0, 0] = 0.0; r[:, 0, 0] = 0.0
    Ph = tridv(sub, diag, sup, r)
    Ph[:, 0, 0] = 0.0                                  # press.F90:107-108
    Ph[:, :, pr.nnx // 2] = 0.0                        # press.F90:109
    if pr.nny != 1:
        Ph[:, pr.nny // 2, :] = 0.0                    # press.F90:110
    return ffthrz_inv(Ph, pr.nnx), Ph


def comp2(st: State, pr: Params, dtab1: float):
    """Pressure gradient, RK update, clamp and mean removal (comp2.f90:1-45)."""
    n = pr.nnz
    sl = slice(1 + OF, n + 1 + OF)
    ns = slice(1 + OF, n + OF)
    p = st.p
    st.ru[sl] -= xderiv(p[sl], pr.xl)
    if pr.nny != 1:
        st.rv[sl] -= yderiv(p[sl], pr.yl)
    rw = st.rw[ns] - (p[2 + OF: n + 1 + OF] - p[ns]) * pr.dzm
    rw -= (np.sum(rw, axis=(-2, -1)) * pr.xym)[:, None, None]
    st.rw[ns] = rw
    st.u[sl] += dtab1 * st.ru[sl]
    if pr.nny != 1:
        st.v[sl] += dtab1 * st.rv[sl]
    w = st.w[ns] + dtab1 * st.rw[ns]
    st.e[ns] = np.maximum(st.e[ns] + dtab1 * st.re[ns], EMIN)
    w -= (np.sum(w, axis=(-2, -1)) * pr.xym)[:, None, None]
    st.w[ns] = w


# ----------------------------------------------------------------------------
# Stage / step driver (les.F90:254-285)
# ----------------------------------------------------------------------------
def stage(st: State, pr: Params, istage: int, mnout: bool = False):
    dtab1 = pr.dt * GAMA[istage - 1]
    dtab2 = pr.dt * ZETAS[istage - 1]
    dtabm = 1.0 / dtab1
    lower(st, pr)
    upper(st, pr)
    comp1(st, pr, istage, mnout, dtab2)
    rhs = comprhs(st, pr, dtabm)
    p, ph = press(rhs, pr)
    st.p[1 + OF: pr.nnz + 1 + OF] = p
    st.rhs = rhs
    st.phat = ph
    comp2(st, pr, dtab1)


def step(st: State, pr: Params, mnout: bool = False):
    for s in (1, 2, 3):
        stage(st, pr, s, mnout)


# ----------------------------------------------------------------------------
# Diagnostics: compmn (compmn.F90:1-76) and the x-y spectra (calc_spectra.f90)
# ----------------------------------------------------------------------------
MEAN_NAMES = ("englez", "engsbz", "divz", "uxym", "vxym", "wxym", "uule", "vvle", "wwle",
              "uwle", "vwle", "uvle", "pple", "uusb", "vvsb", "uvsb", "wwsb", "uwsb", "vwsb")
# meanbuf row order: allocate.f90:89-111


def compmn(st: State, pr: Params):
    """Horizontal-mean profiles and resolved moments (compmn.F90:1-64).
    Returns meanbuf as (19, nnz) in the allocate.f90:89-111 row order."""
    n = pr.nnz
    avg = lambda a: np.sum(a, axis=(-2, -1)) * pr.xym  # noqa: E731
    u = st.u; v = st.v; w = st.w
    sl = slice(1 + OF, n + 1 + OF)
    uxym = avg(u[sl]); vxym = avg(v[sl]); wxym = avg(w[sl])
    engsbz = avg(st.e[sl])
    up = u[2 + OF: n + 2 + OF]; vp = v[2 + OF: n + 2 + OF]
    uxp = np.append(uxym[1:], avg(lev(u, n + 1))); vxp = np.append(vxym[1:], avg(lev(v, n + 1)))
    ul = 0.5 * ((u[sl] - uxym[:, None, None]) + up - uxp[:, None, None])
    vl = 0.5 * ((v[sl] - vxym[:, None, None]) + vp - vxp[:, None, None])
    w0 = w[sl]
    uule = avg(ul * ul); vvle = avg(vl * vl); wwle = avg(w0 * w0)
    uwle = avg(ul * w0); vwle = avg(vl * w0); uvle = avg(ul * vl)
    englez = (uule + vvle + wwle) * 0.5
    mb = np.zeros((19, n))
    vals = dict(englez=englez, engsbz=engsbz, divz=st.divz if st.divz is not None else np.zeros(n),
                uxym=uxym, vxym=vxym, wxym=wxym, uule=uule, vvle=vvle, wwle=wwle, uwle=uwle,
                vwle=vwle, uvle=uvle, pple=np.zeros(n))
    for key in ("uusb", "vvsb", "uvsb", "wwsb", "uwsb", "vwsb"):
        vals[key] = st.sgs.get(key, np.zeros(n))
    for i, name in enumerate(MEAN_NAMES):
        mb[i] = vals[name]
    return mb


def calc_spectra(fields, zlevels):
    """x-y energy spectra at fixed levels (calc_spectra.f90:139-243).

    ``fields`` is a list of (u, v, w) arrays of shape (nnz, nny, nnx) holding
    levels 1..nnz (w at faces).  Returns (uhat, vhat, what) of shape
    (nnx/2+1, nlevels), plus (uvar, vvar, wvar) and the integral scales
    (l111, l112, l221, l222, l331, l332), with the reference's pooled
    nCount normalisation (Q21) and the unnormalised r=0 bin."""
    u0 = fields[0][0]
    nnz, nny, nnx = u0.shape
    nxy = nnx * nny
    nl = len(zlevels)
    waveN = np.array([i - 1 if i <= nnx // 2 + 1 else i - nnx - 1 for i in range(1, nnx + 1)], dtype=float)
    nCount = np.zeros(nnx // 2 + 1)
    hats = [np.zeros((nnx // 2 + 1, nl)) for _ in range(3)]
    var = [np.zeros(nl) for _ in range(3)]
    l1 = [np.zeros(nl) for _ in range(3)]
    l2 = [np.zeros(nl) for _ in range(3)]
    ii, jj = np.meshgrid(np.arange(nnx // 2 + 1), np.arange(nny), indexing="ij")
    rr = np.rint(np.sqrt(waveN[ii] ** 2 + waveN[jj] ** 2)).astype(int)   # never a half-integer
    wgt = np.where((ii >= 1) & (ii <= nnx // 2 - 1), 2.0, 1.0)
    keep = rr <= nnx // 2
    for (u, v, w) in fields:
        um = u.sum(axis=(1, 2)) / nxy
        vm = v.sum(axis=(1, 2)) / nxy
        for iz, zl in enumerate(zlevels):
            ins = (0.5 * ((u[zl - 1] - um[zl - 1]) + (u[zl] - um[zl])),
                   0.5 * ((v[zl - 1] - vm[zl - 1]) + (v[zl] - vm[zl])),
                   w[zl - 1])
            for f, a in enumerate(ins):
                F = np.fft.rfft2(a).T / nxy                    # (nnx/2+1, nny): fft_out(i, j)
                P = (F * np.conj(F)).real
                nCount += np.bincount(rr[keep], wgt[keep], minlength=nnx // 2 + 1)
                hats[f][:, iz] += np.bincount(rr[keep], (wgt * P)[keep], minlength=nnx // 2 + 1)
                l1[f][iz] += P[0, :].sum() + P[0, : nnx // 2].sum()
                l2[f][iz] += P[:, 0].sum()
                var[f][iz] += P.sum() + P[1: nnx // 2, :].sum()
    nf = len(fields)
    for i in range(1, nnx // 2 + 1):
        fac = math.pi * ((i + 0.5) ** 2 - (i - 0.5) ** 2) * 3 * nl * nf / nCount[i]
        for f in range(3):
            hats[f][i, :] *= fac
    var = [x / nf for x in var]
    ls = []
    for f in range(3):
        ls.append(l1[f] / nf * math.pi / var[f])
        ls.append(l2[f] / nf * math.pi / var[f])
    return hats, var, ls


# ----------------------------------------------------------------------------
# Helpers for tests
# ----------------------------------------------------------------------------
def interior(st: State, pr: Params, name: str):
    """Levels 1..nnz of a field, shape (nnz, nny, nnx)."""
    return getattr(st, name)[1 + OF: pr.nnz + 1 + OF].copy()


def divergence(st: State, pr: Params):
    """Small-grid discrete divergence dx u + dy v + (w(k)-w(k-1))/dz, k=1..nnz,
    the operator the pressure solve makes vanish (comprhs.f90:17-35)."""
    n = pr.nnz
    sl = slice(1 + OF, n + 1 + OF)
    return (xderiv(st.u[sl], pr.xl) + yderiv(st.v[sl], pr.yl)
            + (st.w[sl] - st.w[0 + OF: n + OF]) * pr.dzm)
