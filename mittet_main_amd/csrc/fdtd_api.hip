// fdtd_api.hip -- host side of libmittet_fdtd.so: the C ABI declared in
// include/mittet_fdtd.h.  Owns the model (grid, eps', CPML profiles), shots
// (fields, psi shells, receivers, DFT accumulators) and batches in HBM, and
// drives the kernels of fdtd_kernels.hip on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include "../../include/mittet_fdtd.h"
#include "fdtd_dev.h"
#include "fdtd_transform.h"

using mt::ShotDev;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(FWI3D_EDEVICE, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_),  \
                        __FILE__, __LINE__);                                            \
    } while (0)

// The HIP runtime refuses an allocation or a synchronous copy while a stream
// captures a graph, in any thread.  Objects may be driven from several host
// threads (a gradient request beside the forward batch), so capture sections,
// allocations, frees and synchronous copies take this lock; kernel launches and
// graph launches do not.
std::recursive_mutex g_rt_mu;

template <typename T>
hipError_t dalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0)
        return hipSuccess;
    std::lock_guard<std::recursive_mutex> lk(g_rt_mu);
    return hipMalloc((void **)p, count * sizeof(T));
}

template <typename T>
void dfree(T *&p)
{
    if (p) {
        std::lock_guard<std::recursive_mutex> lk(g_rt_mu);
        (void)hipFree((void *)p);
    }
    p = nullptr;
}

hipError_t new_stream(hipStream_t *st)
{
    std::lock_guard<std::recursive_mutex> lk(g_rt_mu);
    return hipStreamCreateWithFlags(st, hipStreamNonBlocking);
}

void del_stream(hipStream_t st)
{
    std::lock_guard<std::recursive_mutex> lk(g_rt_mu);
    (void)hipStreamDestroy(st);
}

// synchronous host-to-device copy
hipError_t hcopy(void *dst, const void *src, size_t bytes)
{
    std::lock_guard<std::recursive_mutex> lk(g_rt_mu);
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
}

int axis_len(const fwi3d_grid_params &p, int a) { return a == 0 ? p.ix : (a == 1 ? p.iy : p.iz); }
double axis_step(const fwi3d_grid_params &p, int a) { return a == 0 ? p.dx : (a == 1 ? p.dy : p.dz); }

// CPML profiles of satan_i/fwi3d_cpu_cpmlabc_v1.4.c:7-175 for one axis:
// sigma graded with ((offset)/(nc-1))^(nn+order) from sigma_max
// = (nn+order+1)*cmax*ln(1/R)/(2*nc*h)*factor, kappa graded with exponent nn,
// alpha linear from 0 at the outer edge to alpha_max, and the recursive
// convolution constants b = exp(-(sigma/kappa+alpha)dt),
// c = sigma*(b-1)/(sigma+kappa*alpha)/kappa.  E quantities sit on integer
// offsets, H quantities half a cell inward; offsets are formed in single
// precision as the reference does, so the outermost cell exceeds sigma_max.
// A Fortran real power whose exponent is a whole-number constant: the
// compiler folds x**3.0d0 into the integer power, evaluated by repeated
// squaring (x**3 = x * (x*x)), which differs from pow(x, 3.0) in the last
// place for about half of the ramp values.  Other exponents: pow.
double fortran_pow(double x, double e)
{
    if (e != std::floor(e) || std::fabs(e) > 64.0)
        return std::pow(x, e);
    long long k = (long long)std::fabs(e);
    double r = 1.0, sq = x;
    bool any = false;
    while (k) {
        if (k & 1) {
            r = any ? r * sq : sq;
            any = true;
        }
        k >>= 1;
        if (k)
            sq = sq * sq;
    }
    return e < 0 ? 1.0 / r : r;
}

void cpml_axis(const fwi3d_grid_params &p, int a, std::vector<double> &out, int variant = 0)
{
    const int n = axis_len(p, a), nc = p.ncpml;
    const double h = axis_step(p, a);
    out.assign((size_t)FWI3D_NPROFILES * n, 0.0);
    auto at = [&](int row, int i) -> double & { return out[(size_t)row * n + i]; };
    const double sig_max = (p.nn + p.order + 1.0) * p.cmax * std::log(1.0 / p.rcoef) /
                           (2.0 * (nc * h)) * p.opt_to_max[a];
    // Fortran ver0040 (cpu_cpml4.f90:21-60): 1-based ramps in double, the high
    // side mirrors the low side, no float-typed offsets
    const double den = variant == 1 ? (double)nc - 1.0 : (double)((float)nc - 1.0f);
    for (int i = 0; i < n; ++i) {
        bool shell = true;
        double oe = 0, om = 0, ae = 0, am = 0;  // E/H offsets for sigma/kappa and alpha
        const int lo = i < nc ? i : n - 1 - i;  // variant 1: mirrored 0-based position
        if (variant == 1) {
            if (i < nc || i >= n - nc) {
                const double i1 = (double)(lo + 1);
                oe = (double)nc - i1;
                om = (double)nc - i1 - 0.5;
                ae = i1;
                am = i1 + 0.5;
            } else {
                shell = false;
            }
        } else if (i < nc) {
            oe = (double)(nc - i);
            om = (double)((float)(nc - i) - 0.5f);
            ae = (double)i;
            am = (double)((float)i + 0.5f);
        } else if (i >= n - nc) {
            oe = (double)((float)(i - n) + 1.0f + (float)nc);
            om = (double)((float)(i - n) + 0.5f + (float)nc);
            ae = (double)((float)(n - i) - 1.0f);
            am = (double)((float)(n - i) - 0.5f);
        } else {
            shell = false;
        }
        if (shell) {
            // C: pow at run time (cpmlabc.c:35-40); Fortran: constant exponents
            // (cpu_cpml4.f90:25-30, nn, order and ma are parameters)
            auto pw = [&](double x, double e) {
                return variant == 1 ? fortran_pow(x, e) : std::pow(x, e);
            };
            const double es = sig_max * pw(oe / den, p.nn + p.order);
            const double ms = sig_max * pw(om / den, p.nn + p.order);
            const double ek = 1.0 + (p.kappa_max - 1.0) * pw(oe / den, p.nn);
            const double mk = 1.0 + (p.kappa_max - 1.0) * pw(om / den, p.nn);
            const double ea = p.alpha_max * pw(ae / den, p.ma);
            const double ma = p.alpha_max * pw(am / den, p.ma);
            const double be = std::exp(-(es / ek + ea) * p.dt);
            const double bh = std::exp(-(ms / mk + ma) * p.dt);
            at(FWI3D_P_ESIG, i) = es;
            at(FWI3D_P_MSIG, i) = ms;
            at(FWI3D_P_EKAP, i) = ek;
            at(FWI3D_P_MKAP, i) = mk;
            at(FWI3D_P_AE, i) = ea;
            at(FWI3D_P_AM, i) = ma;
            at(FWI3D_P_BE, i) = be;
            at(FWI3D_P_BH, i) = bh;
            at(FWI3D_P_CE, i) = es * (be - 1.0) / (es + ek * ea) / ek;
            at(FWI3D_P_CH, i) = ms * (bh - 1.0) / (ms + mk * ma) / mk;
        } else {
            at(FWI3D_P_EKAP, i) = 1.0;
            at(FWI3D_P_MKAP, i) = 1.0;
        }
        at(FWI3D_P_KED, i) = at(FWI3D_P_EKAP, i) * h;
        at(FWI3D_P_KHD, i) = at(FWI3D_P_MKAP, i) * h;
    }
}

// Largest stable dt of the 4th-order leapfrog on the model: c_max =
// sqrt(2 omega0 / min(mu sigma)) over the cells, dt_lim = 1/(c_max sqrt(sum
// 1/dl^2)) times 6/7 (satan_i/fwi3d_cpu_lattice_v1.4.c:36-39, the C
// reference's courant*6/7 before its 0.999 safety factor) or 1/1.16667
// (ver0040/cpu_parameter_wafe.f90:118, Fortran, before 0.999).  A step above
// it grows without bound.  mu: per-cell permeability (null: mu0 everywhere).
double stable_dt(const fwi3d_grid_params &p, int variant, const double *sigma, const double *mu,
                 size_t n)
{
    double m = HUGE_VAL;
    for (size_t c = 0; c < n; ++c)
        m = std::min(m, (mu ? mu[c] : p.mu0) * sigma[c]);
    const double cmax = std::sqrt(2.0 * p.omega0 / m);
    const double inv = std::sqrt(1.0 / (p.dx * p.dx) + 1.0 / (p.dy * p.dy) + 1.0 / (p.dz * p.dz));
    return variant == 1 ? 1.0 / cmax / inv / 1.16667 : 1.0 / cmax / inv * 6.0 / 7.0;
}

int check_stable(const fwi3d_grid_params &p, int variant, const double *sigma, const double *mu,
                 size_t n)
{
    const double lim = stable_dt(p, variant, sigma, mu, n);
    if (p.dt > lim * (1.0 + 1e-12))
        return fail(FWI3D_EINVAL,
                    "dt %.9g exceeds the stability limit %.9g of this model (its lowest "
                    "conductivity-permeability product sets the fastest fictitious wave)",
                    p.dt, lim);
    return FWI3D_OK;
}

// host rows (FWI3D_P_*) -> device rows (mt::PR_*)
void pack_device_profile(const std::vector<double> &hp, int n, std::vector<double> &dp)
{
    static const int rows[mt::PR_N] = {FWI3D_P_ESIG, FWI3D_P_MSIG, FWI3D_P_BE, FWI3D_P_BH,
                                       FWI3D_P_CE,   FWI3D_P_CH,   FWI3D_P_KED, FWI3D_P_KHD};
    dp.assign((size_t)mt::PR_N * n, 0.0);
    for (int r = 0; r < mt::PR_N; ++r)
        std::memcpy(&dp[(size_t)r * n], &hp[(size_t)rows[r] * n], sizeof(double) * n);
}

int tiles(int n, int t) { return (n + t - 1) / t; }

// Tiles {x0, y0, k, yend} of the per-cell kernels around an LDS region [lo, hi):
// planes outside the region and rows of 32 x 8 tiles not inside it in y get
// every tile; rows inside it in y get 16 x 16 tiles over the two x-edge
// segments [0, lo_x) and [hi_x, ix) only.  Cells of the region that a tile
// still holds are skipped by the kernels.  Padded to a multiple of 8 entries
// with k = -1.
std::vector<int4> complement_tiles(int ix, int iy, int iz, const int lo[3], const int hi[3])
{
    std::vector<int4> shell;
    const int ftx = tiles(ix, mt::ZX), fty = tiles(iy, mt::ZY);
    const int ya = tiles(lo[1], mt::ZY), yb = hi[1] / mt::ZY;   // tile rows inside [lo_y, hi_y)
    for (int k = 0; k < iz; ++k) {
        const bool inz = k >= lo[2] && k < hi[2];
        for (int ty = 0; ty < fty; ++ty) {
            if (inz && ty >= ya && ty < yb)
                continue;
            for (int tx = 0; tx < ftx; ++tx)
                shell.push_back(int4{tx * mt::ZX, ty * mt::ZY, k, 0});
        }
        if (!inz || ya >= yb)
            continue;
        const int yend = yb * mt::ZY;
        for (int y0 = ya * mt::ZY; y0 < yend; y0 += 16) {
            for (int x0 = 0; x0 < lo[0]; x0 += 16)
                shell.push_back(int4{x0, y0, k, yend});
            for (int x0 = hi[0]; x0 < ix; x0 += 16)
                shell.push_back(int4{x0, y0, k, yend});
        }
    }
    shell.resize((size_t)tiles((int)shell.size(), 8) * 8, int4{0, 0, -1, 0});
    return shell;
}

constexpr int kDefaultChunk = 32;
constexpr int kDefaultZChunk = 32;

// Padding (doubles) between E ring slots and between DFT accumulator rows.
// Unpadded, the slot and row strides of power-of-two-rich grids are multiples
// of 256 KiB, so the many streams a DFT flush reads for one cell would share
// their low address bits.  MITTET_FDTD_SLOT_PAD / MITTET_FDTD_ACC_PAD override.
long long env_pad(const char *name, long long dflt)
{
    const char *v = std::getenv(name);
    return v ? std::atoll(v) : dflt;
}
// z-chunk length of the LDS kernels over a region of ltx x lty tiles and nz
// planes: 16 planes while that gives at least 2048 blocks (eight per CU),
// else 8 while that gives at least 512 (two per CU), else 4.  Measured on
// single shots with a 12-cell CPML (Gcell-updates/s for 16 / 8 / 4 planes):
// 120^3 13.6 / 14.5 / 15.0, 152^3 20.6 / 21.2 / 20.8, 216^3 19.9 / 20.3 /
// 20.2, 280^3 22.1 / 22.0 / 22.0.  MITTET_FDTD_LKB overrides.
int lds_chunk(int ltx, int lty, int nz)
{
    const long long e = env_pad("MITTET_FDTD_LKB", 0);
    if (e > 0)
        return (int)std::min<long long>(e, nz);
    const long long cols = (long long)ltx * lty;
    if (mt::LKB >= 16 && cols * tiles(nz, 16) >= 2048)
        return 16;
    if (cols * tiles(nz, 8) >= 512)
        return 8;
    return 4;
}

// Grids from this many cells up take the LDS family.  Graph-captured steps
// with a 4-frequency volume DFT, 16 receivers, Gcell-updates/s of the LDS /
// per-cell family: 84^3 9.8 / 12.7, 120^3 14.3 / 16.1, 152^3 16.9 / 15.7,
// 216^3 18.1 / 16.5.  MITTET_FDTD_LDS=1 takes it on any grid.
constexpr long long kLdsMinCells = 2500000;

// Launches whose steps touch more bytes than the 256 MB Infinity Cache load
// and store their once-touched field values non-temporally (stream_nt).
constexpr long long kNtMinBytes = 256LL << 20;

// Whether a C-variant shot takes the LDS family (family 3): size, no z-chunk
// or x-pair override, and a CPML-free interior to march.
bool takes_lds(const fwi3d_grid_params &p, int zchunk)
{
    const long long N = (long long)p.ix * p.iy * p.iz;
    const int xlo = (p.ncpml + 15) / 16 * 16;
    const bool pair = p.ix % 2 == 0 && env_pad("MITTET_FDTD_PAIR", 0);
    const bool lds_size = N >= kLdsMinCells || env_pad("MITTET_FDTD_LDS", 0);
    return zchunk == 0 && !pair && lds_size && p.ncpml >= 2 && xlo < p.ix - p.ncpml &&
           p.ncpml < p.iy - p.ncpml && p.ncpml < p.iz - p.ncpml && !env_pad("MITTET_FDTD_NOLDS", 0);
}

// With MITTET_FDTD_FUSE=1 at shot creation, family-3 shots with E in a ring
// of more than one slot run the fused E+H interior kernel (k_ehfield_l),
// which reads E and H of the chunk's planes k0 - 3 .. k1 + 2 (so ncpml >= 3)
// and indexes H planes with 32-bit offsets.  It is opt-in: at 288^3 it moves
// 104 B per interior cell instead of 152 but takes 580 us against 531 us
// for k_efield_l + k_hfield_l at 288^3 (register-bound at two blocks per CU,
// two barriers per plane and a 1.42x E halo recomputation).
bool takes_fused(const fwi3d_grid_params &p, int zchunk, int ring)
{
    const long long N = (long long)p.ix * p.iy * p.iz;
    return takes_lds(p, zchunk) && ring > 1 && p.ncpml >= 3 && 3 * N < (1LL << 31) &&
           env_pad("MITTET_FDTD_FUSE", 0);
}

constexpr long long kSlotPad = 0;
constexpr long long kAccPad = 0;

}  // namespace

struct fwi3d_model {
    fwi3d_grid_params p;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::vector<double> prof[3];
    double *d_prof[3] = {nullptr, nullptr, nullptr};
    double *d_sig = nullptr;
    double *d_eps = nullptr;
    double *d_cbi = nullptr;
    std::vector<double> sig;  // host copy (source scaling)
    int version = 0;          // bumps on every media_coeff_3d (and set_mu)
    int nshots = 0;
    int variant = 0;          // 0: C v1.4 semantics; 1: Fortran ver0040
    double *d_mu = nullptr;   // variant 1: per-cell permeability
    std::vector<double> mu;   // host copy (source scaling)
};

struct fwi3d_shot {
    fwi3d_model *m = nullptr;
    hipStream_t st = nullptr;
    int it = 0, nrec = 0, nf = 0, ring = 1;
    int src_comp = 0, src_cell = -1;
    std::vector<int> src_h;   // variant 1: Hz source cells
    std::vector<double> signal, omega;
    std::vector<int> rec;
    double *d_e = nullptr, *d_h = nullptr, *d_psi[3] = {nullptr, nullptr, nullptr};
    double *d_h2 = nullptr;   // second H buffer of the fused E+H step (null: not fused)
    bool fused = false;       // steps run the fused E+H interior kernel
    double *d_amp = nullptr, *d_traces = nullptr, *d_w = nullptr, *d_acc = nullptr;
    double *d_acc_h = nullptr;   // H volume spectra (fwi3d_enable_volume_dft_h)
    double *d_hring = nullptr;   // H after each step of a chunk, for the H sums
    int *d_rec = nullptr, *d_step = nullptr, *d_pre = nullptr, *d_flag = nullptr;
    ShotDev desc{};
    ShotDev *d_desc = nullptr;
    bool desc_dirty = true;
    int kind = 0;          // fused kernel family (see launch_eh)
    int nblocks = 0;       // fused blocks (padded to a multiple of 8)
    int nblocks_l = 0;     // LDS-kernel blocks of family 3 (padded to a multiple of 8)
    int4 *d_shell = nullptr;  // family 3: tiles of the per-cell kernels
    int nblocks_op = 0;    // per-cell blocks of the stand-alone operators
    int ncellblocks = 0;
    int step = 0;          // steps done (host mirror of *d_step)
    int amp_version = -1;
    hipGraphExec_t graph = nullptr;
    fwi3d_batch *batch = nullptr;
};

struct fwi3d_batch {
    int max_shots = 0, chunk = kDefaultChunk;
    hipStream_t st = nullptr;
    bool own_stream = false;
    std::vector<fwi3d_shot *> shots;
    ShotDev *d_desc = nullptr;
    int *d_pre = nullptr, *d_pre_l = nullptr, *d_pre_cells = nullptr;
    int nblocks = 0, nblocks_l = 0, ncellblocks = 0;
    int kind = 0;         // fused kernel family of the members (all equal)
    bool dirty = true;
    hipGraphExec_t graph = nullptr;
};

namespace {

int upload_desc(fwi3d_shot *s)
{
    HIP_TRY(hipMemcpyAsync(s->d_desc, &s->desc, sizeof(ShotDev), hipMemcpyHostToDevice, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    s->desc_dirty = false;
    return FWI3D_OK;
}

// source decrement per step, satan_i/fwi3d_cpu_function_v1.4.c:524-527
int ensure_amp(fwi3d_shot *s)
{
    fwi3d_model *m = s->m;
    if (m->version == 0)
        return fail(FWI3D_ESTATE, "fwi3d_media_coeff_3d has not been called on the model");
    if (s->amp_version == m->version || (s->src_cell < 0 && s->src_h.empty()))
        return FWI3D_OK;
    const fwi3d_grid_params &p = m->p;
    if (m->variant == 1) {
        // Hz -= signal(istep)*dt/myu/dx/dy/dz at every point (ver0040/cpu_waveform.f90:46-51);
        // step n takes s[n] (the Fortran istep = n + 1)
        const size_t L = (size_t)std::max(s->it - 1, 1);
        std::vector<double> amp(4 * L, 0.0);
        for (size_t q = 0; q < s->src_h.size(); ++q) {
            const double mu = m->mu[(size_t)s->src_h[q]];
            // each division of a complex (the Fortran signal) by a real taken
            // as the reference's amdflang build does: (x c) / (c c)
            auto cdivr = [](double x, double c) { return (x * c) / (c * c); };
            for (int n = 0; n + 1 < s->it; ++n)
                amp[q * L + n] = cdivr(cdivr(cdivr(cdivr(s->signal[n] * p.dt, mu), p.dx), p.dy), p.dz);
        }
        HIP_TRY(hipMemcpyAsync(s->d_amp, amp.data(), sizeof(double) * amp.size(),
                               hipMemcpyHostToDevice, s->st));
        HIP_TRY(hipStreamSynchronize(s->st));
        s->amp_version = m->version;
        return FWI3D_OK;
    }
    std::vector<double> amp((size_t)std::max(s->it - 1, 1), 0.0);
    const double sg = m->sig[(size_t)s->src_cell];
    for (int n = 0; n + 1 < s->it; ++n)
        amp[n] = (s->signal[n] + s->signal[n + 1]) / 2.0 * p.dt * 2.0 * p.omega0 / sg / p.dx /
                 p.dy / p.dz;
    HIP_TRY(hipMemcpyAsync(s->d_amp, amp.data(), sizeof(double) * amp.size(),
                           hipMemcpyHostToDevice, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    s->amp_version = m->version;
    return FWI3D_OK;
}

int check_shot_ready(fwi3d_shot *s)
{
    if (!s)
        return fail(FWI3D_EINVAL, "null shot");
    int rc = ensure_amp(s);
    if (rc)
        return rc;
    if (s->desc_dirty && (rc = upload_desc(s)))
        return rc;
    return FWI3D_OK;
}

// A fused kernel family and its blocks: 0 per-cell, 1 per-cell x-pairs,
// 2 z-marching, 3 LDS z-marching interior (pre_l / nblocks_l) plus per-cell
// kernels over the shell tiles (pre / nblocks).  `pre` are exclusive prefixes
// of per-shot block counts.
struct Fam {
    int kind;
    const int *pre;
    int nblocks;
    const int *pre_l;
    int nblocks_l;
    bool rings;   // every shot keeps E in a ring of more than one slot
    bool fused;   // family 3, every shot with a second H buffer: the fused E+H step
    bool nt;      // E/H launches with non-temporal once-touched values (stream_nt)
};

// Families whose E half-step has a per-cell launch that can carry the
// receiver gather of the previous step.  Only where every shot has a ring of
// more than one slot: step j reads slot (cur + j) % ring and writes the next
// one, so the gather of step j - 1 (slot (cur + j) % ring) is not written by
// that launch.  With one slot E is updated in place, and the leading gather
// blocks would race the tile blocks that overwrite it.
bool rec_folds(const Fam &f)
{
    return (f.kind == 0 || f.kind == 3) && f.nblocks > 0 && f.rings;
}

// rec_prev: the E launch of step j also gathers step j - 1's receivers
// (rec_folds(f) must hold).  Fused family 3: the E half-step is the shell E
// launch plus the fused E+H interior launch, the H half-step the shell H
// launch alone.
int launch_eh(const Fam &f, const ShotDev *d, int ns, int j, bool h, hipStream_t st,
              bool rec_prev = false)
{
    hipError_t e;
    if (f.kind == 4) {
        if (!h)
            HIP_TRY(mt::launch_source_hz(d, ns, j, st));
        e = h ? mt::launch_hfield_v(d, ns, f.pre, f.nblocks, j, st)
              : mt::launch_efield_v(d, ns, f.pre, f.nblocks, j, st);
        HIP_TRY(e);
        return FWI3D_OK;
    }
    if (f.kind == 2)
        e = h ? mt::launch_hfield_z(d, ns, f.pre, f.nblocks, j, 1, st)
              : mt::launch_efield_z(d, ns, f.pre, f.nblocks, j, 1, st);
    else if (f.kind == 1)
        e = h ? mt::launch_hfield_p(d, ns, f.pre, f.nblocks, j, 1, st)
              : mt::launch_efield_p(d, ns, f.pre, f.nblocks, j, 1, st);
    else
        e = h ? mt::launch_hfield_f(d, ns, f.pre, f.nblocks, j, 1, st, f.fused, f.nt)
              : mt::launch_efield_f(d, ns, f.pre, f.nblocks, j, 1, st, rec_prev, f.fused, f.nt);
    HIP_TRY(e);
    if (f.kind == 3 && f.fused) {
        if (!h)
            HIP_TRY(mt::launch_ehfield_l(d, ns, f.pre_l, f.nblocks_l, j, st));
    } else if (f.kind == 3) {
        HIP_TRY(h ? mt::launch_hfield_l(d, ns, f.pre_l, f.nblocks_l, j, 1, st, f.nt)
                  : mt::launch_efield_l(d, ns, f.pre_l, f.nblocks_l, j, 1, st, f.nt));
    }
    return FWI3D_OK;
}

// launches for local steps j in [0, L) followed by the DFT flush and step advance
int launch_chunk(const ShotDev *d, int ns, const Fam &f, const int *pre_cells, int ncellblocks,
                 int L, bool any_dft, bool any_rec, hipStream_t st, bool any_hdft = false)
{
    int rc;
    const bool fold = any_rec && rec_folds(f);   // step j's gather rides on step j + 1's E launch
    const int hf = f.fused ? 1 : 0;
    for (int j = 0; j < L; ++j) {
        if ((rc = launch_eh(f, d, ns, j, false, st, fold && j > 0)) ||
            (rc = launch_eh(f, d, ns, j, true, st)))
            return rc;
        if (any_rec && !(fold && j + 1 < L))
            HIP_TRY(mt::launch_receivers(d, ns, j, -1, 0, st, hf));
        if (any_hdft)
            HIP_TRY(mt::launch_hstore(d, ns, pre_cells, ncellblocks, j, st, hf));
    }
    if (any_dft)
        HIP_TRY(mt::launch_flush(d, ns, pre_cells, ncellblocks, L, st));
    if (any_hdft)
        HIP_TRY(mt::launch_flush(d, ns, pre_cells, ncellblocks, L, st, true));
    HIP_TRY(mt::launch_advance(d, ns, L, st, hf));
    return FWI3D_OK;
}

int capture_chunk(hipGraphExec_t *out, const ShotDev *d, int ns, const Fam &f,
                  const int *pre_cells, int ncellblocks, int L, bool any_dft, bool any_rec,
                  hipStream_t st, bool any_hdft = false)
{
    hipGraph_t g = nullptr;
    std::lock_guard<std::recursive_mutex> lk(g_rt_mu);
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = launch_chunk(d, ns, f, pre_cells, ncellblocks, L, any_dft, any_rec, st, any_hdft);
    hipError_t e = hipStreamEndCapture(st, &g);
    if (rc)
        return rc;
    if (e != hipSuccess)
        return fail(FWI3D_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess)
        return fail(FWI3D_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    return FWI3D_OK;
}

}  // namespace

extern "C" {

const char *fwi3d_last_error(void) { return g_err.c_str(); }
const char *fwi3d_version(void) { return "mittet_fdtd 0.1 gfx950"; }

double fwi3d_lattice_time_3d(double minval, double omega0, double mu0, double dx, double dy,
                             double dz, double factor, double *cmax_out)
{
    const double cmax = std::sqrt(2.0 * omega0 / mu0 / minval);
    const double courant =
        1.0 / cmax / std::sqrt(1.0 / std::pow(dx, 2) + 1.0 / std::pow(dy, 2) + 1.0 / std::pow(dz, 2));
    if (cmax_out)
        *cmax_out = cmax;
    return courant * 6.0 / 7.0 * factor;
}

int fwi3d_fmax_check(double maxval, double omega0, double mu0, double glim, double dx, double dy,
                     double dz, double fmax_w, double *fmax_out)
{
    if (!(maxval > 0) || !(omega0 > 0) || !(mu0 > 0) || !(glim > 0) || !(dx > 0) || !(dy > 0) ||
        !(dz > 0))
        return fail(FWI3D_EINVAL, "fmax_check: every argument must be > 0");
    const double cmin = std::sqrt(2.0 * omega0 / mu0 / maxval);
    const double fmax = cmin / glim / std::max(std::max(dx, dy), dz);
    if (fmax_out)
        *fmax_out = fmax;
    if (fmax < fmax_w)
        return fail(FWI3D_EINVAL, "fmax %.6g Hz of the grid is below the waveform's %.6g Hz", fmax,
                    fmax_w);
    return FWI3D_OK;
}

double fwi3d_stable_dt(const fwi3d_grid_params *p, int variant, const double *sigma,
                       const double *mu)
{
    if (!p || !sigma || (variant != 0 && variant != 1))
        return -1.0;
    return stable_dt(*p, variant, sigma, mu, (size_t)p->ix * p->iy * p->iz);
}

int fwi3d_model_create(const fwi3d_grid_params *p, void *stream, fwi3d_model **out)
{
    return fwi3d_model_create_variant(p, 0, stream, out);
}

int fwi3d_model_create_variant(const fwi3d_grid_params *p, int variant, void *stream,
                               fwi3d_model **out)
{
    if (!p || !out)
        return fail(FWI3D_EINVAL, "null argument");
    *out = nullptr;
    if (variant != 0 && variant != 1)
        return fail(FWI3D_EINVAL, "variant must be 0 (C v1.4) or 1 (Fortran ver0040)");
    if (p->ncpml < 2)
        return fail(FWI3D_EINVAL, "ncpml must be >= 2 (got %d)", p->ncpml);
    for (int a = 0; a < 3; ++a) {
        const int n = axis_len(*p, a);
        if (n < 2 * p->ncpml + 4)
            return fail(FWI3D_EINVAL, "axis %d: %d cells cannot hold two %d-cell CPML shells", a,
                        n, p->ncpml);
        if (!(axis_step(*p, a) > 0))
            return fail(FWI3D_EINVAL, "grid step of axis %d must be > 0", a);
    }
    if ((long long)p->ix * p->iy * p->iz >= (1LL << 31))
        return fail(FWI3D_EINVAL, "grid of %lld cells exceeds 2^31", (long long)p->ix * p->iy * p->iz);
    if (!(p->dt > 0) || !(p->omega0 > 0) || !(p->mu0 > 0) || !(p->cmax > 0))
        return fail(FWI3D_EINVAL, "dt, omega0, mu0 and cmax must be > 0");
    fwi3d_model *m = new fwi3d_model();
    m->p = *p;
    m->variant = variant;
    if (stream) {
        m->stream = (hipStream_t)stream;
    } else {
        hipError_t e = new_stream(&m->stream);
        if (e != hipSuccess) {
            delete m;
            return fail(FWI3D_EDEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
        }
        m->own_stream = true;
    }
    for (int a = 0; a < 3; ++a) {
        cpml_axis(*p, a, m->prof[a], variant);
        std::vector<double> dp;
        const int n = axis_len(*p, a);
        pack_device_profile(m->prof[a], n, dp);
        if (dalloc(&m->d_prof[a], dp.size()) != hipSuccess ||
            hcopy(m->d_prof[a], dp.data(), sizeof(double) * dp.size()) !=
                hipSuccess) {
            fwi3d_model_destroy(m);
            return fail(FWI3D_ENOMEM, "profile upload failed");
        }
    }
    const size_t N = (size_t)p->ix * p->iy * p->iz;
    if (dalloc(&m->d_sig, N) != hipSuccess || dalloc(&m->d_eps, N) != hipSuccess ||
        dalloc(&m->d_cbi, N) != hipSuccess) {
        fwi3d_model_destroy(m);
        return fail(FWI3D_ENOMEM, "cannot allocate %zu-cell model", N);
    }
    if (variant == 1) {
        // permeability mu0 everywhere until fwi3d_model_set_mu
        m->mu.assign(N, p->mu0);
        if (dalloc(&m->d_mu, N) != hipSuccess ||
            hcopy(m->d_mu, m->mu.data(), sizeof(double) * N) != hipSuccess) {
            fwi3d_model_destroy(m);
            return fail(FWI3D_ENOMEM, "cannot allocate the permeability volume");
        }
    }
    *out = m;
    return FWI3D_OK;
}

int fwi3d_model_set_mu(fwi3d_model *m, const double *mu)
{
    if (!m || !mu)
        return fail(FWI3D_EINVAL, "null argument");
    if (m->variant != 1)
        return fail(FWI3D_EINVAL, "per-cell permeability belongs to the Fortran variant (1)");
    const size_t N = (size_t)m->p.ix * m->p.iy * m->p.iz;
    for (size_t c = 0; c < N; ++c)
        if (!(mu[c] > 0) || !std::isfinite(mu[c]))
            return fail(FWI3D_EINVAL, "mu[%zu] must be finite and > 0", c);
    int rc;
    if (m->sig.size() == N && (rc = check_stable(m->p, m->variant, m->sig.data(), mu, N)))
        return rc;
    m->mu.assign(mu, mu + N);
    HIP_TRY(hipStreamSynchronize(m->stream));
    if (hcopy(m->d_mu, mu, sizeof(double) * N) != hipSuccess)
        return fail(FWI3D_EDEVICE, "permeability upload failed");
    m->version++;   // source amplitudes depend on mu
    return FWI3D_OK;
}

int fwi3d_model_destroy(fwi3d_model *m)
{
    if (!m)
        return FWI3D_OK;
    if (m->nshots > 0)
        return fail(FWI3D_ESTATE, "model still has %d shots", m->nshots);
    for (auto &q : m->d_prof)
        dfree(q);
    dfree(m->d_sig);
    dfree(m->d_eps);
    dfree(m->d_cbi);
    dfree(m->d_mu);
    if (m->own_stream && m->stream)
        del_stream(m->stream);
    delete m;
    return FWI3D_OK;
}

int fwi3d_model_get_profile(const fwi3d_model *m, int axis, int which, double *out)
{
    if (!m || !out || axis < 0 || axis > 2 || which < 0 || which >= FWI3D_NPROFILES)
        return fail(FWI3D_EINVAL, "bad profile request");
    const int n = axis_len(m->p, axis);
    std::memcpy(out, &m->prof[axis][(size_t)which * n], sizeof(double) * n);
    return FWI3D_OK;
}

int fwi3d_media_coeff_3d(fwi3d_model *m, const double *sigma)
{
    if (!m || !sigma)
        return fail(FWI3D_EINVAL, "null argument");
    const size_t N = (size_t)m->p.ix * m->p.iy * m->p.iz;
    for (size_t c = 0; c < N; ++c)
        if (!(sigma[c] > 0) || !std::isfinite(sigma[c]))
            return fail(FWI3D_EINVAL, "sigma[%zu] = %g: conductivity must be finite and > 0", c,
                        sigma[c]);
    int rc = check_stable(m->p, m->variant, sigma, m->variant == 1 ? m->mu.data() : nullptr, N);
    if (rc)
        return rc;
    m->sig.assign(sigma, sigma + N);
    HIP_TRY(hipMemcpyAsync(m->d_sig, sigma, sizeof(double) * N, hipMemcpyHostToDevice, m->stream));
    HIP_TRY(mt::launch_eps(m->d_sig, m->d_eps, m->d_cbi, (int)N, m->p.omega0, m->p.dt, m->stream));
    HIP_TRY(hipStreamSynchronize(m->stream));
    m->version++;
    return FWI3D_OK;
}

long long fwi3d_shot_bytes(const fwi3d_grid_params *p, const fwi3d_shot_params *sp)
{
    if (!p || !sp)
        return -1;
    const long long N = (long long)p->ix * p->iy * p->iz;
    const long long nc2 = 2LL * p->ncpml;
    const int ring = sp->nf_vol > 0 ? (sp->ring > 0 ? sp->ring : kDefaultChunk) : 1;
    long long b = 8LL * (3LL * N * ring + 3LL * N);                          // E ring, H
    if (takes_fused(*p, sp->zchunk, ring))
        b += 8LL * 3 * N;                                                     // second H
    b += 8LL * 4 * ((long long)p->iz * p->iy * nc2 + (long long)p->iz * nc2 * p->ix +
                    nc2 * p->iy * p->ix);                                     // psi slabs
    b += 8LL * sp->it + 8LL * 6 * sp->nrec * (long long)sp->it + 4LL * sp->nrec;
    b += 16LL * sp->nf_vol * (long long)sp->it + 16LL * 3 * sp->nf_vol * N;  // weights, acc
    return b;
}

long long fwi3d_volume_dft_h_bytes(const fwi3d_grid_params *p, const fwi3d_shot_params *sp)
{
    if (!p || !sp)
        return -1;
    if (sp->nf_vol <= 0)
        return 0;
    const long long N = (long long)p->ix * p->iy * p->iz;
    const long long ring = sp->ring > 0 ? sp->ring : kDefaultChunk;
    return 8LL * 3 * N * ring + 16LL * 3 * sp->nf_vol * N;
}

int fwi3d_shot_create(fwi3d_model *m, const fwi3d_shot_params *sp, fwi3d_shot **out)
{
    if (!m || !sp || !out)
        return fail(FWI3D_EINVAL, "null argument");
    *out = nullptr;
    const fwi3d_grid_params &p = m->p;
    if (sp->it < 2)
        return fail(FWI3D_EINVAL, "it must be >= 2");
    const bool magnetic = m->variant == 1;
    if (!magnetic && (sp->src_comp < 0 || sp->src_comp > 2))
        return fail(FWI3D_EINVAL, "src_comp must be 0, 1 or 2");
    if (magnetic && sp->src_comp != FWI3D_SRC_HZ)
        return fail(FWI3D_EINVAL, "Fortran-variant shots take the magnetic Hz source (src_comp %d)",
                    FWI3D_SRC_HZ);
    const bool has_src = sp->signal != nullptr;
    if (has_src && (sp->src_i < 0 || sp->src_i >= p.ix || sp->src_j < 0 || sp->src_j >= p.iy ||
                    sp->src_k < 0 || sp->src_k >= p.iz))
        return fail(FWI3D_EINVAL, "source cell (%d,%d,%d) outside the %dx%dx%d grid", sp->src_i,
                    sp->src_j, sp->src_k, p.ix, p.iy, p.iz);
    if (sp->nrec < 0 || (sp->nrec > 0 && !sp->rec_ijk))
        return fail(FWI3D_EINVAL, "bad receiver list");
    for (int r = 0; r < sp->nrec; ++r) {
        const int *q = sp->rec_ijk + 3 * r;
        if (q[0] < 0 || q[0] >= p.ix || q[1] < 0 || q[1] >= p.iy || q[2] < 0 || q[2] >= p.iz)
            return fail(FWI3D_EINVAL, "receiver %d at (%d,%d,%d) outside the grid", r, q[0], q[1],
                        q[2]);
    }
    if (sp->zchunk < 0)
        return fail(FWI3D_EINVAL, "zchunk must be >= 0");
    if (sp->nf_vol < 0 || (sp->nf_vol > 0 && !sp->omega_vol))
        return fail(FWI3D_EINVAL, "bad volume DFT frequency list");
    for (int f = 0; f < sp->nf_vol; ++f)
        if (!(sp->omega_vol[f] >= 0) || !std::isfinite(sp->omega_vol[f]))
            return fail(FWI3D_EINVAL, "omega_vol[%d] must be finite and >= 0", f);
    int ring = 1;
    if (sp->nf_vol > 0) {
        ring = sp->ring > 0 ? sp->ring : kDefaultChunk;
        if (ring > mt::RING_MAX)
            return fail(FWI3D_EINVAL, "ring %d exceeds %d", ring, mt::RING_MAX);
    }

    fwi3d_shot *s = new fwi3d_shot();
    s->m = m;
    s->st = m->stream;
    s->it = sp->it;
    s->nrec = sp->nrec;
    s->nf = sp->nf_vol;
    s->ring = ring;
    s->src_comp = sp->src_comp;
    if (has_src) {
        const int cell = (sp->src_k * p.iy + sp->src_j) * p.ix + sp->src_i;
        if (magnetic)
            s->src_h.push_back(cell);
        else
            s->src_cell = cell;
        s->signal.assign(sp->signal, sp->signal + sp->it);
    }
    for (int r = 0; r < sp->nrec; ++r) {
        const int *q = sp->rec_ijk + 3 * r;
        s->rec.push_back((q[2] * p.iy + q[1]) * p.ix + q[0]);
    }
    if (s->nf)
        s->omega.assign(sp->omega_vol, sp->omega_vol + s->nf);

    const size_t N = (size_t)p.ix * p.iy * p.iz;
    const size_t nc2 = 2 * (size_t)p.ncpml;
    const size_t slab[3] = {(size_t)p.iz * p.iy * nc2, (size_t)p.iz * nc2 * p.ix,
                            nc2 * p.iy * p.ix};
    const long long sstride = 3 * (long long)N + (ring > 1 ? env_pad("MITTET_FDTD_SLOT_PAD", kSlotPad) : 0);
    const long long astride = (long long)N + env_pad("MITTET_FDTD_ACC_PAD", kAccPad);
    if (sstride < 3 * (long long)N || astride < (long long)N) {
        delete s;
        return fail(FWI3D_EINVAL, "negative stride padding");
    }
    bool ok = dalloc(&s->d_e, (size_t)sstride * ring) == hipSuccess &&
              dalloc(&s->d_h, 3 * N) == hipSuccess &&
              dalloc(&s->d_psi[0], mt::PSI_N * slab[0]) == hipSuccess &&
              dalloc(&s->d_psi[1], mt::PSI_N * slab[1]) == hipSuccess &&
              dalloc(&s->d_psi[2], mt::PSI_N * slab[2]) == hipSuccess &&
              dalloc(&s->d_amp, (size_t)sp->it * (magnetic ? 4 : 1)) == hipSuccess &&
              dalloc(&s->d_traces, std::max<size_t>(1, 6 * (size_t)sp->nrec * sp->it)) == hipSuccess &&
              dalloc(&s->d_rec, std::max<size_t>(1, (size_t)sp->nrec)) == hipSuccess &&
              dalloc(&s->d_step, 1) == hipSuccess && dalloc(&s->d_flag, 1) == hipSuccess &&
              dalloc(&s->d_pre, 2) == hipSuccess &&
              dalloc(&s->d_desc, 1) == hipSuccess;
    if (ok && s->nf)
        ok = dalloc(&s->d_w, 2 * (size_t)s->nf * sp->it) == hipSuccess &&
             dalloc(&s->d_acc, 2 * 3 * (size_t)s->nf * astride) == hipSuccess;
    if (!ok) {
        m->nshots++;
        fwi3d_shot_destroy(s);
        return fail(FWI3D_ENOMEM, "device allocation failed for a %zu-cell shot (%lld bytes)", N,
                    fwi3d_shot_bytes(&p, sp));
    }
    m->nshots++;

    int rc = FWI3D_OK;
    const int zero2[2] = {0, 0};
    if (hcopy(s->d_pre, zero2, sizeof zero2) != hipSuccess)
        rc = fail(FWI3D_EDEVICE, "prefix upload failed");
    if (!rc && sp->nrec &&
        hcopy(s->d_rec, s->rec.data(), sizeof(int) * sp->nrec) != hipSuccess)
        rc = fail(FWI3D_EDEVICE, "receiver upload failed");
    if (!rc && s->nf) {
        // w[n][f] = dt*exp(-a n dt)*(cos, sin)(a n dt), a = sqrt(omega0*omega_f)
        std::vector<double> w(2 * (size_t)s->nf * sp->it);
        for (int f = 0; f < s->nf; ++f) {
            const double a = std::sqrt(p.omega0 * s->omega[f]);
            for (int n = 0; n < sp->it; ++n) {
                const double x = a * n * p.dt;
                const double e = p.dt * std::exp(-x);
                double sn, cs;
                sincos(x, &sn, &cs);  // one call: sin/cos and sincos may differ by an ulp
                w[2 * ((size_t)n * s->nf + f)] = e * cs;
                w[2 * ((size_t)n * s->nf + f) + 1] = e * sn;
            }
        }
        if (hcopy(s->d_w, w.data(), sizeof(double) * w.size()) != hipSuccess)
            rc = fail(FWI3D_EDEVICE, "weight upload failed");
    }
    if (rc) {
        fwi3d_shot_destroy(s);
        return rc;
    }

    ShotDev &d = s->desc;
    d.ix = p.ix;
    d.iy = p.iy;
    d.iz = p.iz;
    d.nc = p.ncpml;
    d.n = (int)N;
    d.nbx = tiles(p.ix, mt::TX);
    d.nby = tiles(p.iy, mt::TY);
    d.it = sp->it;
    d.dt = p.dt;
    d.omega0 = p.omega0;
    d.mu0 = p.mu0;
    d.c1 = p.c1;
    d.c2 = p.c2;
    d.dl[0] = p.dx;
    d.dl[1] = p.dy;
    d.dl[2] = p.dz;
    for (int a = 0; a < 3; ++a) {
        d.prof[a] = m->d_prof[a];
        d.psi[a] = s->d_psi[a];
        d.slab[a] = (long long)slab[a];
    }
    d.eps = m->d_eps;
    d.cbi = m->d_cbi;
    d.kb = sp->zchunk > 0 ? std::min(sp->zchunk, p.iz) : std::min(kDefaultZChunk, p.iz);
    d.ztx = tiles(p.ix, mt::ZX);
    d.zty = tiles(p.iy, mt::ZY);
    d.nkz = tiles(p.iz, d.kb);
    d.zb = d.ztx * d.zty * d.nkz;
    d.zb_pad = tiles(d.zb, 8) * 8;
    // x-pair kernels (two cells per lane) are opt-in: on 288^3 they run ~1.3x
    // slower than one cell per lane (more registers, fewer waves in flight)
    d.pair = (p.ix % 2 == 0 && env_pad("MITTET_FDTD_PAIR", 0)) ? 1 : 0;
    d.ftx = d.pair ? tiles(p.ix, 32) : tiles(p.ix, mt::ZX);
    d.fty = d.pair ? tiles(p.iy, 16) : tiles(p.iy, mt::ZY);
    d.fb = d.ftx * d.fty * p.iz;
    d.fb_pad = tiles(d.fb, 8) * 8;
    d.e = s->d_e;
    d.sstride = sstride;
    d.astride = astride;
    d.ring = ring;
    d.cur = 0;
    d.h = s->d_h;
    d.h2 = nullptr;
    d.hpar = 0;
    d.src_comp = s->src_comp;
    d.src_cell = s->src_cell;
    d.amp = s->d_amp;
    d.nrec = s->nrec;
    d.rec_cell = s->d_rec;
    d.traces = s->d_traces;
    d.nf = s->nf;
    d.w = s->d_w;
    d.acc = s->d_acc;
    d.acc_h = nullptr;
    d.hring = nullptr;
    d.step = s->d_step;
    d.flag = s->d_flag;
    d.variant = m->variant;
    d.mu = m->d_mu;
    d.nsrc_h = (int)s->src_h.size();
    for (int q = 0; q < d.nsrc_h; ++q)
        d.src_h[q] = s->src_h[q];
    s->nblocks_op = d.nbx * d.nby * p.iz;
    s->kind = sp->zchunk > 0 ? 2 : d.pair;
    if (magnetic) {
        // every 32 x 8 tile, one cell per lane (fdtd_fortran.hip)
        d.pair = 0;
        d.ftx = tiles(p.ix, mt::ZX);
        d.fty = tiles(p.iy, mt::ZY);
        d.fb = d.ftx * d.fty * p.iz;
        d.fb_pad = tiles(d.fb, 8) * 8;
        s->kind = 4;
    }
    // family 3: LDS z-marching kernels over the CPML-free interior (rows start
    // 128-byte aligned, at ncpml rounded up to 16 cells), per-cell kernels over
    // the tiles that hold any other cell.  Below kLdsMinCells the per-cell
    // family alone (two launches per step instead of four) is faster.
    d.shell = nullptr;
    const int xlo = (p.ncpml + 15) / 16 * 16;
    if (s->kind == 0 && takes_lds(p, sp->zchunk)) {
        const int lo[3] = {xlo, p.ncpml, p.ncpml};
        const int hi[3] = {p.ix - p.ncpml, p.iy - p.ncpml, p.iz - p.ncpml};
        for (int a = 0; a < 3; ++a) {
            d.lo[a] = lo[a];
            d.hi[a] = hi[a];
        }
        d.ltx = tiles(hi[0] - lo[0], mt::LX);
        d.lty = tiles(hi[1] - lo[1], mt::LY);
        d.lkb = lds_chunk(d.ltx, d.lty, hi[2] - lo[2]);
        d.lkz = tiles(hi[2] - lo[2], d.lkb);
        d.lb = d.ltx * d.lty * d.lkz;
        d.lb_pad = tiles(d.lb, 8) * 8;
        std::vector<int4> shell = complement_tiles(p.ix, p.iy, p.iz, lo, hi);
        d.fb = (int)shell.size();
        d.fb_pad = d.fb;
        hipError_t e = dalloc(&s->d_shell, shell.size());
        if (e == hipSuccess)
            e = hcopy(s->d_shell, shell.data(), sizeof(int4) * shell.size());
        if (e != hipSuccess) {
            fwi3d_shot_destroy(s);
            return fail(FWI3D_ENOMEM, "shell tile list: %s", hipGetErrorString(e));
        }
        d.shell = s->d_shell;
        s->kind = 3;
        s->nblocks_l = d.lb_pad;
        // the fused kernel takes one ked / khd per axis over the region (kappa = 1)
        bool flat = true;
        for (int a = 0; a < 3; ++a) {
            const int na = axis_len(p, a);
            const double *ked = &m->prof[a][(size_t)FWI3D_P_KED * na];
            const double *khd = &m->prof[a][(size_t)FWI3D_P_KHD * na];
            for (int q = lo[a]; q < hi[a]; ++q)
                flat = flat && ked[q] == ked[lo[a]] && khd[q] == khd[lo[a]];
        }
        if (flat && takes_fused(p, sp->zchunk, ring)) {
            if (dalloc(&s->d_h2, 3 * N) != hipSuccess) {
                fwi3d_shot_destroy(s);
                return fail(FWI3D_ENOMEM, "second H buffer (%zu cells)", N);
            }
            s->fused = true;
            d.h2 = s->d_h2;
        }
    }
    s->nblocks = s->kind == 2 ? d.zb_pad : d.fb_pad;
    s->ncellblocks = tiles((int)N, mt::FLUSH_CELLS);
    if ((rc = upload_desc(s)) || (rc = fwi3d_set_zero_eh(s))) {
        fwi3d_shot_destroy(s);
        return rc;
    }
    *out = s;
    return FWI3D_OK;
}

int fwi3d_shot_add_source(fwi3d_shot *s, int i, int j, int k)
{
    if (!s)
        return fail(FWI3D_EINVAL, "null shot");
    if (s->m->variant != 1)
        return fail(FWI3D_EINVAL, "extra source points belong to the Fortran variant (1)");
    if (s->batch)
        return fail(FWI3D_ESTATE, "shot is in a batch");
    if (s->signal.empty())
        return fail(FWI3D_ESTATE, "shot has no source waveform");
    const fwi3d_grid_params &p = s->m->p;
    if (i < 0 || i >= p.ix || j < 0 || j >= p.iy || k < 0 || k >= p.iz)
        return fail(FWI3D_EINVAL, "source cell (%d,%d,%d) outside the grid", i, j, k);
    if (s->src_h.size() >= 4)
        return fail(FWI3D_EINVAL, "at most 4 source points");
    s->src_h.push_back((k * p.iy + j) * p.ix + i);
    s->desc.nsrc_h = (int)s->src_h.size();
    s->desc.src_h[s->desc.nsrc_h - 1] = s->src_h.back();
    s->amp_version = -1;
    return upload_desc(s);
}

int fwi3d_shot_destroy(fwi3d_shot *s)
{
    if (!s)
        return FWI3D_OK;
    if (s->batch)
        fwi3d_batch_remove(s->batch, s);
    if (s->st)
        (void)hipStreamSynchronize(s->st);
    if (s->graph)
        (void)hipGraphExecDestroy(s->graph);
    dfree(s->d_e);
    dfree(s->d_h);
    dfree(s->d_h2);
    for (auto &q : s->d_psi)
        dfree(q);
    dfree(s->d_amp);
    dfree(s->d_traces);
    dfree(s->d_w);
    dfree(s->d_acc);
    dfree(s->d_acc_h);
    dfree(s->d_hring);
    dfree(s->d_rec);
    dfree(s->d_step);
    dfree(s->d_flag);
    dfree(s->d_pre);
    dfree(s->d_shell);
    dfree(s->d_desc);
    if (s->m)
        s->m->nshots--;
    delete s;
    return FWI3D_OK;
}

int fwi3d_set_zero_eh(fwi3d_shot *s)
{
    if (!s)
        return fail(FWI3D_EINVAL, "null shot");
    if (s->batch)
        return fail(FWI3D_ESTATE, "shot is in a batch");
    const fwi3d_grid_params &p = s->m->p;
    const size_t N = (size_t)p.ix * p.iy * p.iz;
    HIP_TRY(mt::launch_fill(s->d_e, (size_t)s->desc.sstride * s->ring, 0.0, s->st));
    HIP_TRY(mt::launch_fill(s->d_h, 3 * N, 0.0, s->st));
    if (s->d_h2)
        HIP_TRY(mt::launch_fill(s->d_h2, 3 * N, 0.0, s->st));
    for (int a = 0; a < 3; ++a)
        HIP_TRY(mt::launch_fill(s->d_psi[a], mt::PSI_N * (size_t)s->desc.slab[a], 0.0, s->st));
    HIP_TRY(mt::launch_fill(s->d_traces, std::max<size_t>(1, 6 * (size_t)s->nrec * s->it), 0.0, s->st));
    if (s->nf)
        HIP_TRY(mt::launch_fill(s->d_acc, 2 * 3 * (size_t)s->nf * s->desc.astride, 0.0, s->st));
    if (s->d_acc_h)
        HIP_TRY(mt::launch_fill(s->d_acc_h, 2 * 3 * (size_t)s->nf * s->desc.astride, 0.0, s->st));
    static const int kZero = 0;
    HIP_TRY(hipMemcpyAsync(s->d_step, &kZero, sizeof(int), hipMemcpyHostToDevice, s->st));
    HIP_TRY(hipMemcpyAsync(s->d_flag, &kZero, sizeof(int), hipMemcpyHostToDevice, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    s->step = 0;
    if (s->desc.cur != 0 || s->desc.hpar != 0) {
        s->desc.cur = 0;
        s->desc.hpar = 0;
        s->desc_dirty = true;
    }
    return FWI3D_OK;
}

static int single_op(fwi3d_shot *s, int which, int sign, int step)
{
    int rc = check_shot_ready(s);
    if (rc)
        return rc;
    if (s->m->variant != 0)
        return fail(FWI3D_EINVAL, "the per-operator calls follow the C reference; Fortran-variant "
                                  "shots run through fwi3d_step_n / fwi3d_run");
    if (s->batch)
        return fail(FWI3D_ESTATE, "shot is in a batch");
    if (sign != 1 && sign != -1)
        return fail(FWI3D_EINVAL, "sign must be +1 or -1");
    switch (which) {
    case 0:
        HIP_TRY(mt::launch_efield(mt::M_UPD | mt::M_INPLACE, s->d_desc, 1, s->d_pre, s->nblocks_op, 0,
                                  0, sign, s->st));
        break;
    case 1:
        HIP_TRY(mt::launch_efield(mt::M_PSI | mt::M_INPLACE, s->d_desc, 1, s->d_pre, s->nblocks_op, 0,
                                  0, 1, s->st));
        break;
    case 2:
        HIP_TRY(mt::launch_hfield(mt::M_UPD | mt::M_INPLACE, s->d_desc, 1, s->d_pre, s->nblocks_op, 0,
                                  sign, s->st));
        break;
    case 3:
        HIP_TRY(mt::launch_hfield(mt::M_PSI | mt::M_INPLACE, s->d_desc, 1, s->d_pre, s->nblocks_op, 0,
                                  1, s->st));
        break;
    case 4:
        if (step < 0 || step >= s->it - 1)
            return fail(FWI3D_EINVAL, "step %d outside [0, %d)", step, s->it - 1);
        HIP_TRY(mt::launch_source(s->d_desc, step, s->st));
        break;
    case 5:
        if (step < 0 || step >= s->it - 1)
            return fail(FWI3D_EINVAL, "step %d outside [0, %d)", step, s->it - 1);
        HIP_TRY(mt::launch_receivers(s->d_desc, 1, 0, step, 1, s->st));
        break;
    }
    return FWI3D_OK;
}

int fwi3d_e_field_cpml42(fwi3d_shot *s, int sign) { return single_op(s, 0, sign, 0); }
int fwi3d_epml_abcs4(fwi3d_shot *s) { return single_op(s, 1, 1, 0); }
int fwi3d_h_field_cpml42(fwi3d_shot *s, int sign) { return single_op(s, 2, sign, 0); }
int fwi3d_hpml_abcs4(fwi3d_shot *s) { return single_op(s, 3, 1, 0); }
int fwi3d_read_source_3d(fwi3d_shot *s, int step) { return single_op(s, 4, 1, step); }
int fwi3d_output(fwi3d_shot *s, int step) { return single_op(s, 5, 1, step); }

// Bytes per cell one step touches: E read, E written (to the next ring slot,
// or in place with one slot), H, and the two coefficient arrays.
long long step_touch_bytes(const fwi3d_shot *s)
{
    return (long long)s->desc.n * (s->ring > 1 ? 88 : 64);
}

// Non-temporal field traffic pays once a launch's fields no longer stay in the
// Infinity Cache from one half-step to the next: then L2 and that cache keep
// the stencil neighbours that other tiles still read.  Single shots with a
// 12-cell CPML and a 4-frequency volume DFT (88 B per cell), Gcell-updates/s
// without / with: 96^3 17.7 / 15.2, 120^3 19.0 / 15.6, 136^3 (211 MiB) 16.3 /
// 15.5, 152^3 (295 MiB) 18.4 / 18.8, 184^3 17.9 / 19.0, 216^3 19.3 / 20.6,
// 280^3 20.2 / 21.4-23.1; 84^3 with a 10-cell CPML 14.8 / 12.9.  Without DFT
// (one slot, 64 B per cell): 184^3 20.3 / 21.4, 216^3 22.2 / 23.5, 280^3
// 23.1 / 25.2, while eight 152^3 shots (214 MiB) run their forward in 4.79 s
// without, 5.23 s with.  MITTET_FDTD_NT=0 or 1 overrides the choice.
bool stream_nt(long long touch_bytes)
{
    const long long e = env_pad("MITTET_FDTD_NT", -1);
    if (e >= 0)
        return e != 0;
    return touch_bytes > kNtMinBytes;
}

static Fam shot_fam(const fwi3d_shot *s)
{
    return {s->kind, s->d_pre, s->nblocks, s->d_pre, s->nblocks_l, s->ring > 1,
            s->kind == 3 && s->fused, stream_nt(step_touch_bytes(s))};
}

// After `steps` steps of a fused schedule the current H is in the other buffer
// when `steps` is odd, and E sits `steps` ring slots further on.  Returns
// whether the device descriptor changed.
static bool advance_desc(fwi3d_shot *s, int steps, bool fused)
{
    bool moved = false;
    if (s->ring > 1 && steps % s->ring) {
        s->desc.cur = (s->desc.cur + steps) % s->ring;
        moved = true;
    }
    if (fused && (steps & 1)) {
        s->desc.hpar ^= 1;
        moved = true;
    }
    return moved;
}

int fwi3d_step_n(fwi3d_shot *s, int nsteps)
{
    int rc = check_shot_ready(s);
    if (rc)
        return rc;
    if (s->batch)
        return fail(FWI3D_ESTATE, "shot is in a batch");
    if (nsteps < 0)
        return fail(FWI3D_EINVAL, "nsteps < 0");
    const int chunk = s->nf ? s->ring : kDefaultChunk;
    const bool dft = s->nf > 0, rec = s->nrec > 0;
    int left = std::min(nsteps, s->it - 1 - s->step);
    while (left > 0) {
        const int L = std::min(chunk, left);
        if (L == chunk) {
            if (!s->graph && (rc = capture_chunk(&s->graph, s->d_desc, 1, shot_fam(s), s->d_pre,
                                                 s->ncellblocks, L, dft, rec, s->st,
                                                 s->d_acc_h != nullptr)))
                return rc;
            HIP_TRY(hipGraphLaunch(s->graph, s->st));
        } else {
            if ((rc = launch_chunk(s->d_desc, 1, shot_fam(s), s->d_pre, s->ncellblocks, L, dft, rec,
                                   s->st, s->d_acc_h != nullptr)))
                return rc;
        }
        s->step += L;
        left -= L;
        if (advance_desc(s, L, shot_fam(s).fused) && (rc = upload_desc(s)))
            return rc;
    }
    return FWI3D_OK;
}

int fwi3d_profile_steps(fwi3d_shot *s, int nsteps, double *ms)
{
    int rc = check_shot_ready(s);
    if (rc)
        return rc;
    if (!ms || nsteps <= 0)
        return fail(FWI3D_EINVAL, "bad profile request");
    if (s->batch)
        return fail(FWI3D_ESTATE, "shot is in a batch");
    const int chunk = s->nf ? s->ring : kDefaultChunk;
    hipEvent_t ev[2];
    HIP_TRY(hipEventCreate(&ev[0]));
    HIP_TRY(hipEventCreate(&ev[1]));
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const Fam f = shot_fam(s);
    const int hf = f.fused ? 1 : 0;
    // the E half-step's launches without the fused interior launch, timed apart
    Fam fe = f;
    fe.nblocks_l = f.fused ? 0 : f.nblocks_l;
    auto timed = [&](int slot, auto &&launch) -> int {
        HIP_TRY(hipEventRecord(ev[0], s->st));
        HIP_TRY(launch());
        HIP_TRY(hipEventRecord(ev[1], s->st));
        HIP_TRY(hipEventSynchronize(ev[1]));
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, ev[0], ev[1]));
        acc[slot] += t;
        cnt[slot] += 1;
        return FWI3D_OK;
    };
    int left = std::min(nsteps, s->it - 1 - s->step);
    const int steps_run = left;
    while (left > 0 && !rc) {
        const int L = std::min(chunk, left);
        for (int j = 0; j < L && !rc; ++j) {
            rc = timed(0, [&] {
                return launch_eh(fe, s->d_desc, 1, j, false, s->st) ? hipErrorUnknown : hipSuccess;
            });
            if (!rc && f.fused)
                rc = timed(7, [&] {
                    return mt::launch_ehfield_l(s->d_desc, 1, s->d_pre, s->nblocks_l, j, s->st);
                });
            if (!rc)
                rc = timed(1, [&] {
                    return launch_eh(f, s->d_desc, 1, j, true, s->st) ? hipErrorUnknown : hipSuccess;
                });
            if (!rc && s->nrec)
                rc = timed(2, [&] { return mt::launch_receivers(s->d_desc, 1, j, -1, 0, s->st, hf); });
            if (!rc && s->d_acc_h)
                rc = timed(5, [&] {
                    return mt::launch_hstore(s->d_desc, 1, s->d_pre, s->ncellblocks, j, s->st, hf);
                });
        }
        if (!rc && s->nf)
            rc = timed(3, [&] { return mt::launch_flush(s->d_desc, 1, s->d_pre, s->ncellblocks, L, s->st); });
        if (!rc && s->d_acc_h)
            rc = timed(6, [&] {
                return mt::launch_flush(s->d_desc, 1, s->d_pre, s->ncellblocks, L, s->st, true);
            });
        if (!rc)
            rc = timed(4, [&] { return mt::launch_advance(s->d_desc, 1, L, s->st, hf); });
        if (rc)
            break;
        s->step += L;
        left -= L;
        if (advance_desc(s, L, f.fused))
            rc = upload_desc(s);
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (rc)
        return rc;
    for (int q = 0; q < 8; ++q)
        ms[q] = cnt[q] ? acc[q] / cnt[q] : 0.0;
    double tot = 0.0;
    for (int q = 0; q < 8; ++q)
        tot += acc[q];
    ms[4] = steps_run ? tot / steps_run : 0.0;
    ms[8] = f.fused ? 1.0 : 0.0;
    return FWI3D_OK;
}

int fwi3d_diag_copy(long long bytes, int reps, double *ms_out)
{
    if (bytes < 16 || reps <= 0)
        return fail(FWI3D_EINVAL, "bad copy size");
    const size_t count = (size_t)(bytes / 16) * 2;
    double *a = nullptr, *b = nullptr;
    if (dalloc(&a, count) != hipSuccess || dalloc(&b, count) != hipSuccess) {
        dfree(a);
        return fail(FWI3D_ENOMEM, "copy buffers");
    }
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = FWI3D_OK;
    float total = 0.f;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = fail(FWI3D_EDEVICE, "stream/event creation");
    if (!rc && mt::launch_fill(a, count, 1.0, st) != hipSuccess)
        rc = fail(FWI3D_EDEVICE, "fill");
    for (int r = 0; r < reps && !rc; ++r) {
        float t = 0.f;
        if (hipEventRecord(e0, st) != hipSuccess || mt::launch_diag_copy(a, b, count, st) != hipSuccess ||
            hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&t, e0, e1) != hipSuccess)
            rc = fail(FWI3D_EDEVICE, "copy launch");
        total += t;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    dfree(a);
    dfree(b);
    if (!rc && ms_out)
        *ms_out = total / reps;
    return rc;
}

int fwi3d_run(fwi3d_shot *s)
{
    if (!s)
        return fail(FWI3D_EINVAL, "null shot");
    return fwi3d_step_n(s, s->it - 1 - s->step);
}

int fwi3d_shot_sync(fwi3d_shot *s)
{
    if (!s)
        return fail(FWI3D_EINVAL, "null shot");
    HIP_TRY(hipStreamSynchronize(s->st));
    return FWI3D_OK;
}

int fwi3d_shot_step(const fwi3d_shot *s) { return s ? s->step : -1; }

int fwi3d_shot_diverged(fwi3d_shot *s)
{
    if (!s)
        return fail(FWI3D_EINVAL, "null shot");
    hipStream_t st = s->batch ? s->batch->st : s->st;
    int flag = 0;
    HIP_TRY(hipMemcpyAsync(&flag, s->d_flag, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return flag ? 1 : 0;
}

int fwi3d_get_field(fwi3d_shot *s, int which, double *out)
{
    if (!s || !out || which < 0 || which > 5)
        return fail(FWI3D_EINVAL, "bad field request");
    const size_t N = (size_t)s->desc.n;
    const double *src = which < 3 ? s->d_e + (size_t)s->desc.cur * s->desc.sstride + (size_t)which * N
                                  : mt::hsel(s->desc, 0, 0) + (size_t)(which - 3) * N;
    HIP_TRY(hipMemcpyAsync(out, src, sizeof(double) * N, hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    return FWI3D_OK;
}

int fwi3d_set_field(fwi3d_shot *s, int which, const double *in)
{
    if (!s || !in || which < 0 || which > 5)
        return fail(FWI3D_EINVAL, "bad field request");
    const size_t N = (size_t)s->desc.n;
    double *dst = which < 3 ? s->d_e + (size_t)s->desc.cur * s->desc.sstride + (size_t)which * N
                            : mt::hsel(s->desc, 0, 0) + (size_t)(which - 3) * N;
    HIP_TRY(hipMemcpyAsync(dst, in, sizeof(double) * N, hipMemcpyHostToDevice, s->st));
    // both H buffers of a fused shot: cells no step updates must agree in them
    if (which >= 3 && s->d_h2) {
        double *other = (dst == s->d_h + (size_t)(which - 3) * N ? s->d_h2 : s->d_h) +
                        (size_t)(which - 3) * N;
        HIP_TRY(hipMemcpyAsync(other, in, sizeof(double) * N, hipMemcpyHostToDevice, s->st));
    }
    HIP_TRY(hipStreamSynchronize(s->st));
    return FWI3D_OK;
}

int fwi3d_get_traces(fwi3d_shot *s, double *out)
{
    if (!s || !out)
        return fail(FWI3D_EINVAL, "null argument");
    if (s->nrec == 0)
        return FWI3D_OK;
    HIP_TRY(hipMemcpyAsync(out, s->d_traces, sizeof(double) * 6 * (size_t)s->nrec * s->it,
                           hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    return FWI3D_OK;
}

int fwi3d_get_volume_dft(fwi3d_shot *s, double *out)
{
    if (!s || !out)
        return fail(FWI3D_EINVAL, "null argument");
    if (!s->nf)
        return fail(FWI3D_ESTATE, "shot has no volume DFT frequencies");
    HIP_TRY(hipMemcpy2DAsync(out, sizeof(double) * 2 * (size_t)s->desc.n, s->d_acc,
                             sizeof(double) * 2 * (size_t)s->desc.astride,
                             sizeof(double) * 2 * (size_t)s->desc.n, 3 * (size_t)s->nf,
                             hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    return FWI3D_OK;
}

int fwi3d_enable_volume_dft_h(fwi3d_shot *s)
{
    if (!s)
        return fail(FWI3D_EINVAL, "null shot");
    if (!s->nf)
        return fail(FWI3D_ESTATE, "shot has no volume DFT frequencies");
    if (s->batch)
        return fail(FWI3D_ESTATE, "shot is in a batch");
    if (s->step != 0)
        return fail(FWI3D_ESTATE, "enable the H volume DFT before the first step (step %d)",
                    s->step);
    if (s->d_acc_h)
        return FWI3D_OK;
    const size_t count = 2 * 3 * (size_t)s->nf * s->desc.astride;
    const size_t ring = (size_t)s->desc.sstride * s->ring;
    if (dalloc(&s->d_acc_h, count) != hipSuccess)
        return fail(FWI3D_ENOMEM, "H volume accumulators (%zu doubles)", count);
    if (dalloc(&s->d_hring, ring) != hipSuccess) {
        dfree(s->d_acc_h);
        s->d_acc_h = nullptr;
        return fail(FWI3D_ENOMEM, "H ring (%zu doubles)", ring);
    }
    HIP_TRY(mt::launch_fill(s->d_acc_h, count, 0.0, s->st));
    s->desc.acc_h = s->d_acc_h;
    s->desc.hring = s->d_hring;
    s->desc_dirty = true;
    if (s->graph) {   // recaptured with the H DFT launches
        (void)hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
    }
    return upload_desc(s);
}

int fwi3d_get_volume_dft_h(fwi3d_shot *s, double *out)
{
    if (!s || !out)
        return fail(FWI3D_EINVAL, "null argument");
    if (!s->d_acc_h)
        return fail(FWI3D_ESTATE, "shot has no H volume DFT (fwi3d_enable_volume_dft_h)");
    HIP_TRY(hipMemcpy2DAsync(out, sizeof(double) * 2 * (size_t)s->desc.n, s->d_acc_h,
                             sizeof(double) * 2 * (size_t)s->desc.astride,
                             sizeof(double) * 2 * (size_t)s->desc.n, 3 * (size_t)s->nf,
                             hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    return FWI3D_OK;
}

// Scratch of fwi3d_receiver_dft_noalloc: a fixed device array instead of
// per-call buffers.  The name is a host-side stand-in, not a device address:
// host code takes the address from hipGetSymbolAddress, once per device with
// that device current, and uses only that pointer.
__device__ __align__(256) double scratch[FWI3D_NOALLOC_SCRATCH_DOUBLES];
static constexpr int kScratchDevices = 64;
static double *g_scratch[kScratchDevices];
static std::mutex g_scratch_mu;

static int noalloc_scratch(double **out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices)
        return fail(FWI3D_EDEVICE, "device %d beyond the %d the scratch table holds", dev,
                    kScratchDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        double *p = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(FWI3D_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return FWI3D_OK;
}

static int device_rows_dft(hipStream_t st, const double *d_rows, int nrows, int it, double dt,
                           double omega0, int nfreq, const double *omegas, double *out_host)
{
    double *d_om = nullptr, *d_out = nullptr;
    if (dalloc(&d_om, (size_t)nfreq) != hipSuccess ||
        dalloc(&d_out, 2 * (size_t)nrows * nfreq) != hipSuccess) {
        dfree(d_om);
        return fail(FWI3D_ENOMEM, "spectrum buffers");
    }
    hipError_t e = hipMemcpyAsync(d_om, omegas, sizeof(double) * nfreq, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = mt::launch_receiver_dft(d_rows, nrows, it, dt, omega0, nfreq, d_om, d_out, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out_host, d_out, sizeof(double) * 2 * nrows * nfreq,
                           hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    dfree(d_om);
    dfree(d_out);
    if (e != hipSuccess)
        return fail(FWI3D_EDEVICE, "receiver DFT: %s", hipGetErrorString(e));
    return FWI3D_OK;
}

static std::complex<double> h_factor(double omega0, double om)
{
    // sqrt(-2*omega0/(i*omega)), ver0040/cpu_f_to_d_h.f90:146-148.  At omega = 0
    // (the reference's bin k = 0) the factor divides by zero and grows as
    // omega^(-1/2): no finite value exists, so the bin is NaN, not inf.
    if (om == 0.0)
        return {std::nan(""), std::nan("")};
    return std::sqrt(std::complex<double>(-2.0 * omega0, 0.0) / std::complex<double>(0.0, 1.0) / om);
}

// the f_to_d_h factor on the magnetic rows of receiver spectra (host)
static void h_factor_rows(const fwi3d_shot *s, int nfreq, const double *omegas, double *out)
{
    for (int f = 0; f < nfreq; ++f) {
        const std::complex<double> k = h_factor(s->m->p.omega0, omegas[f]);
        for (int row = 3 * s->nrec; row < 6 * s->nrec; ++row) {
            std::complex<double> v(out[2 * ((size_t)row * nfreq + f)],
                                   out[2 * ((size_t)row * nfreq + f) + 1]);
            v *= k;
            out[2 * ((size_t)row * nfreq + f)] = v.real();
            out[2 * ((size_t)row * nfreq + f) + 1] = v.imag();
        }
    }
}

int fwi3d_receiver_dft(fwi3d_shot *s, int nfreq, const double *omegas, int hfac, double *out)
{
    if (!s || !omegas || !out || nfreq <= 0)
        return fail(FWI3D_EINVAL, "bad receiver DFT request");
    if (s->nrec == 0)
        return FWI3D_OK;
    int rc = device_rows_dft(s->st, s->d_traces, 6 * s->nrec, s->it, s->m->p.dt, s->m->p.omega0,
                             nfreq, omegas, out);
    if (rc || !hfac)
        return rc;
    h_factor_rows(s, nfreq, omegas, out);
    return FWI3D_OK;
}

int fwi3d_receiver_dft_noalloc(fwi3d_shot *s, int nfreq, const double *omegas, int hfac,
                               double *out)
{
    if (!s || !omegas || !out || nfreq <= 0)
        return fail(FWI3D_EINVAL, "bad receiver DFT request");
    if (s->nrec == 0)
        return FWI3D_OK;
    const size_t rows = 6 * (size_t)s->nrec;
    const size_t off = ((size_t)nfreq + 31) / 32 * 32;   // spectra start 256-byte aligned
    const size_t need = off + 2 * rows * (size_t)nfreq;
    if (need > (size_t)FWI3D_NOALLOC_SCRATCH_DOUBLES)
        return fail(FWI3D_EINVAL,
                    "receiver DFT of %zu rows x %d frequencies needs %zu scratch doubles; the "
                    "allocation-free call holds %d",
                    rows, nfreq, need, FWI3D_NOALLOC_SCRATCH_DOUBLES);
    double *d = nullptr;
    int rc = noalloc_scratch(&d);
    if (rc)
        return rc;
    // every output element is written by the kernel: nothing needs clearing
    HIP_TRY(hipMemcpyAsync(d, omegas, sizeof(double) * nfreq, hipMemcpyHostToDevice, s->st));
    HIP_TRY(mt::launch_receiver_dft(s->d_traces, (int)rows, s->it, s->m->p.dt, s->m->p.omega0,
                                    nfreq, d, d + off, s->st));
    HIP_TRY(hipMemcpyAsync(out, d + off, sizeof(double) * 2 * rows * nfreq, hipMemcpyDeviceToHost,
                           s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    if (hfac)
        h_factor_rows(s, nfreq, omegas, out);
    return FWI3D_OK;
}

int fwi3d_f_to_d_h(double omega0, int nfreq, const double *omegas, long long count, double *sp)
{
    if (!omegas || !sp || nfreq <= 0 || count < 0)
        return fail(FWI3D_EINVAL, "bad f_to_d_h request");
    for (long long q = 0; q < count; ++q)
        for (int f = 0; f < nfreq; ++f) {
            const std::complex<double> k = h_factor(omega0, omegas[f]);
            std::complex<double> v(sp[2 * (q * nfreq + f)], sp[2 * (q * nfreq + f) + 1]);
            v *= k;
            sp[2 * (q * nfreq + f)] = v.real();
            sp[2 * (q * nfreq + f) + 1] = v.imag();
        }
    return FWI3D_OK;
}

int fwi3d_source_spectrum(double dt, double omega0, int it, const double *signal, int nfreq,
                          const double *omegas, double *out)
{
    if (!signal || !omegas || !out || it < 1 || nfreq <= 0)
        return fail(FWI3D_EINVAL, "bad source spectrum request");
    double *d_sig = nullptr;
    hipStream_t st = nullptr;
    if (new_stream(&st) != hipSuccess)
        return fail(FWI3D_EDEVICE, "stream for the source spectrum");
    int rc = FWI3D_OK;
    if (dalloc(&d_sig, (size_t)it) != hipSuccess)
        rc = fail(FWI3D_ENOMEM, "signal buffer");
    else if (hcopy(d_sig, signal, sizeof(double) * it) != hipSuccess)
        rc = fail(FWI3D_EDEVICE, "signal upload");
    if (!rc)
        rc = device_rows_dft(st, d_sig, 1, it, dt, omega0, nfreq, omegas, out);
    dfree(d_sig);
    del_stream(st);
    if (rc)
        return rc;
    for (int f = 0; f < nfreq; ++f) {
        std::complex<double> v(out[2 * f], out[2 * f + 1]);
        v = omegas[f] == 0.0 ? std::complex<double>(2.0 * omega0, 0.0) : v * h_factor(omega0, omegas[f]);
        out[2 * f] = v.real();
        out[2 * f + 1] = v.imag();
    }
    return FWI3D_OK;
}

}  // extern "C"

// ------------------------------------------------- complex-row DFTs on device

// DFT of `rows` complex device rows of length nd into d_out, on stream st
static int dft_rows_dev(hipStream_t st, const double *d_in, int rows, int nd, int sign,
                        double scale, double *d_out)
{
    double *d_tw = nullptr;
    if (dalloc(&d_tw, 2 * (size_t)nd) != hipSuccess)
        return fail(FWI3D_ENOMEM, "twiddle table");
    hipError_t e = mt::launch_twiddle(nd, d_tw, st);
    if (e == hipSuccess)
        e = mt::launch_dft_rows(d_in, rows, nd, sign, scale, d_tw, d_out, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    dfree(d_tw);
    if (e != hipSuccess)
        return fail(FWI3D_EDEVICE, "row DFT: %s", hipGetErrorString(e));
    return FWI3D_OK;
}

// Host rows in -> device DFT -> host rows out, on a private stream.  With
// `jrow` (complex, nd), that row is transformed the same way and multiplies
// every transformed row before the download.
static int dft_rows_host(const double *in, int rows, int nd, int sign, double scale, double *out,
                         const double *jrow = nullptr)
{
    const size_t n = 2 * (size_t)rows * nd;
    hipStream_t st = nullptr;
    double *d_in = nullptr, *d_out = nullptr, *d_j = nullptr, *d_jw = nullptr;
    int rc = FWI3D_OK;
    if (new_stream(&st) != hipSuccess)
        return fail(FWI3D_EDEVICE, "stream for a row DFT");
    if (dalloc(&d_in, n) != hipSuccess || dalloc(&d_out, n) != hipSuccess ||
        (jrow && (dalloc(&d_j, 2 * (size_t)nd) != hipSuccess ||
                  dalloc(&d_jw, 2 * (size_t)nd) != hipSuccess)))
        rc = fail(FWI3D_ENOMEM, "row DFT buffers (%d x %d)", rows, nd);
    if (!rc && (hcopy(d_in, in, sizeof(double) * n) != hipSuccess ||
                (jrow && hcopy(d_j, jrow, sizeof(double) * 2 * nd) != hipSuccess)))
        rc = fail(FWI3D_EDEVICE, "row DFT upload");
    if (!rc)
        rc = dft_rows_dev(st, d_in, rows, nd, sign, scale, d_out);
    if (!rc && jrow)
        rc = dft_rows_dev(st, d_j, 1, nd, sign, scale, d_jw);
    if (!rc) {
        hipError_t e = jrow ? mt::launch_cmul_rows(d_out, d_jw, rows, nd, st) : hipSuccess;
        if (e == hipSuccess)
            e = hipMemcpyAsync(out, d_out, sizeof(double) * n, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
        if (e != hipSuccess)
            rc = fail(FWI3D_EDEVICE, "row DFT: %s", hipGetErrorString(e));
    }
    dfree(d_in);
    dfree(d_out);
    dfree(d_j);
    dfree(d_jw);
    del_stream(st);
    return rc;
}

extern "C" {

int fwi3d_dft_rows(const double *in, int rows, int nd, int sign, double scale, double *out)
{
    if (!in || !out || rows <= 0 || nd <= 0 || (sign != 1 && sign != -1))
        return fail(FWI3D_EINVAL, "bad row DFT request");
    if (rows > 65535)
        return fail(FWI3D_EINVAL, "row DFT: %d rows exceed 65535", rows);
    return dft_rows_host(in, rows, nd, sign, scale, out);
}

int fwi3d_convolve_gj(const double *g_t, int rows, int nd, const double *j_t, double *out)
{
    if (!g_t || !j_t || !out || rows <= 0 || nd <= 0)
        return fail(FWI3D_EINVAL, "bad G*J request");
    if (rows > 65535)
        return fail(FWI3D_EINVAL, "G*J: %d rows exceed 65535", rows);
    std::vector<double> jc(2 * (size_t)nd, 0.0);
    for (int n = 0; n < nd; ++n)
        jc[2 * n] = j_t[n];
    return dft_rows_host(g_t, rows, nd, 1, 1.0, out, jc.data());
}

int fwi3d_laplace_fft(fwi3d_shot *s, double *e_t, double *j_t, double *g_t)
{
    if (!s || !e_t || !j_t || !g_t)
        return fail(FWI3D_EINVAL, "null argument");
    if (s->signal.empty())
        return fail(FWI3D_ESTATE, "shot has no source waveform");
    if (s->nrec == 0)
        return fail(FWI3D_ESTATE, "shot has no receivers");
    const int it = s->it, rows = 3 * s->nrec;
    if (2 * rows + 1 > 65535)
        return fail(FWI3D_EINVAL, "laplace_fft: too many receivers");
    const double dt = s->m->p.dt, omega0 = s->m->p.omega0;
    const double om = 2.0 * M_PI / it / dt;
    std::vector<double> omg(it);
    for (int n = 0; n < it; ++n)
        omg[n] = om * n;
    // spectra at every bin: rows of E, then G = E/J, then J
    std::vector<double> spec(2 * (size_t)(2 * rows + 1) * it);
    double *ew = spec.data(), *gw = ew + 2 * (size_t)rows * it, *jw = gw + 2 * (size_t)rows * it;
    int rc = device_rows_dft(s->st, s->d_traces, rows, it, dt, omega0, it, omg.data(), ew);
    if (!rc)
        rc = fwi3d_source_spectrum(dt, omega0, it, s->signal.data(), it, omg.data(), jw);
    if (rc)
        return rc;
    for (int r = 0; r < rows; ++r)
        for (int n = 0; n < it; ++n) {
            const size_t q = (size_t)r * it + n;
            const std::complex<double> g = std::complex<double>(ew[2 * q], ew[2 * q + 1]) /
                                           std::complex<double>(jw[2 * n], jw[2 * n + 1]);
            gw[2 * q] = g.real();
            gw[2 * q + 1] = g.imag();
        }
    std::vector<double> t(spec.size());
    if ((rc = dft_rows_host(spec.data(), 2 * rows + 1, it, -1, 2.0 / it / dt, t.data())))
        return rc;
    const size_t blk = 2 * (size_t)rows * it;
    std::memcpy(e_t, t.data(), sizeof(double) * blk);
    std::memcpy(g_t, t.data() + blk, sizeof(double) * blk);
    std::memcpy(j_t, t.data() + 2 * blk, sizeof(double) * 2 * it);
    return FWI3D_OK;
}

}  // extern "C"

// FFTW_BACKWARD (un-normalised exp(+2 pi i q n / it)) of the true waveform at
// bins n0 <= n < n1
static std::vector<std::complex<double>> true_spectrum(const double *strue, int it, int n0, int n1)
{
    std::vector<std::complex<double>> tru(std::max(0, n1 - n0));
    for (int n = n0; n < n1; ++n) {
        double re = 0.0, im = 0.0;
        for (int q = 0; q < it; ++q) {
            const long long mm = ((long long)q * n) % it;
            const double ang = 2.0 * M_PI * (double)mm / (double)it;
            re += strue[q] * std::cos(ang);
            im += strue[q] * std::sin(ang);
        }
        tru[n - n0] = {re, im};
    }
    return tru;
}

extern "C" {

int fwi3d_laplace_to_freq(fwi3d_shot *s, int nbins, const double *strue, double *e_w, double *j_w,
                          double *g_w, double *obs)
{
    if (!s || !strue || !e_w || !j_w || !g_w || !obs || nbins <= 0 || nbins > s->it)
        return fail(FWI3D_EINVAL, "bad laplace_to_freq request");
    if (s->signal.empty())
        return fail(FWI3D_ESTATE, "shot has no source waveform");
    const int it = s->it;
    const double dt = s->m->p.dt, omega0 = s->m->p.omega0;
    const double om = 2.0 * M_PI / it / dt;
    std::vector<double> omg(nbins);
    for (int n = 0; n < nbins; ++n)
        omg[n] = om * n;
    int rc = FWI3D_OK;
    if (s->nrec)
        rc = device_rows_dft(s->st, s->d_traces, 3 * s->nrec, it, dt, omega0, nbins, omg.data(), e_w);
    if (!rc)
        rc = fwi3d_source_spectrum(dt, omega0, it, s->signal.data(), nbins, omg.data(), j_w);
    if (rc)
        return rc;
    const std::vector<std::complex<double>> tru = true_spectrum(strue, it, 0, nbins);
    for (int row = 0; row < 3 * s->nrec; ++row)
        for (int n = 0; n < nbins; ++n) {
            const size_t q = (size_t)row * nbins + n;
            const std::complex<double> e(e_w[2 * q], e_w[2 * q + 1]), jv(j_w[2 * n], j_w[2 * n + 1]);
            const std::complex<double> g = e / jv, o = g * tru[n];
            g_w[2 * q] = g.real();
            g_w[2 * q + 1] = g.imag();
            obs[2 * q] = o.real();
            obs[2 * q + 1] = o.imag();
        }
    return FWI3D_OK;
}


// ---------------------------------------------------------------- gradients

// Checkpoint plan of a gradient (recursive recomputation).  The forward pass
// stores the state at the start of every top segment (length seglen[0]).  The
// backward sweep takes the segments in reverse; at depth l < levels - 1 it
// re-runs a segment once from its stored start, storing the starts of its
// pieces (length seglen[l + 1]) in that depth's slots, and recurses into the
// pieces in reverse; at depth levels - 1 it replays the segment keeping every
// state and runs the adjoint steps over it.  levels = 1 is the classic
// sqrt(Nt) scheme; every extra level costs one more forward sweep and divides
// the stored states by about Nt^(1/(levels(levels+1))).
struct GradPlan {
    int levels = 1;
    int k = 1;                        // pieces per segment below the top
    std::vector<int> seglen;          // [levels]: segment length at each depth
    int nseg = 0;                     // top segments
    size_t states = 0;                // stored states: top + inner depths + leaf buffer + adjoint
};

static GradPlan grad_plan(int nt, int levels, int segment)
{
    GradPlan P;
    P.levels = std::max(1, levels);
    if (P.levels == 1) {
        const int seg = segment > 0 ? std::min(segment, nt)
                                    : std::max(1, (int)std::ceil(std::sqrt((double)nt)));
        P.k = seg;
        P.seglen = {seg};
    } else {
        // smallest k with k^(levels + 1) >= nt
        int k = std::max(1, (int)std::floor(std::pow((double)nt, 1.0 / (P.levels + 1))));
        while (std::pow((double)k, P.levels + 1) < (double)nt)
            ++k;
        P.k = k;
        P.seglen.resize(P.levels);
        for (int l = P.levels - 1; l >= 0; --l) {
            const double len = std::pow((double)k, P.levels - l);
            P.seglen[l] = (int)std::min<double>(len, nt);
        }
    }
    P.nseg = (nt + P.seglen[0] - 1) / P.seglen[0];
    size_t inner = 0;
    for (int l = 1; l < P.levels; ++l)
        inner += (size_t)((P.seglen[l - 1] + P.seglen[l] - 1) / P.seglen[l]);
    P.states = (size_t)P.nseg + inner + (size_t)P.seglen[P.levels - 1] + 1 + 1;
    return P;
}

struct fwi3d_grad {
    fwi3d_shot *s = nullptr;
    GradPlan plan;
    int seg = 0, nseg = 0, nt = 0;    // top segment length and count
    int leaf = 0;                     // steps replayed with every state kept
    size_t state = 0;                 // doubles per stored state
    size_t off_h = 0, off_psi[3] = {0, 0, 0};
    double *ckpt = nullptr;           // nseg states
    std::vector<double *> inner;      // depth l >= 1: slots for the starts of a segment's pieces
    double *buf = nullptr;            // leaf + 1 states
    double *adj = nullptr;            // adjoint state (one state)
    double *tmp = nullptr;            // 3N + psi slabs
    double *grad = nullptr;           // N
    double *grad_src = nullptr;       // 1
    double *g = nullptr;              // [it][3][nrec] trace sensitivities
    double *rp = nullptr;             // per axis: 1/ked, 1/khd
    int kappa_one = 0;                // ked = khd = dl on every axis
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};   // LDS region of the adjoint (empty: lo = hi)
    int4 *tiles = nullptr;            // per-cell tiles outside it (ntiles, multiple of 8)
    int ntiles = 0;
    hipGraphExec_t replay[2] = {nullptr, nullptr};   // segment replays: full length, last
    int replay_len[2] = {0, 0};
    ShotDev rdesc{};                  // the shot's descriptor with E in buf as a ring (replays)
    ShotDev *d_rdesc = nullptr;
};

// doubles between stored states: the content rounded up to 256 bytes, so that
// each state (whose E a replay writes as a ring slot) starts as aligned as the
// shot's own buffers
static size_t state_pad(size_t doubles)
{
    return (doubles + 31) / 32 * 32;
}

static void state_layout(fwi3d_grad *g)
{
    const ShotDev &d = g->s->desc;
    const size_t N = (size_t)d.n;
    g->off_h = 3 * N;
    size_t o = 6 * N;
    for (int a = 0; a < 3; ++a) {
        g->off_psi[a] = o;
        o += mt::PSI_N * (size_t)d.slab[a];
    }
    g->state = state_pad(o);
}

// the shot's live state (E slot 0 -- ring 1 -- H, the psi slabs) in the
// stored-state order
static mt::StateSegs state_segs(const fwi3d_grad *g)
{
    const fwi3d_shot *s = g->s;
    const size_t N = (size_t)s->desc.n;
    mt::StateSegs v{};
    v.live[0] = s->d_e;
    v.count[0] = 3 * N;
    v.live[1] = mt::hsel(s->desc, 0, 0);
    v.count[1] = 3 * N;
    for (int a = 0; a < 3; ++a) {
        v.live[2 + a] = s->d_psi[a];
        v.count[2 + a] = mt::PSI_N * (size_t)s->desc.slab[a];
    }
    v.nseg = 5;
    return v;
}

// the segments from `first` on (1: H and the psi slabs, stored from offset
// off_h of a state; 2: the psi slabs, from off_psi[0])
static mt::StateSegs state_segs_from(const fwi3d_grad *g, int first)
{
    const mt::StateSegs all = state_segs(g);
    mt::StateSegs v{};
    for (int k = first; k < all.nseg; ++k) {
        v.live[k - first] = all.live[k];
        v.count[k - first] = all.count[k];
    }
    v.nseg = all.nseg - first;
    return v;
}

// copy the shot's live state to / from a stored state (one launch)
static int state_copy(fwi3d_grad *g, double *dst_state, bool save)
{
    HIP_TRY(mt::launch_state_copy(state_segs(g), dst_state, save, g->s->st));
    return FWI3D_OK;
}

static int set_step(fwi3d_shot *s, int n)
{
    HIP_TRY(hipMemcpyAsync(s->d_step, &n, sizeof(int), hipMemcpyHostToDevice, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    s->step = n;
    return FWI3D_OK;
}

// L forward steps from the shot's current step, storing the state after step
// t into buffer slot t + 1; captured once per segment length as a HIP graph.
// Slot 0 holds the state at the start.  The steps run on a copy of the shot's
// descriptor whose E ring is the buffer itself (slot stride one state, more
// slots than a leaf has steps): step t reads E of slot t and writes E of slot
// t + 1 in place of the shot's E, and each step's receiver gather rides on the
// next step's E launch as in a ring shot's chunk.  For the per-cell and LDS
// kernel families H and the psi slabs go the same way (rings over the states'
// H and psi parts, see hsel and ShotDev::pstride), so the steps write every
// stored state themselves and nothing is copied; the other families update H
// and psi in place and copy them into each slot.  The last slot is copied back
// to the shot afterwards (E, or the whole state with the rings), which leaves
// the shot's state at the segment's end as an in-place replay would.
static int replay_segment(fwi3d_grad *g, int L)
{
    fwi3d_shot *s = g->s;
    const int slot = L == g->leaf ? 0 : 1;
    if (g->replay[slot] && g->replay_len[slot] != L) {
        (void)hipGraphExecDestroy(g->replay[slot]);
        g->replay[slot] = nullptr;
    }
    const Fam f0 = shot_fam(s);
    const bool hring = (f0.kind == 0 || f0.kind == 3) && !f0.fused;
    g->rdesc = s->desc;
    g->rdesc.e = g->buf;
    g->rdesc.sstride = (long long)g->state;
    g->rdesc.ring = g->leaf + 1;
    g->rdesc.cur = 0;
    if (hring) {
        g->rdesc.h = g->buf + g->off_h;
        g->rdesc.hstride = (long long)g->state;
        for (int a = 0; a < 3; ++a)
            g->rdesc.psi[a] = g->buf + g->off_psi[a];
        g->rdesc.pstride = (long long)g->state;
    }
    HIP_TRY(hipMemcpyAsync(g->d_rdesc, &g->rdesc, sizeof(ShotDev), hipMemcpyHostToDevice, s->st));
    if (!g->replay[slot]) {
        Fam f = f0;
        f.rings = true;   // the stored states: an E ring of leaf + 1 slots
        // a replay step reads one stored state and writes the next: E and H
        // (and psi) both ways plus the coefficients, 112 B per cell (88 where H
        // is updated in place and copied)
        f.nt = stream_nt((long long)s->desc.n * (hring ? 112 : 88));
        const bool fold = s->nrec && rec_folds(f);
        const mt::StateSegs segs = state_segs_from(g, 1);   // H and psi, without the rings
        hipGraph_t graph = nullptr;
        std::lock_guard<std::recursive_mutex> lk(g_rt_mu);
        HIP_TRY(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
        int rc = FWI3D_OK;
        for (int j = 0; j < L && !rc; ++j) {
            if ((rc = launch_eh(f, g->d_rdesc, 1, j, false, s->st, fold && j > 0)) ||
                (rc = launch_eh(f, g->d_rdesc, 1, j, true, s->st)))
                break;
            if (s->nrec && !(fold && j + 1 < L) &&
                mt::launch_receivers(g->d_rdesc, 1, j, -1, 0, s->st) != hipSuccess)
                rc = fail(FWI3D_EDEVICE, "receiver launch");
            else if (!hring && mt::launch_state_copy(segs, g->buf + g->state * (j + 1) + g->off_h,
                                                     true, s->st) != hipSuccess)
                rc = fail(FWI3D_EDEVICE, "state copy launch");
        }
        if (!rc && mt::launch_advance(g->d_rdesc, 1, L, s->st) != hipSuccess)
            rc = fail(FWI3D_EDEVICE, "advance launch");
        hipError_t e = hipStreamEndCapture(s->st, &graph);
        if (rc)
            return rc;
        if (e != hipSuccess)
            return fail(FWI3D_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&g->replay[slot], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess)
            return fail(FWI3D_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
        g->replay_len[slot] = L;
    }
    HIP_TRY(hipGraphLaunch(g->replay[slot], s->st));
    mt::StateSegs back = state_segs(g);   // the whole state with the rings, else E
    if (!hring)
        back.nseg = 1;
    HIP_TRY(mt::launch_state_copy(back, g->buf + g->state * L, false, s->st));
    s->step += L;
    return FWI3D_OK;
}

extern "C" {

static size_t state_doubles(const fwi3d_shot *s)
{
    size_t o = 6 * (size_t)s->desc.n;
    for (int a = 0; a < 3; ++a)
        o += mt::PSI_N * (size_t)s->desc.slab[a];
    return o;
}

// ked = khd = dl on every axis (kappa = 1): the transposed stencils then
// multiply by 1/dl
static bool kappa_is_one(const fwi3d_shot *s)
{
    bool one = true;
    for (int a = 0; a < 3; ++a) {
        const int n = a == 0 ? s->desc.ix : (a == 1 ? s->desc.iy : s->desc.iz);
        for (int row : {FWI3D_P_KED, FWI3D_P_KHD})
            for (int q = 0; q < n; ++q)
                one &= s->m->prof[a][(size_t)row * n + q] == s->desc.dl[a];
    }
    return one;
}

// LDS region of the adjoint: [ncpml + 1, n - ncpml - 1) on each axis (x start
// rounded up to 8 cells), where the update is CPML-free and no transposed psi
// stencil (offsets -2..+2 from a CPML face cell) reaches; empty without
// kappa = 1 or when the LDS kernels are switched off
static bool adj_region(const fwi3d_shot *s, int lo[3], int hi[3])
{
    const ShotDev &d = s->desc;
    const int n[3] = {d.ix, d.iy, d.iz};
    bool ok = d.nc >= 2 && kappa_is_one(s) && !env_pad("MITTET_FDTD_NOLDS", 0);
    for (int a = 0; a < 3; ++a) {
        lo[a] = a == 0 ? (d.nc + 1 + 7) / 8 * 8 : d.nc + 1;
        hi[a] = n[a] - d.nc - 1;
        ok = ok && lo[a] < hi[a];
    }
    if (!ok)
        for (int a = 0; a < 3; ++a)
            lo[a] = hi[a] = 0;
    return ok;
}

// per-cell tiles of the adjoint: around its LDS region, or none (every tile)
static std::vector<int4> adj_tiles(const fwi3d_shot *s)
{
    int lo[3], hi[3];
    if (!adj_region(s, lo, hi))
        return {};
    return complement_tiles(s->desc.ix, s->desc.iy, s->desc.iz, lo, hi);
}

// device bytes a gradient with this plan allocates (states, scratch, gradient,
// trace sensitivities, profile reciprocals, per-cell tile list; the replay's
// copy of the shot descriptor, under 1 KB, is not counted)
static long long grad_plan_bytes(const fwi3d_shot *s, const GradPlan &P)
{
    const size_t N = (size_t)s->desc.n, st = state_doubles(s);
    const size_t tmp = st - 3 * N;
    const size_t other = tmp + N + 1 + std::max<size_t>(1, 3 * (size_t)s->nrec * s->it) +
                         2 * ((size_t)s->desc.ix + s->desc.iy + s->desc.iz) +
                         2 * adj_tiles(s).size();
    return 8LL * (long long)(P.states * state_pad(st) + other);
}

static int grad_new(fwi3d_shot *s, int levels, int segment, fwi3d_grad **out)
{
    if (!s || !out)
        return fail(FWI3D_EINVAL, "null argument");
    *out = nullptr;
    if (s->nf || s->ring != 1)
        return fail(FWI3D_EINVAL, "gradient shots must not carry a volume DFT");
    if (s->m->variant != 0)
        return fail(FWI3D_EINVAL, "the adjoint follows the C reference's step (variant 0)");
    if (s->batch)
        return fail(FWI3D_ESTATE, "shot is in a batch");
    if (levels < 1 || levels > 8)
        return fail(FWI3D_EINVAL, "levels must be in [1, 8] (got %d)", levels);
    const int nt = s->it - 1;
    fwi3d_grad *g = new fwi3d_grad();
    g->s = s;
    g->nt = nt;
    g->plan = grad_plan(nt, levels, segment);
    g->seg = g->plan.seglen[0];
    g->nseg = g->plan.nseg;
    g->leaf = g->plan.seglen[levels - 1];
    state_layout(g);
    const size_t N = (size_t)s->desc.n;
    size_t tmp = 3 * N;
    for (int a = 0; a < 3; ++a)
        tmp += mt::PSI_N * (size_t)s->desc.slab[a];
    bool ok = dalloc(&g->ckpt, g->state * g->nseg) == hipSuccess;
    for (int l = 1; l < levels && ok; ++l) {
        const size_t pieces =
            (size_t)((g->plan.seglen[l - 1] + g->plan.seglen[l] - 1) / g->plan.seglen[l]);
        double *p = nullptr;
        ok = dalloc(&p, g->state * pieces) == hipSuccess;
        g->inner.push_back(p);
    }
    if (!ok || dalloc(&g->buf, g->state * (g->leaf + 1)) != hipSuccess ||
        dalloc(&g->adj, g->state) != hipSuccess || dalloc(&g->tmp, tmp) != hipSuccess ||
        dalloc(&g->grad, N) != hipSuccess || dalloc(&g->grad_src, 1) != hipSuccess ||
        dalloc(&g->g, std::max<size_t>(1, 3 * (size_t)s->nrec * s->it)) != hipSuccess ||
        dalloc(&g->rp, 2 * ((size_t)s->desc.ix + s->desc.iy + s->desc.iz)) != hipSuccess ||
        dalloc(&g->d_rdesc, 1) != hipSuccess) {
        fwi3d_grad_destroy(g);
        return fail(FWI3D_ENOMEM, "gradient storage (%zu states of %zu doubles)",
                    g->plan.states, g->state);
    }
    // reciprocal ked / khd rows: the transposed stencils multiply where the
    // forward divides
    std::vector<double> rp;
    for (int a = 0; a < 3; ++a) {
        const int n = a == 0 ? s->desc.ix : (a == 1 ? s->desc.iy : s->desc.iz);
        for (int row : {FWI3D_P_KED, FWI3D_P_KHD})
            for (int q = 0; q < n; ++q)
                rp.push_back(1.0 / s->m->prof[a][(size_t)row * n + q]);
    }
    g->kappa_one = kappa_is_one(s);
    adj_region(s, g->lo, g->hi);
    const std::vector<int4> tl = adj_tiles(s);
    g->ntiles = (int)tl.size();
    if (g->ntiles && dalloc(&g->tiles, tl.size()) != hipSuccess) {
        fwi3d_grad_destroy(g);
        return fail(FWI3D_ENOMEM, "adjoint tile list");
    }
    if (hcopy(g->rp, rp.data(), sizeof(double) * rp.size()) != hipSuccess ||
        (g->ntiles && hcopy(g->tiles, tl.data(), sizeof(int4) * tl.size()) != hipSuccess)) {
        fwi3d_grad_destroy(g);
        return fail(FWI3D_EDEVICE, "profile upload failed");
    }
    *out = g;
    return FWI3D_OK;
}

int fwi3d_grad_create(fwi3d_shot *s, int segment, fwi3d_grad **out)
{
    return grad_new(s, 1, segment, out);
}

int fwi3d_grad_create_levels(fwi3d_shot *s, int levels, fwi3d_grad **out)
{
    return grad_new(s, levels, 0, out);
}

int fwi3d_grad_plan(int nt, int levels, int *seglen, int *nseg, long long *states)
{
    if (nt < 1 || levels < 1 || levels > 8 || !seglen || !nseg || !states)
        return fail(FWI3D_EINVAL, "bad gradient plan request");
    const GradPlan P = grad_plan(nt, levels, 0);
    for (int l = 0; l < levels; ++l)
        seglen[l] = P.seglen[l];
    *nseg = P.nseg;
    *states = (long long)P.states;
    return FWI3D_OK;
}

long long fwi3d_grad_plan_bytes(fwi3d_shot *s, int levels)
{
    if (!s || levels < 1 || levels > 8 || s->it < 2)
        return -1;
    return grad_plan_bytes(s, grad_plan(s->it - 1, levels, 0));
}

int fwi3d_grad_destroy(fwi3d_grad *g)
{
    if (!g)
        return FWI3D_OK;
    if (g->s)
        (void)hipStreamSynchronize(g->s->st);
    dfree(g->ckpt);
    for (double *p : g->inner)
        dfree(p);
    dfree(g->buf);
    dfree(g->adj);
    dfree(g->tmp);
    dfree(g->grad);
    dfree(g->grad_src);
    dfree(g->g);
    dfree(g->rp);
    dfree(g->tiles);
    dfree(g->d_rdesc);
    for (auto &r : g->replay)
        if (r)
            (void)hipGraphExecDestroy(r);
    delete g;
    return FWI3D_OK;
}

long long fwi3d_grad_bytes(fwi3d_grad *g)
{
    if (!g)
        return -1;
    return grad_plan_bytes(g->s, g->plan);
}

int fwi3d_grad_forward(fwi3d_grad *g)
{
    if (!g)
        return fail(FWI3D_EINVAL, "null gradient");
    fwi3d_shot *s = g->s;
    int rc = fwi3d_set_zero_eh(s);
    for (int q = 0; q < g->nseg && !rc; ++q) {
        if ((rc = state_copy(g, g->ckpt + g->state * q, true)))
            break;
        rc = fwi3d_step_n(s, std::min(g->seg, g->nt - q * g->seg));
    }
    if (!rc)
        rc = fwi3d_shot_sync(s);
    return rc;
}

// Adjoint steps over [n0, n0 + L) in reverse, at recursion depth `depth` of
// the checkpoint plan; the stored state `start` is the forward state at n0.
static int adj_range(fwi3d_grad *g, int depth, int n0, int L, double *start, mt::AdjArgs &A,
                     const std::vector<double> &amp)
{
    fwi3d_shot *s = g->s;
    const int nrec = s->nrec;
    int rc;
    if ((rc = state_copy(g, start, false)) || (rc = set_step(s, n0)))
        return rc;
    if (depth == g->plan.levels - 1) {
        // replay the segment keeping every state, then step the adjoint back
        if ((rc = state_copy(g, g->buf, true)) || (rc = replay_segment(g, L)))
            return rc;
        for (int t = L - 1; t >= 0; --t) {
            const int n = n0 + t;
            const double *u0 = g->buf + g->state * t, *u1 = g->buf + g->state * (t + 1);
            A.uE0 = u0;
            A.uH0 = u0 + g->off_h;
            A.uE1 = u1;
            for (int a = 0; a < 3; ++a)
                A.upsi1[a] = u1 + g->off_psi[a];
            A.amp_n = amp[n];
            HIP_TRY(mt::launch_adjoint_step(A, g->g + (size_t)n * 3 * nrec, nrec, s->d_rec, s->st));
        }
        return FWI3D_OK;
    }
    // one forward pass over the segment storing the start of each piece
    const int P = g->plan.seglen[depth + 1];
    const int np = (L + P - 1) / P;
    double *slots = g->inner[depth];
    for (int p = 0; p < np; ++p) {
        if ((rc = state_copy(g, slots + g->state * p, true)))
            return rc;
        if (p + 1 < np && (rc = fwi3d_step_n(s, P)))
            return rc;
    }
    for (int p = np - 1; p >= 0; --p)
        if ((rc = adj_range(g, depth + 1, n0 + p * P, std::min(P, L - p * P),
                            slots + g->state * p, A, amp)))
            return rc;
    return FWI3D_OK;
}

int fwi3d_grad_backward(fwi3d_grad *g, const double *djdt, double *djds)
{
    if (!g || !djdt || !djds)
        return fail(FWI3D_EINVAL, "null argument");
    fwi3d_shot *s = g->s;
    fwi3d_model *m = s->m;
    int rc = check_shot_ready(s);
    if (rc)
        return rc;
    const int nrec = s->nrec, it = s->it;
    const size_t N = (size_t)s->desc.n;
    // sensitivities regrouped per step: g[n][comp][rec]
    std::vector<double> gt(std::max<size_t>(1, 3 * (size_t)nrec * it));
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < nrec; ++r)
            for (int n = 0; n < it; ++n)
                gt[((size_t)n * 3 + c) * nrec + r] = djdt[((size_t)c * nrec + r) * it + n];
    HIP_TRY(hipMemcpyAsync(g->g, gt.data(), sizeof(double) * gt.size(), hipMemcpyHostToDevice, s->st));
    HIP_TRY(mt::launch_fill(g->adj, g->state, 0.0, s->st));
    HIP_TRY(mt::launch_fill(g->grad, N, 0.0, s->st));
    HIP_TRY(mt::launch_fill(g->grad_src, 1, 0.0, s->st));
    std::vector<double> amp(it, 0.0);
    const double sig_src = s->src_cell >= 0 ? m->sig[(size_t)s->src_cell] : 1.0;
    if (s->src_cell >= 0) {
        const fwi3d_grid_params &p = m->p;
        for (int n = 0; n + 1 < it; ++n)
            amp[n] = (s->signal[n] + s->signal[n + 1]) / 2.0 * p.dt * 2.0 * p.omega0 / sig_src / p.dx /
                     p.dy / p.dz;
    }
    mt::AdjArgs A{};
    const ShotDev &d = s->desc;
    A.ix = d.ix;
    A.iy = d.iy;
    A.iz = d.iz;
    A.nc = d.nc;
    A.n = d.n;
    A.dt = d.dt;
    A.mu0 = d.mu0;
    A.c1 = d.c1;
    A.c2 = d.c2;
    A.ftx = tiles(d.ix, mt::ZX);
    A.fty = tiles(d.iy, mt::ZY);
    A.fb = A.ftx * A.fty * d.iz;
    A.fb_pad = tiles(A.fb, 8) * 8;
    A.shell = nullptr;
    for (int a = 0; a < 3; ++a) {
        A.lo[a] = g->lo[a];
        A.hi[a] = g->hi[a];
    }
    A.ltx = A.lty = A.lkz = A.lb = A.lb_pad = A.lkb = 0;
    if (g->ntiles) {
        A.shell = g->tiles;
        A.fb = A.fb_pad = g->ntiles;
        A.ltx = tiles(A.hi[0] - A.lo[0], mt::LX);
        A.lty = tiles(A.hi[1] - A.lo[1], mt::LY);
        A.lkb = lds_chunk(A.ltx, A.lty, A.hi[2] - A.lo[2]);
        A.lkz = tiles(A.hi[2] - A.lo[2], A.lkb);
        A.lb = A.ltx * A.lty * A.lkz;
        A.lb_pad = tiles(A.lb, 8) * 8;
    }
    // [nc, n - nc) is free of CPML terms when nc >= 2 (the update ranges start
    // at index 1 or 2); with a thinner CPML no cell takes the fast paths
    for (int a = 0; a < 3; ++a) {
        const int n = a == 0 ? d.ix : (a == 1 ? d.iy : d.iz);
        A.inner_hi[a] = d.nc >= 2 ? n - d.nc : 0;
    }
    A.kappa_one = g->kappa_one;
    const double *rp = g->rp;
    for (int a = 0; a < 3; ++a) {
        const int n = a == 0 ? d.ix : (a == 1 ? d.iy : d.iz);
        A.dl[a] = d.dl[a];
        A.rdl[a] = 1.0 / d.dl[a];
        A.rp[a] = rp;
        rp += 2 * n;
        A.prof[a] = d.prof[a];
        A.slab[a] = d.slab[a];
        A.lpsi[a] = g->adj + g->off_psi[a];
    }
    A.eps = d.eps;
    A.cbi = d.cbi;
    A.lE = g->adj;
    A.lH = g->adj + g->off_h;
    A.t3 = g->tmp;
    size_t o = 3 * N;
    for (int a = 0; a < 3; ++a) {
        A.tpsi[a] = g->tmp + o;
        o += mt::PSI_N * (size_t)d.slab[a];
    }
    A.grad = g->grad;
    A.grad_src = g->grad_src;
    A.src_cell = s->src_cell;
    A.src_comp = s->src_comp;
    A.sig_src = sig_src;
    for (int q = g->nseg - 1; q >= 0; --q) {
        const int n0 = q * g->seg, L = std::min(g->seg, g->nt - n0);
        if ((rc = adj_range(g, 0, n0, L, g->ckpt + g->state * q, A, amp)))
            return rc;
    }
    std::vector<double> ge(N);
    double gs = 0.0;
    HIP_TRY(hipMemcpyAsync(ge.data(), g->grad, sizeof(double) * N, hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipMemcpyAsync(&gs, g->grad_src, sizeof(double), hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    const double de = 1.0 / (2.0 * m->p.omega0);
    for (size_t c = 0; c < N; ++c)
        djds[c] = ge[c] * de;
    if (s->src_cell >= 0)
        djds[s->src_cell] += gs;
    return FWI3D_OK;
}

}  // extern "C"


// ---------------------------------------------------------------- reference-style FWI

struct fwi3d_compat {
    fwi3d_model *m = nullptr;
    hipStream_t st = nullptr;
    int nsrc = 0, nrec = 0, f0 = 0, f1 = 0, nb = 0, it = 0;
    std::vector<int> rec;                          // receiver cells
    std::vector<mt::cplx> jw, tw;                  // J and T at the bins f0..f1-1
    std::vector<std::vector<mt::cplx>> de, effl;   // per source [3][nrec][f1]: residual, E_ff at receivers
    double scaler = 0.0;
    int *d_rec = nullptr;
    double *d_jw = nullptr, *d_tw = nullptr, *d_eff = nullptr, *d_eback = nullptr, *d_G = nullptr;
    double *d_sense = nullptr, *d_grad = nullptr, *d_dconj = nullptr, *d_gather = nullptr;
    double *d_work = nullptr, *d_blockmax = nullptr;
};

static constexpr int kPhiBlocks = 1024;

static size_t compat_vol(const fwi3d_compat *c)
{
    const fwi3d_grid_params &p = c->m->p;
    return (size_t)3 * c->nb * p.ix * p.iy * p.iz;   // complex values per spectra volume
}

// receiver E_ff (laplaceToFreq_2 / _3 at the receiver cells) of a shot whose
// volume DFT ran at the object's bins: [3][nrec][f1] complex, 0 below f0
// the shot's grid is the object's and its volume DFT ran at the object's bins
static int compat_shot_ok(const fwi3d_compat *c, const fwi3d_shot *s)
{
    const fwi3d_grid_params &p = c->m->p, &q = s->m->p;
    bool ok = p.ix == q.ix && p.iy == q.iy && p.iz == q.iz && s->nf == c->nb && s->it == c->it;
    const double om = 2.0 * M_PI / c->it / p.dt;
    for (int b = 0; ok && b < c->nb; ++b)
        ok = s->omega[b] == om * (c->f0 + b);
    return ok ? FWI3D_OK
              : fail(FWI3D_EINVAL, "shot must share the object's grid, it and its %d DFT bins "
                                   "(fwi3d_compat_omegas)", c->nb);
}

static int compat_receivers(fwi3d_compat *c, fwi3d_shot *s, std::vector<mt::cplx> &out)
{
    if (!s)
        return fail(FWI3D_EINVAL, "null shot");
    int rc = compat_shot_ok(c, s);
    if (rc)
        return rc;
    const int nrec = c->nrec, nb = c->nb, f1 = c->f1;
    std::vector<mt::cplx> raw((size_t)3 * nrec * nb);
    HIP_TRY(mt::launch_compat_gather(s->d_acc, s->desc.astride, nb, nrec, c->d_rec, c->d_gather, s->st));
    HIP_TRY(hipMemcpyAsync(raw.data(), c->d_gather, sizeof(mt::cplx) * raw.size(), hipMemcpyDeviceToHost,
                           s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    out.assign((size_t)3 * nrec * f1, mt::cplx{0.0, 0.0});
    for (int d = 0; d < 3; ++d)
        for (int r = 0; r < nrec; ++r)
            for (int b = 0; b < nb; ++b)
                out[((size_t)d * nrec + r) * f1 + c->f0 + b] =
                    mt::cmul(mt::cdiv(raw[((size_t)d * nrec + r) * nb + b], c->jw[b]), c->tw[b]);
    return FWI3D_OK;
}

extern "C" {

int fwi3d_compat_create(fwi3d_model *m, int nsrc, int nrec, const int *rec_ijk, int start_f,
                        int upper_f, int it, const double *signal_x, const double *signal_true,
                        fwi3d_compat **out)
{
    if (!m || !out || !rec_ijk || !signal_x || !signal_true || nsrc <= 0 || nrec <= 0 ||
        start_f < 1 || upper_f <= start_f || it < 2 || upper_f > it)
        return fail(FWI3D_EINVAL, "bad compat request (need nsrc, nrec > 0 and 1 <= start_f < upper_f <= it)");
    *out = nullptr;
    const fwi3d_grid_params &p = m->p;
    auto *c = new fwi3d_compat();
    c->m = m;
    c->st = m->stream;
    c->nsrc = nsrc;
    c->nrec = nrec;
    c->f0 = start_f;
    c->f1 = upper_f;
    c->nb = upper_f - start_f;
    c->it = it;
    for (int r = 0; r < nrec; ++r) {
        const int i = rec_ijk[3 * r], j = rec_ijk[3 * r + 1], k = rec_ijk[3 * r + 2];
        if (i < 0 || j < 0 || k < 0 || i >= p.ix || j >= p.iy || k >= p.iz) {
            delete c;
            return fail(FWI3D_EINVAL, "receiver %d outside the grid", r);
        }
        c->rec.push_back((k * p.iy + j) * p.ix + i);
    }
    // J at the bins (fictitious transform of signal_x with sqrt(-2 omega0/(i omega))) and T
    const double om = 2.0 * M_PI / it / p.dt;
    std::vector<double> omg(c->nb), jv(2 * (size_t)c->nb);
    for (int b = 0; b < c->nb; ++b)
        omg[b] = om * (start_f + b);
    int rc = fwi3d_source_spectrum(p.dt, p.omega0, it, signal_x, c->nb, omg.data(), jv.data());
    if (rc) {
        delete c;
        return rc;
    }
    const auto tru = true_spectrum(signal_true, it, start_f, upper_f);
    for (int b = 0; b < c->nb; ++b) {
        c->jw.push_back({jv[2 * b], jv[2 * b + 1]});
        c->tw.push_back({tru[b].real(), tru[b].imag()});
    }
    c->de.assign(nsrc, {});
    c->effl.assign(nsrc, {});
    const size_t N = (size_t)p.ix * p.iy * p.iz, vol = compat_vol(c);
    if (dalloc(&c->d_rec, (size_t)nrec) != hipSuccess || dalloc(&c->d_jw, 2 * (size_t)c->nb) != hipSuccess ||
        dalloc(&c->d_tw, 2 * (size_t)c->nb) != hipSuccess ||
        dalloc(&c->d_eff, 2 * vol * nsrc) != hipSuccess || dalloc(&c->d_eback, 2 * vol * nsrc) != hipSuccess ||
        dalloc(&c->d_G, 2 * vol) != hipSuccess || dalloc(&c->d_sense, N * nsrc) != hipSuccess ||
        dalloc(&c->d_grad, N) != hipSuccess || dalloc(&c->d_dconj, 2 * (size_t)nsrc * c->nb) != hipSuccess ||
        dalloc(&c->d_gather, 2 * (size_t)3 * nrec * c->nb) != hipSuccess || dalloc(&c->d_work, N) != hipSuccess ||
        dalloc(&c->d_blockmax, 2 * (size_t)kPhiBlocks) != hipSuccess) {
        fwi3d_compat_destroy(c);
        return fail(FWI3D_ENOMEM, "compat buffers (%zu complex spectra per source)", vol);
    }
    if (hcopy(c->d_rec, c->rec.data(), sizeof(int) * nrec) != hipSuccess ||
        hcopy(c->d_jw, c->jw.data(), sizeof(mt::cplx) * c->nb) != hipSuccess ||
        hcopy(c->d_tw, c->tw.data(), sizeof(mt::cplx) * c->nb) != hipSuccess) {
        fwi3d_compat_destroy(c);
        return fail(FWI3D_EDEVICE, "compat upload failed");
    }
    if ((rc = fwi3d_compat_reset(c))) {
        fwi3d_compat_destroy(c);
        return rc;
    }
    *out = c;
    return FWI3D_OK;
}

int fwi3d_compat_destroy(fwi3d_compat *c)
{
    if (!c)
        return FWI3D_OK;
    (void)hipStreamSynchronize(c->st);
    dfree(c->d_rec);
    dfree(c->d_jw);
    dfree(c->d_tw);
    dfree(c->d_eff);
    dfree(c->d_eback);
    dfree(c->d_G);
    dfree(c->d_sense);
    dfree(c->d_grad);
    dfree(c->d_dconj);
    dfree(c->d_gather);
    dfree(c->d_work);
    dfree(c->d_blockmax);
    delete c;
    return FWI3D_OK;
}

int fwi3d_compat_omegas(const fwi3d_compat *c, double *out)
{
    if (!c || !out)
        return fail(FWI3D_EINVAL, "null argument");
    const double om = 2.0 * M_PI / c->it / c->m->p.dt;
    for (int b = 0; b < c->nb; ++b)
        out[b] = om * (c->f0 + b);
    return FWI3D_OK;
}

int fwi3d_compat_reset(fwi3d_compat *c)
{
    if (!c)
        return fail(FWI3D_EINVAL, "null compat");
    const size_t N = (size_t)c->m->p.ix * c->m->p.iy * c->m->p.iz;
    HIP_TRY(mt::launch_fill(c->d_grad, N, 0.0, c->st));
    HIP_TRY(mt::launch_fill(c->d_eback, 2 * compat_vol(c) * c->nsrc, 0.0, c->st));
    HIP_TRY(mt::launch_fill(c->d_sense, N * c->nsrc, 0.0, c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    return FWI3D_OK;
}

int fwi3d_compat_source(fwi3d_compat *c, int isrc, fwi3d_shot *s, const double *eobs, double *de,
                        double *err_sum)
{
    if (!c || !s || !eobs || isrc < 0 || isrc >= c->nsrc)
        return fail(FWI3D_EINVAL, "bad compat source request");
    int rc = check_shot_ready(s);
    if (rc || (rc = compat_receivers(c, s, c->effl[isrc])))
        return rc;
    const size_t N = (size_t)s->desc.n;
    HIP_TRY(mt::launch_compat_eff(s->d_acc, s->desc.astride, (int)N, c->nb, c->d_jw, c->d_tw,
                                  c->d_eff + 2 * compat_vol(c) * isrc, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    // residualE: dE = Eobs - E_ff at the receivers over bins [0, upperF);
    // err += |dEx| + |dEy| + |dEz| (receiver, then bin)
    const int nrec = c->nrec, f1 = c->f1;
    std::vector<mt::cplx> &d = c->de[isrc];
    d.assign((size_t)3 * nrec * f1, mt::cplx{0.0, 0.0});
    const mt::cplx *ob = reinterpret_cast<const mt::cplx *>(eobs);
    for (size_t q = 0; q < d.size(); ++q)
        d[q] = {ob[q].re - c->effl[isrc][q].re, ob[q].im - c->effl[isrc][q].im};
    double err = err_sum ? *err_sum : 0.0;
    for (int r = 0; r < nrec; ++r)
        for (int w = 0; w < f1; ++w)
            for (int comp = 0; comp < 3; ++comp) {
                const mt::cplx v = d[((size_t)comp * nrec + r) * f1 + w];
                err += std::hypot(v.re, v.im);
            }
    if (err_sum)
        *err_sum = err;
    if (de)
        std::memcpy(de, d.data(), sizeof(mt::cplx) * d.size());
    return FWI3D_OK;
}

int fwi3d_compat_green(fwi3d_compat *c, int irec, int idir, fwi3d_shot *g)
{
    if (!c || !g || irec < 0 || irec >= c->nrec || idir < 0 || idir > 2)
        return fail(FWI3D_EINVAL, "bad compat Green's function request");
    int rc = compat_shot_ok(c, g);
    if (rc || (rc = check_shot_ready(g)))
        return rc;
    for (int s = 0; s < c->nsrc; ++s)
        if (c->de[s].empty())
            return fail(FWI3D_ESTATE, "source %d has no residual yet", s);
    std::vector<mt::cplx> dc((size_t)c->nsrc * c->nb);
    for (int s = 0; s < c->nsrc; ++s)
        for (int b = 0; b < c->nb; ++b) {
            const mt::cplx v = c->de[s][((size_t)idir * c->nrec + irec) * c->f1 + c->f0 + b];
            dc[(size_t)s * c->nb + b] = {v.re, -v.im};
        }
    HIP_TRY(hipMemcpyAsync(c->d_dconj, dc.data(), sizeof(mt::cplx) * dc.size(), hipMemcpyHostToDevice, g->st));
    HIP_TRY(mt::launch_compat_green(g->d_acc, g->desc.astride, g->desc.n, c->nb, c->d_jw, c->nsrc, c->d_eff,
                                    c->d_dconj, c->d_G, c->d_sense, c->d_eback, g->st));
    HIP_TRY(hipStreamSynchronize(g->st));
    return FWI3D_OK;
}

int fwi3d_compat_gamma(fwi3d_compat *c, double *grad)
{
    if (!c)
        return fail(FWI3D_EINVAL, "null compat");
    const fwi3d_grid_params &p = c->m->p;
    HIP_TRY(mt::launch_compat_gamma(p.ix, p.iy, p.iz, p.ncpml, c->nb, c->f0, c->nsrc, c->d_eff, c->d_eback,
                                    c->d_sense, c->d_grad, c->st));
    if (grad)
        HIP_TRY(hipMemcpyAsync(grad, c->d_grad, sizeof(double) * p.ix * p.iy * p.iz, hipMemcpyDeviceToHost,
                               c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    return FWI3D_OK;
}

int fwi3d_compat_receivers(fwi3d_compat *c, fwi3d_shot *s, double *eff)
{
    if (!c || !s || !eff)
        return fail(FWI3D_EINVAL, "null argument");
    int rc = check_shot_ready(s);
    std::vector<mt::cplx> v;
    if (rc || (rc = compat_receivers(c, s, v)))
        return rc;
    std::memcpy(eff, v.data(), sizeof(mt::cplx) * v.size());
    return FWI3D_OK;
}

int fwi3d_compat_phi(fwi3d_compat *c, int iseabed, double *scaler, double *sig2)
{
    if (!c)
        return fail(FWI3D_EINVAL, "null compat");
    fwi3d_model *m = c->m;
    const fwi3d_grid_params &p = m->p;
    const int N = p.ix * p.iy * p.iz;
    std::vector<double> bm(2 * (size_t)kPhiBlocks);
    double g0 = 0.0;
    HIP_TRY(mt::launch_phi_max(m->d_sig, c->d_grad, N, p.ix * p.iy, iseabed, c->d_blockmax, kPhiBlocks, c->st));
    HIP_TRY(hipMemcpyAsync(bm.data(), c->d_blockmax, sizeof(double) * bm.size(), hipMemcpyDeviceToHost, c->st));
    HIP_TRY(hipMemcpyAsync(&g0, c->d_grad, sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    // the reference starts both maxima from cell 0
    double sigmax = m->sig[0], gradmax = g0;
    for (int b = 0; b < kPhiBlocks; ++b) {
        sigmax = bm[2 * b] > sigmax ? bm[2 * b] : sigmax;
        gradmax = bm[2 * b + 1] > gradmax ? bm[2 * b + 1] : gradmax;
    }
    c->scaler = (double)0.01f * sigmax / gradmax;
    HIP_TRY(mt::launch_phi_apply(m->d_sig, c->d_grad, c->scaler, N, c->d_work, c->st));
    if (sig2)
        HIP_TRY(hipMemcpyAsync(sig2, c->d_work, sizeof(double) * N, hipMemcpyDeviceToHost, c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    if (scaler)
        *scaler = c->scaler;
    return FWI3D_OK;
}

int fwi3d_compat_alpha(fwi3d_compat *c, int isrc, const double *eff2, double *k1, double *k2)
{
    if (!c || !eff2 || !k1 || !k2 || isrc < 0 || isrc >= c->nsrc || c->de[isrc].empty())
        return fail(FWI3D_EINVAL, "bad compat alpha request");
    const mt::cplx *e2 = reinterpret_cast<const mt::cplx *>(eff2);
    mt::cplx a = {k1[0], k1[1]}, b = {k2[0], k2[1]};
    const int nrec = c->nrec, f1 = c->f1;
    for (int r = 0; r < nrec; ++r)
        for (int w = 0; w < f1; ++w)
            for (int comp = 0; comp < 3; ++comp) {
                const size_t q = ((size_t)comp * nrec + r) * f1 + w;
                const mt::cplx t = {e2[q].re - c->effl[isrc][q].re, e2[q].im - c->effl[isrc][q].im};
                const mt::cplx u = mt::cmul(c->de[isrc][q], t), v = mt::cmul(t, t);
                a = {a.re + u.re, a.im + u.im};
                b = {b.re + v.re, b.im + v.im};
            }
    k1[0] = a.re;
    k1[1] = a.im;
    k2[0] = b.re;
    k2[1] = b.im;
    return FWI3D_OK;
}

int fwi3d_compat_update(fwi3d_compat *c, const double *k1, const double *k2, double minval, double maxval,
                        int iseabed, double *alpha, double *sig)
{
    if (!c || !k1 || !k2)
        return fail(FWI3D_EINVAL, "null argument");
    fwi3d_model *m = c->m;
    const fwi3d_grid_params &p = m->p;
    const size_t N = (size_t)p.ix * p.iy * p.iz;
    const double a = (double)3.f * std::hypot(k1[0], k1[1]) / std::hypot(k2[0], k2[1]);
    HIP_TRY(mt::launch_copy(m->d_sig, c->d_work, N, c->st));
    HIP_TRY(mt::launch_update(c->d_work, c->d_grad, a * c->scaler, minval, maxval, p.ix, p.iy, p.iz, p.ncpml,
                              iseabed, c->st));
    std::vector<double> ns(N);
    HIP_TRY(hipMemcpyAsync(ns.data(), c->d_work, sizeof(double) * N, hipMemcpyDeviceToHost, c->st));
    HIP_TRY(mt::launch_fill(c->d_grad, N, 0.0, c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    int rc = fwi3d_media_coeff_3d(m, ns.data());
    if (rc)
        return rc;
    if (alpha)
        *alpha = a;
    if (sig)
        std::memcpy(sig, ns.data(), sizeof(double) * N);
    return FWI3D_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- batches

int fwi3d_batch_create(int max_shots, int chunk, void *stream, fwi3d_batch **out)
{
    if (!out || max_shots <= 0)
        return fail(FWI3D_EINVAL, "bad batch request");
    if (chunk <= 0)
        chunk = kDefaultChunk;
    if (chunk > mt::RING_MAX)
        return fail(FWI3D_EINVAL, "chunk %d exceeds %d", chunk, mt::RING_MAX);
    fwi3d_batch *b = new fwi3d_batch();
    b->max_shots = max_shots;
    b->chunk = chunk;
    if (stream) {
        b->st = (hipStream_t)stream;
    } else if (new_stream(&b->st) != hipSuccess) {
        delete b;
        return fail(FWI3D_EDEVICE, "hipStreamCreate failed");
    } else {
        b->own_stream = true;
    }
    if (dalloc(&b->d_desc, (size_t)max_shots) != hipSuccess ||
        dalloc(&b->d_pre, (size_t)max_shots + 1) != hipSuccess ||
        dalloc(&b->d_pre_l, (size_t)max_shots + 1) != hipSuccess ||
        dalloc(&b->d_pre_cells, (size_t)max_shots + 1) != hipSuccess) {
        fwi3d_batch_destroy(b);
        return fail(FWI3D_ENOMEM, "batch descriptors");
    }
    *out = b;
    return FWI3D_OK;
}

int fwi3d_batch_destroy(fwi3d_batch *b)
{
    if (!b)
        return FWI3D_OK;
    if (b->st)
        (void)hipStreamSynchronize(b->st);
    for (auto *s : b->shots)
        s->batch = nullptr;
    if (b->graph)
        (void)hipGraphExecDestroy(b->graph);
    dfree(b->d_desc);
    dfree(b->d_pre);
    dfree(b->d_pre_l);
    dfree(b->d_pre_cells);
    if (b->own_stream && b->st)
        del_stream(b->st);
    delete b;
    return FWI3D_OK;
}

int fwi3d_batch_add(fwi3d_batch *b, fwi3d_shot *s)
{
    if (!b || !s)
        return fail(FWI3D_EINVAL, "null argument");
    if (s->batch)
        return fail(FWI3D_ESTATE, "shot already in a batch");
    if ((int)b->shots.size() >= b->max_shots)
        return fail(FWI3D_EBUDGET, "batch holds at most %d shots", b->max_shots);
    if (s->ring != 1 && s->ring != b->chunk)
        return fail(FWI3D_EINVAL, "shot ring %d must equal the batch chunk %d", s->ring, b->chunk);
    // families 0 and 3 share launches (a family-0 member has no LDS blocks and
    // no shell list); pairs and z-marching need members of their own family
    const bool mix03 = (s->kind == 0 || s->kind == 3) && (b->kind == 0 || b->kind == 3);
    if (!b->shots.empty() && s->kind != b->kind && !mix03)
        return fail(FWI3D_EINVAL, "batch members must use the same fused kernel family "
                                  "(x-pairs, zchunk)");
    b->kind = b->shots.empty() ? s->kind : std::max(b->kind, s->kind);
    int rc = ensure_amp(s);
    if (rc)
        return rc;
    // the shot's own stream must be idle before the batch stream touches it
    HIP_TRY(hipStreamSynchronize(s->st));
    s->batch = b;
    b->shots.push_back(s);
    b->dirty = true;
    return FWI3D_OK;
}

int fwi3d_batch_remove(fwi3d_batch *b, fwi3d_shot *s)
{
    if (!b || !s || s->batch != b)
        return fail(FWI3D_EINVAL, "shot is not in this batch");
    HIP_TRY(hipStreamSynchronize(b->st));
    b->shots.erase(std::find(b->shots.begin(), b->shots.end(), s));
    s->batch = nullptr;
    s->desc_dirty = true;
    b->dirty = true;
    return FWI3D_OK;
}

int fwi3d_batch_size(const fwi3d_batch *b) { return b ? (int)b->shots.size() : -1; }

static int batch_upload(fwi3d_batch *b)
{
    const int ns = (int)b->shots.size();
    std::vector<ShotDev> descs(ns);
    std::vector<int> pre(ns + 1, 0), prel(ns + 1, 0), prec(ns + 1, 0);
    for (int q = 0; q < ns; ++q) {
        descs[q] = b->shots[q]->desc;
        pre[q + 1] = pre[q] + b->shots[q]->nblocks;
        prel[q + 1] = prel[q] + b->shots[q]->nblocks_l;
        prec[q + 1] = prec[q] + b->shots[q]->ncellblocks;
    }
    b->nblocks = pre[ns];
    b->nblocks_l = prel[ns];
    b->ncellblocks = prec[ns];
    if (ns) {
        HIP_TRY(hipMemcpyAsync(b->d_desc, descs.data(), sizeof(ShotDev) * ns, hipMemcpyHostToDevice,
                               b->st));
        HIP_TRY(hipMemcpyAsync(b->d_pre, pre.data(), sizeof(int) * (ns + 1), hipMemcpyHostToDevice,
                               b->st));
        HIP_TRY(hipMemcpyAsync(b->d_pre_l, prel.data(), sizeof(int) * (ns + 1),
                               hipMemcpyHostToDevice, b->st));
        HIP_TRY(hipMemcpyAsync(b->d_pre_cells, prec.data(), sizeof(int) * (ns + 1),
                               hipMemcpyHostToDevice, b->st));
        HIP_TRY(hipStreamSynchronize(b->st));
    }
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    b->dirty = false;
    return FWI3D_OK;
}

static Fam batch_fam(const fwi3d_batch *b)
{
    bool rings = true, fused = b->kind == 3;
    long long touch = 0;
    for (const auto *s : b->shots) {
        rings = rings && s->ring > 1;
        fused = fused && s->kind == 3 && s->fused;
        touch += step_touch_bytes(s);
    }
    return {b->kind, b->d_pre, b->nblocks, b->d_pre_l, b->nblocks_l, rings, fused, stream_nt(touch)};
}

int fwi3d_batch_step(fwi3d_batch *b, int nsteps)
{
    if (!b || nsteps < 0)
        return fail(FWI3D_EINVAL, "bad batch step");
    int rc;
    if (b->dirty && (rc = batch_upload(b)))
        return rc;
    const int ns = (int)b->shots.size();
    if (ns == 0)
        return FWI3D_OK;
    int left = 0;
    bool dft = false, rec = false, hdft = false;
    for (auto *s : b->shots) {
        left = std::max(left, std::min(nsteps, s->it - 1 - s->step));
        dft = dft || s->nf > 0;
        rec = rec || s->nrec > 0;
        hdft = hdft || s->d_acc_h != nullptr;
    }
    while (left > 0) {
        const int L = std::min(b->chunk, left);
        if (L == b->chunk) {
            if (!b->graph && (rc = capture_chunk(&b->graph, b->d_desc, ns, batch_fam(b),
                                                 b->d_pre_cells, b->ncellblocks, L, dft, rec,
                                                 b->st, hdft)))
                return rc;
            HIP_TRY(hipGraphLaunch(b->graph, b->st));
        } else if ((rc = launch_chunk(b->d_desc, ns, batch_fam(b), b->d_pre_cells,
                                      b->ncellblocks, L, dft, rec, b->st, hdft))) {
            return rc;
        }
        left -= L;
        bool moved = false;
        const bool fused = batch_fam(b).fused;
        for (auto *s : b->shots) {
            const int done = std::min(L, s->it - 1 - s->step);
            s->step += done;
            moved = advance_desc(s, done, fused) || moved;
            s->desc_dirty = true;
        }
        if (moved && (rc = batch_upload(b)))
            return rc;
    }
    return FWI3D_OK;
}

int fwi3d_batch_sync(fwi3d_batch *b)
{
    if (!b)
        return fail(FWI3D_EINVAL, "null batch");
    HIP_TRY(hipStreamSynchronize(b->st));
    return FWI3D_OK;
}

}  // extern "C"
