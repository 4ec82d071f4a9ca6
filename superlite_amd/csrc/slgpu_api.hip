// C ABI of libslgpu.so (include/slgpu.h): context, uploads, per-iteration entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "slgpu_device.h"
#include "slgpu_internal.h"
#include "slgpu_sample.h"

using namespace slgpu;

// Device memory of slgpu_opacity_build_noalloc: the staging of the per-call ion populations
// (oc and ginv, 2 * nion * nc doubles), so that entry point needs no allocation.  Each GPU
// holds its own copy until the process ends; every call on a GPU shares it (calls that
// overlap on one GPU race).  Host code never uses the array's name as an address: its
// device address is resolved once per device at slgpu_create (hipGetSymbolAddress) and
// kept in g_scratch_ptr.
constexpr size_t SLGPU_SCRATCH_BYTES = (size_t)4 << 20;
static __device__ __align__(256) unsigned char scratch[SLGPU_SCRATCH_BYTES];

namespace {

thread_local std::string g_err;

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// device address of `scratch` on the current device, resolved on first use per device
hipError_t scratch_address(void **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(SLGPU_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
int dalloc(slgpu_ctx *c, T **p, size_t n)
{
    void *q = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&q, n * sizeof(T));
    if (e != hipSuccess) return fail(SLGPU_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    (void)hipMemsetAsync(q, 0, n * sizeof(T), c->stream);
    c->allocations.push_back(q);
    *p = (T *)q;
    return 0;
}

void dfree(slgpu_ctx *c, void *p)
{
    if (!p) return;
    auto it = std::find(c->allocations.begin(), c->allocations.end(), p);
    if (it != c->allocations.end()) c->allocations.erase(it);
    hipFree(p);
}

struct PhaseTimer {
    slgpu_ctx *c;
    int slot;
    PhaseTimer(slgpu_ctx *c_, int s) : c(c_), slot(s) { hipEventRecord(c->ev[2 * s], c->stream); }
    int stop()
    {
        hipEventRecord(c->ev[2 * slot + 1], c->stream);
        hipError_t e = hipEventSynchronize(c->ev[2 * slot + 1]);
        if (e != hipSuccess) return fail(SLGPU_EHIP, std::string("device: ") + hipGetErrorString(e));
        float ms = 0.f;
        hipEventElapsedTime(&ms, c->ev[2 * slot], c->ev[2 * slot + 1]);
        c->t_ms[slot] += ms;
        return 0;
    }
};

// SOURCE/sourcenumbers.f90:83-145 with one rank, in_srcepwr = 1, tot_esurf = 0
int64_t source_numbers(const std::vector<double> &emit, int64_t nstot, std::vector<int> &nvol)
{
    const int nc = (int)emit.size();
    double etot = 0.0;
    int64_t ncactive = 0;
    for (double e : emit) {
        etot += std::pow(e, 1.0);
        if (e > 0.0) ncactive++;
    }
    nvol.assign(nc, 0);
    if (ncactive == 0) return -1;
    const int64_t nsavail = nstot;
    const double invn = 1.0 / (double)nsavail;
    const double einv = etot == 0.0 ? 0.0 : 1.0 / etot;
    double edone = 0.0;
    int64_t ndone = 0;
    for (int l = 0; l < nc; ++l) {
        const double en = std::pow(emit[l], 1.0);
        if (en == 0.0) continue;
        int64_t n = std::max<int64_t>(1, (int64_t)(en * (double)nsavail * einv));
        if (edone * einv > (double)ndone * invn) n = n + 1;
        nvol[l] = (int)n;
        edone += en;
        ndone += n;
    }
    int64_t s = 0;
    for (int v : nvol) s += v;
    int64_t nextra = nstot - s;
    if (nextra > nc) return -2;       // sourcenumbers.f90:130-131
    const int64_t nsmean = nstot / ncactive;
    while (nextra != 0) {
        for (int l = 0; l < nc; ++l) {
            if (nextra == 0) break;
            if (nvol[l] < nsmean) continue;
            const int sg = nextra > 0 ? 1 : -1;
            nvol[l] += sg;
            nextra -= sg;
        }
    }
    s = 0;
    for (int v : nvol) s += v;
    return s;
}

}  // namespace

extern "C" {

const char *slgpu_last_error_message(void) { return g_err.c_str(); }

int slgpu_create(const slgpu_config *cfg, slgpu_ctx **out)
{
    if (!cfg || !out) return fail(SLGPU_EINVAL, "null argument");
    if (cfg->nx < 1 || cfg->nx > 8192) return fail(SLGPU_EINVAL, "nx must be in 1..8192");
    if (cfg->ng < 2 || cfg->ng > 32767) return fail(SLGPU_EINVAL, "ng must be in 2..32767 (groupmod.f:38)");
    if (cfg->nep < 1 || cfg->nep > 1024 || cfg->nepg < 1)
        return fail(SLGPU_EINVAL, "nep must be in 1..1024 and nepg >= 1");
    if (cfg->npartmax < 1 || cfg->npartmax > (int64_t)1 << 31)
        return fail(SLGPU_EINVAL, "npartmax must be in 1..2^31");
    if (cfg->precision != SLGPU_FP64 && cfg->precision != SLGPU_MIXED)
        return fail(SLGPU_EINVAL, "precision must be SLGPU_FP64 or SLGPU_MIXED");
    if ((cfg->ampfact != 0 && cfg->ampfact != 1) || (cfg->thm_aniso != 0 && cfg->thm_aniso != 1))
        return fail(SLGPU_EINVAL, "ampfact and thm_aniso must be 0 or 1");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(SLGPU_EHIP, "no HIP device visible");
    void *scr = nullptr;
    {
        hipError_t e = scratch_address(&scr);
        if (e != hipSuccess)
            return fail(SLGPU_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    slgpu_ctx *c = new slgpu_ctx();
    c->scratch = scr;
    c->cfg = *cfg;
    c->nx = cfg->nx;
    c->ng = cfg->ng;
    Dev &d = c->d;
    std::memset(&d, 0, sizeof(d));
    d.nx = cfg->nx;
    d.ng = cfg->ng;
    d.nep = cfg->nep;
    d.nepg = cfg->nepg;
    d.tauddmc = cfg->tauddmc;
    d.taulump = cfg->taulump;
    d.puretran = cfg->puretran;
    d.precision = cfg->precision;
    d.ampfact = cfg->ampfact != 0;
    d.thm_aniso = cfg->thm_aniso != 0;
    d.max_steps = cfg->max_steps > 0 ? cfg->max_steps : ((int64_t)1 << 26);
    d.seed = cfg->seed;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(SLGPU_EHIP, "hipStreamCreate failed");
    }
    for (int k = 0; k < 24; ++k) hipEventCreate(&c->ev[k]);
    const size_t nc = c->nx, ng = c->ng, cg = nc * ng, np = cfg->npartmax;
    int rc = 0;
    rc |= dalloc(c, (double **)&d.xarr, nc + 1);
    rc |= dalloc(c, (double **)&d.vxarr, nc + 1);
    rc |= dalloc(c, (double **)&d.divv, nc);
    rc |= dalloc(c, (double **)&d.vol, nc);
    rc |= dalloc(c, (double **)&d.dvdr, nc);
    rc |= dalloc(c, (double **)&d.c3divv, nc);
    rc |= dalloc(c, (double **)&d.wl, ng + 1);
    rc |= dalloc(c, (double **)&d.wlinv, ng + 1);
    rc |= dalloc(c, (double **)&d.gff, 81 * 41);
    rc |= dalloc(c, &d.hckt, nc);
    rc |= dalloc(c, &d.hlparr, nc);
    rc |= dalloc(c, &d.ffB, (size_t)nc * FF_NU);
    rc |= dalloc(c, &d.ffclamp, nc);
    rc |= dalloc(c, &d.ffC, (size_t)nc * FF_NU);
    rc |= dalloc(c, &d.ffW, (size_t)nc * FF_NU);
    rc |= dalloc(c, &d.sig, nc);
    rc |= dalloc(c, &d.tempinv, nc);
    rc |= dalloc(c, &d.temp, nc);
    rc |= dalloc(c, &d.mass, nc);
    rc |= dalloc(c, &d.natomfr, 30 * nc);
    rc |= dalloc(c, &d.el_z, 30);
    rc |= dalloc(c, &d.capT, cg);
    rc |= dalloc(c, &d.cap64, cg);
    rc |= dalloc(c, &d.cap32, cg);
    rc |= dalloc(c, &d.emiss, cg);
    rc |= dalloc(c, &d.capgrey, nc);
    rc |= dalloc(c, &d.capross, nc);
    d.capem32 = d.cap32;
    d.capemgrey = d.capgrey;
    d.nlte = 0;
    rc |= dalloc(c, &d.specarr, cg);
    rc |= dalloc(c, &d.opaclump, 11 * nc);
    rc |= dalloc(c, &d.emitprob, (size_t)cfg->nep * nc);
    rc |= dalloc(c, &d.emitcum, cg);
    rc |= dalloc(c, &d.emit, nc);
    rc |= dalloc(c, &d.radtemp, nc);
    rc |= dalloc(c, &d.srccum, cg);
    rc |= dalloc(c, &d.cell, nc);
    rc |= dalloc(c, &d.natom_in, nc);
    rc |= dalloc(c, &d.vol_in, nc);
    rc |= dalloc(c, &d.nelec_in, nc);
    rc |= dalloc(c, &d.gcapT, cg);
    rc |= dalloc(c, &d.gcell, 4 * nc);
    rc |= dalloc(c, &d.gmpart, (size_t)2 * ((ng + 63) / 64) * nc);
    rc |= dalloc(c, &d.ctpart, (size_t)2 * ((ng + 1023) / 1024) * nc);
    rc |= dalloc(c, &d.gfr, 30 * nc);
    rc |= dalloc(c, &d.cprob, nc);
    rc |= dalloc(c, &d.gprob, cg);
    rc |= dalloc(c, &d.gorder, cg);
    rc |= dalloc(c, &d.cumL, cg);
    rc |= dalloc(c, &d.cumR, cg);
    rc |= dalloc(c, &d.cumD, cg);
    rc |= dalloc(c, &d.cumS, cg);
    rc |= dalloc(c, &d.btab, (size_t)NTAB * nc * BT_K);
    rc |= dalloc(c, (int64_t **)&d.pstart, nc + 1);
    rc |= dalloc(c, &d.px, np);
    rc |= dalloc(c, &d.pmu, np);
    rc |= dalloc(c, &d.pwl, np);
    rc |= dalloc(c, &d.pe, np);
    rc |= dalloc(c, &d.pe0, np);
    rc |= dalloc(c, &d.pstat, np);
    rc |= dalloc(c, &d.pig, np);
    rc |= dalloc(c, &d.perr, np);
    rc |= dalloc(c, &d.pnstep, np);
    rc |= dalloc(c, &d.pix, np);
    rc |= dalloc(c, &d.tally, nc);
    rc |= dalloc(c, &d.eamp, nc);
    // grd_jrad gets one atomic per event; on small grids every lane hits a few hot
    // (cell, group) entries, so blocks add into one of R replicas (R from a 256 MiB budget,
    // 1..32) that a reduction folds into replica 0 after the transport
    {
        const size_t budget = (size_t)256 << 20;
        size_t r = budget / (cg * sizeof(double));
        d.jrad_copies = (int)(r < 1 ? 1 : (r > 32 ? 32 : r));
    }
    rc |= dalloc(c, &d.jrad, (size_t)d.jrad_copies * cg);
    rc |= dalloc(c, &d.tot, 8);
    rc |= dalloc(c, &d.ct, 10);
    rc |= dalloc(c, &d.queue, 1);
    rc |= dalloc(c, &d.lum, ng);
    rc |= dalloc(c, &d.dev, ng);
    rc |= dalloc(c, &d.num, ng);
    rc |= dalloc(c, &d.flags, 1);
    rc |= dalloc(c, &d.eos_ni, 30);
    rc |= dalloc(c, &d.eos_e, 30 * 31);
    rc |= dalloc(c, &d.eos_q, 30 * 31);
    rc |= dalloc(c, &d.eos_g1, 30 * 31);
    rc |= dalloc(c, &d.eos_amass, 30);
    rc |= dalloc(c, &d.eos_sahac2, nc);
    rc |= dalloc(c, &c->eos_temp, nc);
    rc |= dalloc(c, &c->eos_mass, nc);
    rc |= dalloc(c, &c->eos_massfr, 30 * nc);
    rc |= dalloc(c, &c->eos_nelec, nc);
    rc |= dalloc(c, &c->eos_natom, nc);
    rc |= dalloc(c, &c->eos_natomfr, 30 * nc);
    rc |= dalloc(c, &c->eos_niter, nc);
    if (rc) {
        std::string m = g_err;
        slgpu_destroy(c);
        return fail(SLGPU_ENOMEM, m);
    }
    hipStreamSynchronize(c->stream);
    *out = c;
    return 0;
}

int slgpu_destroy(slgpu_ctx *c)
{
    if (!c) return 0;
    if (c->stream) hipStreamSynchronize(c->stream);
    for (void *p : c->allocations) hipFree(p);
    c->allocations.clear();
    for (int k = 0; k < 24; ++k) hipEventDestroy(c->ev[k]);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

void *slgpu_stream(slgpu_ctx *c) { return c ? (void *)c->stream : nullptr; }

int slgpu_set_grid(slgpu_ctx *c, const double *xarr, const double *vxarr, const double *divv,
                   const double *vol, const double *rcell)
{
    if (!c || !xarr || !vxarr || !divv || !vol || !rcell) return fail(SLGPU_EINVAL, "null argument");
    const int nx = c->nx;
    for (int i = 0; i < nx; ++i)
        if (!(xarr[i + 1] > xarr[i])) return fail(SLGPU_EINVAL, "grd_xarr must increase");
    for (int i = 0; i < nx; ++i)
        if (!(vol[i] > 0.0)) return fail(SLGPU_EINVAL, "grd_vol must be positive");
    c->h_xarr.assign(xarr, xarr + nx + 1);
    c->h_vxarr.assign(vxarr, vxarr + nx + 1);
    c->h_divv.assign(divv, divv + nx);
    c->h_vol.assign(vol, vol + nx);
    c->h_rcell.assign(rcell, rcell + nx);
    HIPCHK(hipMemcpyAsync((void *)c->d.xarr, xarr, (nx + 1) * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync((void *)c->d.vxarr, vxarr, (nx + 1) * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync((void *)c->d.divv, divv, nx * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync((void *)c->d.vol, vol, nx * 8, hipMemcpyHostToDevice, c->stream));
    // per-cell constants of the transport, with the kernels' own expression order (one
    // IEEE operation each, so the values are the ones the step would compute)
    std::vector<double> dvdr(nx), c3(nx);
    for (int i = 0; i < nx; ++i) {
        dvdr[i] = (vxarr[i + 1] - vxarr[i]) / (xarr[i + 1] - xarr[i]);
        c3[i] = (3.0 * PC_C) / divv[i];
    }
    HIPCHK(hipMemcpyAsync((void *)c->d.dvdr, dvdr.data(), nx * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync((void *)c->d.c3divv, c3.data(), nx * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->have_grid = true;
    return 0;
}

static int upload_lines(slgpu_ctx *c);
static int build_bfxs(slgpu_ctx *c);

int slgpu_set_groups(slgpu_ctx *c, const double *wl)
{
    if (!c || !wl) return fail(SLGPU_EINVAL, "null argument");
    const int ng = c->ng;
    for (int g = 0; g < ng; ++g)
        if (!(wl[g + 1] > wl[g]) || !(wl[g] > 0.0)) return fail(SLGPU_EINVAL, "grp_wl must be positive and increasing");
    c->h_wl.assign(wl, wl + ng + 1);
    std::vector<double> inv(ng + 1);
    for (int g = 0; g <= ng; ++g) inv[g] = 1.0 / wl[g];     // groupmod.f:50-51
    HIPCHK(hipMemcpyAsync((void *)c->d.wl, wl, (ng + 1) * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync((void *)c->d.wlinv, inv.data(), (ng + 1) * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->have_groups = true;
    int rc = 0;
    if (c->have_lines && c->have_ions) rc = upload_lines(c);
    if (!rc && c->have_bf && c->have_ions) rc = build_bfxs(c);
    return rc;
}

int slgpu_set_ions(slgpu_ctx *c, int nion, const int *ion_iz, const int *ion_ii)
{
    if (!c || nion < 0 || (nion > 0 && (!ion_iz || !ion_ii))) return fail(SLGPU_EINVAL, "bad ion list");
    for (int k = 0; k < nion; ++k) {
        if (ion_iz[k] < 1 || ion_iz[k] > 30 || ion_ii[k] < 1 || ion_ii[k] > ion_iz[k])
            return fail(SLGPU_EINVAL, "ion (iz, ii) out of range");
        if (k > 0 && (ion_iz[k] < ion_iz[k - 1] || (ion_iz[k] == ion_iz[k - 1] && ion_ii[k] <= ion_ii[k - 1])))
            return fail(SLGPU_EINVAL, "ion list must be ascending in (iz, ii)");
    }
    c->nion = nion;
    c->d.nion = nion;
    c->h_ion_iz.assign(ion_iz, ion_iz + nion);
    c->h_ion_ii.assign(ion_ii, ion_ii + nion);
    dfree(c, c->d.ion_z);
    dfree(c, c->d.ion_i);
    dfree(c, c->eos_oc);
    dfree(c, c->eos_ginv);
    dfree(c, c->d.ocg);
    dfree(c, c->d.ocT);
    dfree(c, c->d.ginvT);
    dfree(c, c->d.goc);
    dfree(c, (void *)c->d.bfxs);
    const size_t nc = c->nx;
    int rc = 0;
    rc |= dalloc(c, &c->d.ocg, (size_t)std::max(nion, 1) * nc);
    rc |= dalloc(c, &c->d.ginvT, (size_t)std::max(nion, 1) * nc);
    rc |= dalloc(c, &c->d.goc, (size_t)std::max(nion, 1) * nc);
    rc |= dalloc(c, &c->d.ocT, (size_t)(nion + 30) * nc);     // tail: ff coordinates per element
    rc |= dalloc(c, (double **)&c->d.bfxs, (size_t)std::max(nion, 1) * c->ng);
    rc |= dalloc(c, &c->d.ion_z, std::max(nion, 1));
    rc |= dalloc(c, &c->d.ion_i, std::max(nion, 1));
    rc |= dalloc(c, &c->eos_oc, (size_t)std::max(nion, 1) * nc);
    rc |= dalloc(c, &c->eos_ginv, (size_t)std::max(nion, 1) * nc);
    if (rc) return rc;
    if (nion > 0) {
        HIPCHK(hipMemcpyAsync(c->d.ion_z, ion_iz, nion * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->d.ion_i, ion_ii, nion * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    c->have_ions = true;
    c->have_eos = false;
    if (c->have_lines && c->have_groups) rc = upload_lines(c);
    if (!rc && c->have_bf && c->have_groups) rc = build_bfxs(c);
    return rc;
}

int slgpu_set_lines(slgpu_ctx *c, int64_t nline, const float *wl0, const float *gxs,
                    const float *chilw, const int16_t *iz, const int16_t *ii)
{
    if (!c || nline < 0 || (nline > 0 && (!wl0 || !gxs || !chilw || !iz || !ii)))
        return fail(SLGPU_EINVAL, "bad line list");
    if (nline > ((int64_t)1 << 31) - 1) return fail(SLGPU_EINVAL, "too many lines");
    for (int64_t l = 1; l < nline; ++l)
        if (wl0[l] < wl0[l - 1]) return fail(SLGPU_EINVAL, "line list must be sorted by wl0 (bbxsmod.f:126-179)");
    c->nline = nline;
    c->h_wl0.assign(wl0, wl0 + nline);
    c->h_gxs.assign(gxs, gxs + nline);
    c->h_chilw.assign(chilw, chilw + nline);
    c->h_liz.assign(iz, iz + nline);
    c->h_lii.assign(ii, ii + nline);
    c->have_lines = true;
    if (c->have_groups && c->have_ions) return upload_lines(c);
    return 0;
}

// Lines of group g (1-based) are those with wlsub(g) <= wl0 < wlsub(g+1), the
// pointer walk of physical_opacity.f:118-125 on a sorted list; lines outside the grid and
// lines of ions without populations are dropped (they add nothing, physical_opacity.f:146).
static int upload_lines(slgpu_ctx *c)
{
    const int ng = c->ng;
    int lut[31][32];
    for (auto &r : lut)
        for (int &v : r) v = -1;
    for (int k = 0; k < c->nion; ++k) lut[c->h_ion_iz[k]][c->h_ion_ii[k]] = k;
    std::vector<LineRec> recs;
    recs.reserve(c->nline);
    std::vector<int> cnt(ng + 2, 0);
    int g = 1;
    for (int64_t l = 0; l < c->nline; ++l) {
        const double w = (double)c->h_wl0[l] * PC_ANG;
        if (w < c->h_wl[0] || !(w < c->h_wl[ng])) continue;
        while (g < ng && !(c->h_wl[g] > w)) ++g;      // wlsub(g) <= w < wlsub(g+1)
        const int zi = c->h_liz[l], ii = c->h_lii[l];
        const int k = (zi >= 1 && zi <= 30 && ii >= 1 && ii <= 31) ? lut[zi][ii] : -1;
        if (k < 0) continue;
        recs.push_back(LineRec{c->h_gxs[l], c->h_chilw[l], k, 0});
        cnt[g]++;
    }
    std::vector<int> gstart(ng + 2, 0);
    for (int q = 1; q <= ng; ++q) gstart[q + 1] = gstart[q] + cnt[q];
    c->nline_used = (int64_t)recs.size();
    // the backward's copy: each group's lines ordered by ion (stable), so a run of one ion's
    // lines shares its population and its gradient accumulator
    std::vector<LineRec> recs_ion(recs);
    for (int q = 1; q <= ng; ++q)
        std::stable_sort(recs_ion.begin() + gstart[q], recs_ion.begin() + gstart[q + 1],
                         [](const LineRec &a, const LineRec &b) { return a.ion < b.ion; });
    // pad of an ion-ordered record: the lines left in its run of one ion within its group
    // (itself included), so the backward walks a run as a counted loop
    for (int q = 1; q <= ng; ++q)
        for (int l = gstart[q + 1] - 1; l >= gstart[q]; --l)
            recs_ion[l].pad = (l + 1 < gstart[q + 1] && recs_ion[l + 1].ion == recs_ion[l].ion)
                                  ? recs_ion[l + 1].pad + 1 : 1;
    dfree(c, (void *)c->d.lines);
    dfree(c, (void *)c->d.lines_ion);
    dfree(c, (void *)c->d.gstart);
    int rc = 0;
    rc |= dalloc(c, (LineRec **)&c->d.lines, recs.size());
    rc |= dalloc(c, (LineRec **)&c->d.lines_ion, recs.size());
    rc |= dalloc(c, (int **)&c->d.gstart, ng + 2);
    if (rc) return rc;
    if (!recs.empty()) {
        HIPCHK(hipMemcpyAsync((void *)c->d.lines, recs.data(), recs.size() * sizeof(LineRec),
                              hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync((void *)c->d.lines_ion, recs_ion.data(), recs_ion.size() * sizeof(LineRec),
                              hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(hipMemcpyAsync((void *)c->d.gstart, gstart.data(), (ng + 2) * sizeof(int),
                          hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

static int build_bfxs(slgpu_ctx *c)
{
    float *d1 = nullptr, *d2 = nullptr;
    int *diz = nullptr, *dii = nullptr;
    int rc = 0;
    rc |= dalloc(c, &d1, c->h_ph1.size());
    rc |= dalloc(c, &d2, c->h_ph2.size());
    rc |= dalloc(c, &diz, std::max(c->nion, 1));
    rc |= dalloc(c, &dii, std::max(c->nion, 1));
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(d1, c->h_ph1.data(), c->h_ph1.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d2, c->h_ph2.data(), c->h_ph2.size() * 4, hipMemcpyHostToDevice, c->stream));
    if (c->nion > 0) {
        HIPCHK(hipMemcpyAsync(diz, c->h_ion_iz.data(), c->nion * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(dii, c->h_ion_ii.data(), c->nion * 4, hipMemcpyHostToDevice, c->stream));
    }
    launch_bfxs(c, d1, d2, diz, dii);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    dfree(c, d1); dfree(c, d2); dfree(c, diz); dfree(c, dii);
    return 0;
}

int slgpu_set_bf_tables(slgpu_ctx *c, const float *ph1, const float *ph2)
{
    if (!c || !ph1 || !ph2) return fail(SLGPU_EINVAL, "null argument");
    c->h_ph1.assign(ph1, ph1 + 6 * 7 * 30 * 30);
    c->h_ph2.assign(ph2, ph2 + 7 * 30 * 30);
    c->have_bf = true;
    if (c->have_groups && c->have_ions) return build_bfxs(c);
    return 0;
}

int slgpu_set_ff_table(slgpu_ctx *c, const double *gff)
{
    if (!c || !gff) return fail(SLGPU_EINVAL, "null argument");
    HIPCHK(hipMemcpyAsync((void *)c->d.gff, gff, 81 * 41 * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->have_ff = true;
    return 0;
}

}  // extern "C"

namespace {

// The opacity build (O2-O8) with the per-call staging of the ion populations oc and ginv
// (2 * nion * nc doubles) at `stage`: a fresh allocation for slgpu_opacity_build, the
// library's static scratch array for slgpu_opacity_build_noalloc.
// the NLTE part of a build (physical_opacity_nlte): per-call inputs and extra outputs
struct NlteCall {
    const double *radtemp, *ndens;
    float *out_capemit;
    double *out_capemitgrey, *out_npop;
};

int opacity_build_impl(slgpu_ctx *c, const slgpu_gas *gas, int flags, double *out_cap64,
                       float *out_cap32, double *out_sig, double *out_capgrey,
                       double *out_capross, float *out_emiss, double *stage,
                       const NlteCall *nlc = nullptr)
{
    const int nc = c->nx, ng = c->ng, nion = c->nion;
    // present elements (ff adds exactly zero for the others, physical_opacity.f:240-241)
    int nel = 0, elz[30];
    for (int z = 1; z <= 30; ++z) {
        bool any = false;
        for (int i = 0; i < nc && !any; ++i) any = gas->natomfr[(size_t)i * 30 + z - 1] != 0.0;
        if (any) elz[nel++] = z;
    }
    c->d.nel = nel;
    PhaseTimer tm(c, 0);
    // per-call host inputs (small: nc and nion*nc doubles)
    double *dn = c->d.natom_in, *dv = c->d.vol_in, *de = c->d.nelec_in;   // kept for the backward
    double *doc = stage, *dgi = stage + (size_t)std::max(nion, 1) * nc;
    HIPCHK(hipMemcpyAsync(c->d.temp, gas->temp, nc * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d.mass, gas->mass, nc * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d.natomfr, gas->natomfr, (size_t)nc * 30 * 8, hipMemcpyHostToDevice, c->stream));
    if (nel > 0) HIPCHK(hipMemcpyAsync(c->d.el_z, elz, nel * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(dn, gas->natom, nc * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(dv, gas->vol, nc * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(de, gas->nelec, nc * 8, hipMemcpyHostToDevice, c->stream));
    if (nion > 0) {
        HIPCHK(hipMemcpyAsync(doc, gas->oc, (size_t)nion * nc * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(dgi, gas->ginv, (size_t)nion * nc * 8, hipMemcpyHostToDevice, c->stream));
    }
    if (nlc) {
        HIPCHK(hipMemcpyAsync(c->nl.radtemp, nlc->radtemp, nc * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->nl.ndens, nlc->ndens, (size_t)2 * nc * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemsetAsync(c->nl.flag, 0, sizeof(int), c->stream));
    }
    hipEventRecord(c->ev[22], c->stream);
    launch_cell_prep(c, dn, dv, de, doc, dgi, flags);
    hipEventRecord(c->ev[10], c->stream);
    if (nlc) launch_opacity_nlte(c, flags, ff_prefactor(), rgg_scratch(c));
    else launch_opacity(c, flags);
    hipEventRecord(c->ev[11], c->stream);
    if (nlc) {
        launch_cell_tables(c);
        launch_nlte_means(c);
    } else {
        launch_opacity_finish(c);
    }
    hipEventRecord(c->ev[23], c->stream);
    HIPCHK(hipGetLastError());
    const size_t cg = (size_t)nc * ng;
    if (nlc) {
        if (nlc->out_capemit)
            HIPCHK(hipMemcpyAsync(nlc->out_capemit, c->nl.capemit32, cg * 4, hipMemcpyDeviceToHost, c->stream));
        if (nlc->out_capemitgrey)
            HIPCHK(hipMemcpyAsync(nlc->out_capemitgrey, c->nl.capemitgrey, nc * 8, hipMemcpyDeviceToHost, c->stream));
        if (nlc->out_npop)
            HIPCHK(hipMemcpyAsync(nlc->out_npop, c->nl.pop, (size_t)nc * c->nl.n * 8, hipMemcpyDeviceToHost, c->stream));
    }
    if (out_cap64) HIPCHK(hipMemcpyAsync(out_cap64, c->d.cap64, cg * 8, hipMemcpyDeviceToHost, c->stream));
    if (out_cap32) HIPCHK(hipMemcpyAsync(out_cap32, c->d.cap32, cg * 4, hipMemcpyDeviceToHost, c->stream));
    if (out_emiss) HIPCHK(hipMemcpyAsync(out_emiss, c->d.emiss, cg * 4, hipMemcpyDeviceToHost, c->stream));
    if (out_sig) HIPCHK(hipMemcpyAsync(out_sig, c->d.sig, nc * 8, hipMemcpyDeviceToHost, c->stream));
    if (out_capgrey) HIPCHK(hipMemcpyAsync(out_capgrey, c->d.capgrey, nc * 8, hipMemcpyDeviceToHost, c->stream));
    if (out_capross) HIPCHK(hipMemcpyAsync(out_capross, c->d.capross, nc * 8, hipMemcpyDeviceToHost, c->stream));
    int rc = tm.stop();
    if (!rc && nlc) {
        int flag = 0;
        HIPCHK(hipMemcpy(&flag, c->nl.flag, sizeof(int), hipMemcpyDeviceToHost));
        if (flag) {
            c->have_opacity = false;        // the device opacity is partly overwritten
            c->have_prep = false;
            return fail(SLGPU_EINVAL, "nlte opacity: rate matrix numerically singular");
        }
    }
    float kms = 0.f;
    hipEventElapsedTime(&kms, c->ev[10], c->ev[11]);
    c->t_ms[5] += kms;
    hipEventElapsedTime(&kms, c->ev[22], c->ev[23]);
    c->t_ms[11] += kms;
    c->opac_flags = flags;
    // the emission opacity the DDMC prep and the samplers read (gridmod.f:41,56 grd_capemit,
    // grd_capemitgrey; the absorption opacity in LTE)
    c->d.nlte = nlc ? 1 : 0;
    c->d.capem32 = nlc ? c->nl.capemit32 : c->d.cap32;
    c->d.capemgrey = nlc ? c->nl.capemitgrey : c->d.capgrey;
    c->have_prep = false;
    if (rc) {
        c->have_opacity = false;
        return rc;
    }
    c->have_opacity = true;
    return 0;
}

int check_opacity_inputs(slgpu_ctx *c, const slgpu_gas *gas)
{
    if (!c || !gas || !gas->temp || !gas->natom || !gas->vol || !gas->mass || !gas->nelec ||
        !gas->natomfr || (c->nion > 0 && (!gas->oc || !gas->ginv)))
        return fail(SLGPU_EINVAL, "null gas state");
    if (!(c->have_groups && c->have_lines && c->have_bf && c->have_ff && c->have_ions))
        return fail(SLGPU_EORDER, "groups, lines, ions and bf/ff tables must be set first");
    for (int i = 0; i < c->nx; ++i)
        if (!(gas->temp[i] > 0.0)) return fail(SLGPU_EINVAL, "gas_temp must be positive (gas_update.f:71-72)");
    return 0;
}

}  // namespace

extern "C" {

int slgpu_opacity_build(slgpu_ctx *c, const slgpu_gas *gas, int flags, double *out_cap64,
                        float *out_cap32, double *out_sig, double *out_capgrey,
                        double *out_capross, float *out_emiss)
{
    int rc = check_opacity_inputs(c, gas);
    if (rc) return rc;
    double *stage = nullptr;
    rc = dalloc(c, &stage, (size_t)2 * std::max(c->nion, 1) * c->nx);
    if (rc) return rc;
    rc = opacity_build_impl(c, gas, flags, out_cap64, out_cap32, out_sig, out_capgrey,
                            out_capross, out_emiss, stage);
    hipStreamSynchronize(c->stream);
    dfree(c, stage);
    return rc;
}

int slgpu_opacity_build_noalloc(slgpu_ctx *c, const slgpu_gas *gas, int flags,
                                double *out_cap64, float *out_cap32, double *out_sig,
                                double *out_capgrey, double *out_capross, float *out_emiss)
{
    int rc = check_opacity_inputs(c, gas);
    if (rc) return rc;
    const size_t need = (size_t)2 * std::max(c->nion, 1) * c->nx * sizeof(double);
    if (need > SLGPU_SCRATCH_BYTES)
        return fail(SLGPU_EINVAL, "opacity_build_noalloc: 2*nion*nc doubles exceed the " +
                                      std::to_string(SLGPU_SCRATCH_BYTES) + "-byte scratch array");
    double *stage = (double *)c->scratch;
    if (!stage) return fail(SLGPU_EORDER, "scratch array address not resolved at slgpu_create");
    return opacity_build_impl(c, gas, flags, out_cap64, out_cap32, out_sig, out_capgrey,
                              out_capross, out_emiss, stage);
}

int slgpu_set_nlte_data(slgpu_ctx *c, int nlev, int nlin, int ncoll, const double *glev,
                        const int *llw, const int *lup, const double *wl0, const double *f,
                        const double *gl, const double *gu, const double *aul, const double *bul,
                        const double *blu, const int *llw_coll, const int *lup_coll,
                        const double *coll_c, const double *dele, const int *lev_pi,
                        const double *pi_c, const double *rr_c, const double *dele_pi)
{
    if (!c || !glev || !llw || !lup || !wl0 || !f || !gl || !gu || !aul || !bul || !blu ||
        (ncoll > 0 && (!llw_coll || !lup_coll || !coll_c || !dele)) || !lev_pi || !pi_c ||
        !rr_c || !dele_pi)
        return fail(SLGPU_EINVAL, "null argument");
    if (!(c->have_groups && c->have_ions)) return fail(SLGPU_EORDER, "groups and ions must be set first");
    if (nlev < 2 || nlev > 1024 || nlin < 1 || ncoll < 0)
        return fail(SLGPU_EINVAL, "nlte data: need 2 <= nlev <= 1024, nlin >= 1, ncoll >= 0");
    // distinct level pairs (the reference's rate coefficients are set per pair, so a repeated
    // radiative pair would keep only its last line, a repeated collision pair would add)
    std::vector<char> seen((size_t)nlev * nlev, 0), seenc((size_t)nlev * nlev, 0), seenl(nlev, 0);
    for (int k = 0; k < nlin; ++k) {
        if (llw[k] < 1 || lup[k] > nlev || !(llw[k] < lup[k]))
            return fail(SLGPU_EINVAL, "nlte data: radiative lines need 1 <= llw < lup <= nlev");
        char &v = seen[(size_t)(llw[k] - 1) * nlev + lup[k] - 1];
        if (v) return fail(SLGPU_EINVAL, "nlte data: repeated radiative level pair");
        v = 1;
    }
    for (int k = 0; k < ncoll; ++k) {
        if (llw_coll[k] < 1 || lup_coll[k] > nlev || !(llw_coll[k] < lup_coll[k]))
            return fail(SLGPU_EINVAL, "nlte data: collision lines need 1 <= llw < lup <= nlev");
        char &v = seenc[(size_t)(llw_coll[k] - 1) * nlev + lup_coll[k] - 1];
        if (v) return fail(SLGPU_EINVAL, "nlte data: repeated collision level pair");
        v = 1;
    }
    for (int k = 0; k < nlev; ++k) {
        if (lev_pi[k] < 1 || lev_pi[k] > nlev || seenl[lev_pi[k] - 1])
            return fail(SLGPU_EINVAL, "nlte data: lev_pi must list every level once");
        seenl[lev_pi[k] - 1] = 1;
    }
    const int ng = c->ng, nc = c->nx;
    // group of each line: the first grp_wl edge above wl0 (physical_opacity_nlte.f:166-172),
    // 0 outside the grid; and the lines of each group in file order
    std::vector<int> igs(nlin), cnt(ng + 2, 0);
    for (int k = 0; k < nlin; ++k) {
        const double w = wl0[k] * PC_ANG;
        int g = 0;
        while (g <= ng && !(c->h_wl[g] > w)) ++g;
        igs[k] = (g >= 1 && g <= ng) ? g : 0;
        if (igs[k]) cnt[igs[k]]++;
    }
    std::vector<int> hstart(ng + 2, 0), hidx;
    for (int g = 1; g <= ng; ++g) hstart[g + 1] = hstart[g] + cnt[g];
    std::vector<int> fill(hstart);
    hidx.assign(std::max(hstart[ng + 1], 1), 0);
    for (int k = 0; k < nlin; ++k)
        if (igs[k]) hidx[fill[igs[k]]++] = k;
    int kH = -1;
    for (int k = 0; k < c->nion; ++k)
        if (c->h_ion_iz[k] == 1 && c->h_ion_ii[k] == 1) kH = k;
    NlteDev &q = c->nl;
    if (c->have_nlte) {
        void *ps[] = {q.glev, q.wl0, q.f, q.gl, q.gu, q.aul, q.bul, q.blu, q.cc, q.dele, q.pic, q.rrc,
                      q.delepi, q.llw, q.lup, q.igs, q.cl, q.cu, q.levpi, q.hstart, q.hidx, q.radtemp,
                      q.ndens, q.rc, q.A, q.pop, q.gpi, q.ipiv, q.flag, q.capemitT, q.emissT,
                      q.capemit32, q.capemitgrey};
        for (void *p : ps) dfree(c, p);
        c->have_nlte = false;
        if (c->d.nlte) {            // the last build's emission opacity is gone with them
            c->d.nlte = 0;
            c->d.capem32 = c->d.cap32;
            c->d.capemgrey = c->d.capgrey;
            c->have_opacity = false;
            c->have_prep = false;
        }
    }
    q = NlteDev();
    q.n = nlev; q.nlin = nlin; q.ncoll = ncoll; q.kH = kH;
    const int nc1 = std::max(ncoll, 1);
    int rc = 0;
    rc |= dalloc(c, &q.glev, nlev);
    rc |= dalloc(c, &q.wl0, nlin); rc |= dalloc(c, &q.f, nlin); rc |= dalloc(c, &q.gl, nlin);
    rc |= dalloc(c, &q.gu, nlin); rc |= dalloc(c, &q.aul, nlin); rc |= dalloc(c, &q.bul, nlin);
    rc |= dalloc(c, &q.blu, nlin);
    rc |= dalloc(c, &q.cc, (size_t)4 * nc1); rc |= dalloc(c, &q.dele, nc1);
    rc |= dalloc(c, &q.pic, (size_t)4 * nlev); rc |= dalloc(c, &q.rrc, (size_t)4 * nlev);
    rc |= dalloc(c, &q.delepi, nlev);
    rc |= dalloc(c, &q.llw, nlin); rc |= dalloc(c, &q.lup, nlin); rc |= dalloc(c, &q.igs, nlin);
    rc |= dalloc(c, &q.cl, nc1); rc |= dalloc(c, &q.cu, nc1); rc |= dalloc(c, &q.levpi, nlev);
    rc |= dalloc(c, &q.hstart, ng + 2); rc |= dalloc(c, &q.hidx, hidx.size());
    rc |= dalloc(c, &q.radtemp, nc); rc |= dalloc(c, &q.ndens, (size_t)2 * nc);
    rc |= dalloc(c, &q.rc, (size_t)nc * nlev * nlev); rc |= dalloc(c, &q.A, (size_t)nc * nlev * nlev);
    rc |= dalloc(c, &q.pop, (size_t)nc * nlev); rc |= dalloc(c, &q.gpi, (size_t)nc * nlev);
    rc |= dalloc(c, &q.ipiv, (size_t)nc * nlev); rc |= dalloc(c, &q.flag, 1);
    rc |= dalloc(c, &q.capemitT, (size_t)nc * ng); rc |= dalloc(c, &q.emissT, (size_t)nc * ng);
    rc |= dalloc(c, &q.capemit32, (size_t)nc * ng); rc |= dalloc(c, &q.capemitgrey, nc);
    if (rc) return rc;
    auto up = [&](void *dst, const void *src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream);
    };
    HIPCHK(up(q.glev, glev, nlev * 8));
    HIPCHK(up(q.wl0, wl0, nlin * 8)); HIPCHK(up(q.f, f, nlin * 8)); HIPCHK(up(q.gl, gl, nlin * 8));
    HIPCHK(up(q.gu, gu, nlin * 8)); HIPCHK(up(q.aul, aul, nlin * 8)); HIPCHK(up(q.bul, bul, nlin * 8));
    HIPCHK(up(q.blu, blu, nlin * 8));
    if (ncoll > 0) {
        HIPCHK(up(q.cc, coll_c, (size_t)4 * ncoll * 8)); HIPCHK(up(q.dele, dele, ncoll * 8));
        HIPCHK(up(q.cl, llw_coll, ncoll * 4)); HIPCHK(up(q.cu, lup_coll, ncoll * 4));
    }
    HIPCHK(up(q.pic, pi_c, (size_t)4 * nlev * 8)); HIPCHK(up(q.rrc, rr_c, (size_t)4 * nlev * 8));
    HIPCHK(up(q.delepi, dele_pi, nlev * 8));
    HIPCHK(up(q.llw, llw, nlin * 4)); HIPCHK(up(q.lup, lup, nlin * 4)); HIPCHK(up(q.igs, igs.data(), nlin * 4));
    HIPCHK(up(q.levpi, lev_pi, nlev * 4));
    HIPCHK(up(q.hstart, hstart.data(), (ng + 2) * 4)); HIPCHK(up(q.hidx, hidx.data(), hidx.size() * 4));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->have_nlte = true;
    return 0;
}

int slgpu_opacity_build_nlte(slgpu_ctx *c, const slgpu_gas *gas, const double *radtemp,
                             const double *ndens, int flags, double *out_cap64, float *out_cap32,
                             float *out_capemit, double *out_sig, double *out_capgrey,
                             double *out_capemitgrey, double *out_capross, float *out_emiss,
                             double *out_npop)
{
    int rc = check_opacity_inputs(c, gas);
    if (rc) return rc;
    if (!c->have_nlte) return fail(SLGPU_EORDER, "NLTE data must be set first");
    if (!radtemp || !ndens) return fail(SLGPU_EINVAL, "null argument");
    if (!(flags & SLGPU_NOBB))
        for (int i = 0; i < c->nx; ++i) {
            if (gas->mass[i] <= 0.0) continue;
            if (!(radtemp[i] > 0.0) || !(ndens[2 * (size_t)i] > 0.0) || !(ndens[2 * (size_t)i + 1] >= 0.0))
                return fail(SLGPU_EINVAL, "nlte opacity: radtemp and the H I density must be positive "
                                          "and the H II density non-negative in every massive cell");
        }
    double *stage = nullptr;
    rc = dalloc(c, &stage, (size_t)2 * std::max(c->nion, 1) * c->nx);
    if (rc) return rc;
    NlteCall nlc{radtemp, ndens, out_capemit, out_capemitgrey, out_npop};
    rc = opacity_build_impl(c, gas, flags, out_cap64, out_cap32, out_sig, out_capgrey,
                            out_capross, out_emiss, stage, &nlc);
    hipStreamSynchronize(c->stream);
    dfree(c, stage);
    return rc;
}

int slgpu_nlte_rad_est(slgpu_ctx *c, double *radtemp, double *out_jrad)
{
    if (!c || !radtemp) return fail(SLGPU_EINVAL, "null argument");
    if (!c->have_transport) return fail(SLGPU_EORDER, "slgpu_transport must run first");
    const int nc = c->nx, ng = c->ng;
    for (int i = 0; i < nc; ++i)
        if (!(radtemp[i] > 0.0)) return fail(SLGPU_EINVAL, "grd_radtemp must be positive");
    double *drt = nullptr, *dj = nullptr;
    int rc = dalloc(c, &drt, nc);
    if (!rc && out_jrad) rc = dalloc(c, &dj, (size_t)nc * ng);
    if (!rc) {
        HIPCHK(hipMemcpyAsync(drt, radtemp, nc * 8, hipMemcpyHostToDevice, c->stream));
        launch_nlte_rad_est(c, drt, dj);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(radtemp, drt, nc * 8, hipMemcpyDeviceToHost, c->stream));
        if (dj) HIPCHK(hipMemcpyAsync(out_jrad, dj, (size_t)nc * ng * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    if (drt) dfree(c, drt);
    if (dj) dfree(c, dj);
    return rc;
}

int slgpu_ddmc_prep(slgpu_ctx *c, const double *radtemp, double lbol, double *out_opaclump,
                    double *out_emitprob, double *out_emit, double *out_sthermal)
{
    if (!c || !radtemp) return fail(SLGPU_EINVAL, "null argument");
    if (!c->have_opacity || !c->have_grid) return fail(SLGPU_EORDER, "opacity and grid must be set first");
    const int nc = c->nx;
    for (int i = 0; i < nc; ++i)
        if (!(radtemp[i] > 0.0)) return fail(SLGPU_EINVAL, "grd_radtemp must be positive");
    // SOURCE/sourceenergy.f90:23-50 (surface-like thermal source)
    std::vector<double> emit(nc);
    double s = 0.0;
    for (int i = 0; i < nc; ++i) {
        const double t2 = radtemp[i] * radtemp[i];
        emit[i] = PC_PI * PC_C * (c->h_rcell[i] * c->h_rcell[i]) * PC_ACOEF * (t2 * t2);
        s += emit[i];
    }
    double st = 0.0;
    for (int i = 0; i < nc; ++i) { emit[i] = emit[i] / s * lbol; st += emit[i]; }
    PhaseTimer tm(c, 1);
    HIPCHK(hipMemcpyAsync(c->d.emit, emit.data(), nc * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d.radtemp, radtemp, nc * 8, hipMemcpyHostToDevice, c->stream));
    launch_ddmc_tables(c);
    launch_ddmc_probs(c);
    HIPCHK(hipGetLastError());
    if (out_opaclump) HIPCHK(hipMemcpyAsync(out_opaclump, c->d.opaclump, (size_t)11 * nc * 8, hipMemcpyDeviceToHost, c->stream));
    if (out_emitprob) HIPCHK(hipMemcpyAsync(out_emitprob, c->d.emitprob, (size_t)c->cfg.nep * nc * 8, hipMemcpyDeviceToHost, c->stream));
    int rc = tm.stop();
    if (rc) return rc;
    if (out_emit) std::memcpy(out_emit, emit.data(), nc * 8);
    if (out_sthermal) *out_sthermal = st;
    c->h_emit_cache = emit;
    c->have_prep = true;
    return 0;
}

int slgpu_sources(slgpu_ctx *c, int64_t nstot, uint32_t iteration, int *out_nvol,
                  int64_t *out_nnew)
{
    if (!c) return fail(SLGPU_EINVAL, "null context");
    if (!c->have_prep) return fail(SLGPU_EORDER, "slgpu_ddmc_prep must run first");
    if (nstot < 1) return fail(SLGPU_EINVAL, "nstot must be positive");
    std::vector<int> nvol;
    int64_t nnew = source_numbers(c->h_emit_cache, nstot, nvol);
    if (nnew == -1) return fail(SLGPU_EINVAL, "sourcenumbers: no cell emits");
    if (nnew == -2) return fail(SLGPU_EINVAL, "sourcenumbers: nextra>grd_ncell (sourcenumbers.f90:130-131)");
    if (nnew > c->cfg.npartmax) return fail(SLGPU_EINVAL, "vacancies: nnew>npartmax (vacancies.f90:21)");
    std::vector<int64_t> pstart(c->nx + 1, 0);
    for (int i = 0; i < c->nx; ++i) pstart[i + 1] = pstart[i] + nvol[i];
    PhaseTimer tm(c, 2);
    HIPCHK(hipMemcpyAsync((void *)c->d.pstart, pstart.data(), (c->nx + 1) * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(c->d.flags, 0, 4, c->stream));
    c->d.npart = nnew;
    launch_sources(c, iteration);
    HIPCHK(hipGetLastError());
    int flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, c->d.flags, 4, hipMemcpyDeviceToHost, c->stream));
    int rc = tm.stop();
    if (rc) return rc;
    if (flag) return fail(SLGPU_EINVAL, "interior_source: position rejection did not terminate");
    if (out_nvol) std::memcpy(out_nvol, nvol.data(), c->nx * sizeof(int));
    if (out_nnew) *out_nnew = nnew;
    c->have_sources = true;
    c->have_transport = false;
    return 0;
}

int slgpu_transport(slgpu_ctx *c, uint32_t iteration, double *out_tally, double *out_jrad,
                    double *out_eamp, double *out_totals, int64_t *out_counters)
{
    if (!c) return fail(SLGPU_EINVAL, "null context");
    if (!c->have_sources) return fail(SLGPU_EORDER, "slgpu_sources must run first");
    const int nc = c->nx, ng = c->ng;
    PhaseTimer tm(c, 3);
    // particle_advance.f90:65-67: zero the tallies
    HIPCHK(hipMemsetAsync(c->d.tally, 0, nc * 8, c->stream));
    HIPCHK(hipMemsetAsync(c->d.eamp, 0, nc * 8, c->stream));
    HIPCHK(hipMemsetAsync(c->d.jrad, 0, (size_t)c->d.jrad_copies * nc * ng * 8, c->stream));
    HIPCHK(hipMemsetAsync(c->d.tot, 0, 8 * 8, c->stream));
    HIPCHK(hipMemsetAsync(c->d.ct, 0, 10 * 8, c->stream));
    hipEventRecord(c->ev[12], c->stream);
    HIPCHK(hipMemsetAsync(c->d.queue, 0, 8, c->stream));
    launch_transport(c, iteration);
    hipEventRecord(c->ev[13], c->stream);
    launch_jrad_reduce(c);
    HIPCHK(hipGetLastError());
    double tot[8];
    unsigned long long ct[10];
    HIPCHK(hipMemcpyAsync(tot, c->d.tot, 8 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(ct, c->d.ct, 10 * 8, hipMemcpyDeviceToHost, c->stream));
    if (out_tally) HIPCHK(hipMemcpyAsync(out_tally, c->d.tally, nc * 8, hipMemcpyDeviceToHost, c->stream));
    if (out_eamp) HIPCHK(hipMemcpyAsync(out_eamp, c->d.eamp, nc * 8, hipMemcpyDeviceToHost, c->stream));
    if (out_jrad) HIPCHK(hipMemcpyAsync(out_jrad, c->d.jrad, (size_t)nc * ng * 8, hipMemcpyDeviceToHost, c->stream));
    int rc = tm.stop();
    if (rc) return rc;
    float kms = 0.f;
    hipEventElapsedTime(&kms, c->ev[12], c->ev[13]);
    c->t_ms[6] += kms;
    tot[2] = -tot[0];     // tot_sflux = -sum e(flux) (particle_advance.f90:283)
    if (out_totals)
        for (int k = 0; k < 7; ++k) out_totals[k] = tot[k];
    if (out_counters)
        for (int k = 0; k < 10; ++k) out_counters[k] = (int64_t)ct[k];
    c->have_transport = true;
    if (ct[7] > 0 && c->cfg.errorfatal)
        return fail(SLGPU_ETRANSPORT, "particle_advance: fatal transport error (" + std::to_string(ct[7]) + " particles)");
    return 0;
}

int slgpu_fluxtally(slgpu_ctx *c, double *out_luminos, double *out_lumdev, int *out_lumnum)
{
    if (!c) return fail(SLGPU_EINVAL, "null context");
    if (!c->have_transport) return fail(SLGPU_EORDER, "slgpu_transport must run first");
    const int ng = c->ng;
    PhaseTimer tm(c, 4);
    HIPCHK(hipMemsetAsync(c->d.lum, 0, ng * 8, c->stream));
    HIPCHK(hipMemsetAsync(c->d.dev, 0, ng * 8, c->stream));
    HIPCHK(hipMemsetAsync(c->d.num, 0, ng * 4, c->stream));
    launch_fluxtally(c);
    HIPCHK(hipGetLastError());
    if (out_luminos) HIPCHK(hipMemcpyAsync(out_luminos, c->d.lum, ng * 8, hipMemcpyDeviceToHost, c->stream));
    if (out_lumdev) HIPCHK(hipMemcpyAsync(out_lumdev, c->d.dev, ng * 8, hipMemcpyDeviceToHost, c->stream));
    if (out_lumnum) HIPCHK(hipMemcpyAsync(out_lumnum, c->d.num, ng * 4, hipMemcpyDeviceToHost, c->stream));
    return tm.stop();
}

int slgpu_get_particles(slgpu_ctx *c, int64_t n, uint8_t *stat, double *x, double *mu,
                        double *wl, double *e, int *ig, int64_t *nstep, int *err)
{
    if (!c) return fail(SLGPU_EINVAL, "null context");
    if (n < 0 || n > c->d.npart) return fail(SLGPU_EINVAL, "n exceeds the particle count");
    if (stat) HIPCHK(hipMemcpyAsync(stat, c->d.pstat, n, hipMemcpyDeviceToHost, c->stream));
    if (x) HIPCHK(hipMemcpyAsync(x, c->d.px, n * 8, hipMemcpyDeviceToHost, c->stream));
    if (mu) HIPCHK(hipMemcpyAsync(mu, c->d.pmu, n * 8, hipMemcpyDeviceToHost, c->stream));
    if (wl) HIPCHK(hipMemcpyAsync(wl, c->d.pwl, n * 8, hipMemcpyDeviceToHost, c->stream));
    if (e) HIPCHK(hipMemcpyAsync(e, c->d.pe, n * 8, hipMemcpyDeviceToHost, c->stream));
    if (ig) HIPCHK(hipMemcpyAsync(ig, c->d.pig, n * 4, hipMemcpyDeviceToHost, c->stream));
    if (nstep) HIPCHK(hipMemcpyAsync(nstep, c->d.pnstep, n * 8, hipMemcpyDeviceToHost, c->stream));
    if (err) HIPCHK(hipMemcpyAsync(err, c->d.perr, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

namespace {
// the VJP with its three input gradients on the host (copied in) or, with dev, already on the
// context's device (read in place)
int opacity_vjp_impl(slgpu_ctx *c, bool dev, const double *grad_cap64, const double *grad_capgrey,
                     const double *grad_capross, double *out_grad_temp, double *out_grad_natom,
                     double *out_grad_nelec, double *out_grad_natomfr, double *out_grad_oc)
{
    if (!c) return fail(SLGPU_EINVAL, "null context");
    if (!c->have_opacity) return fail(SLGPU_EORDER, "slgpu_opacity_build must run first");
    if (c->d.nlte) return fail(SLGPU_EORDER, "opacity backward: the last build was NLTE; the backward covers the LTE build");
    const int nc = c->nx, ng = c->ng, nion = c->nion;
    if ((size_t)(nion + c->d.nel) * 64 * sizeof(double) > 65536)
        return fail(SLGPU_EINVAL, "opacity backward: ions + elements exceed 128 per cell");
    PhaseTimer tm(c, 7);
    const size_t cg = (size_t)nc * ng;
    double *dgc = nullptr, *dgp = nullptr, *dgr = nullptr, *out = nullptr;
    int rc = 0;
    if (!dev) {
        if (grad_cap64) rc |= dalloc(c, &dgc, cg);
        if (grad_capgrey) rc |= dalloc(c, &dgp, nc);
        if (grad_capross) rc |= dalloc(c, &dgr, nc);
    }
    const size_t nout = (size_t)nc * (3 + 30 + std::max(nion, 1));
    rc |= dalloc(c, &out, nout);
    if (rc) return rc;
    if (dgc) HIPCHK(hipMemcpyAsync(dgc, grad_cap64, cg * 8, hipMemcpyHostToDevice, c->stream));
    if (dgp) HIPCHK(hipMemcpyAsync(dgp, grad_capgrey, nc * 8, hipMemcpyHostToDevice, c->stream));
    if (dgr) HIPCHK(hipMemcpyAsync(dgr, grad_capross, nc * 8, hipMemcpyHostToDevice, c->stream));
    const double *igc = dev ? grad_cap64 : dgc, *igp = dev ? grad_capgrey : dgp, *igr = dev ? grad_capross : dgr;
    double *gT = out, *gn = out + nc, *ge = out + 2 * nc, *gf = out + 3 * nc, *go = out + 33 * (size_t)nc;
    hipEventRecord(c->ev[20], c->stream);
    launch_opacity_vjp(c, c->opac_flags, igc, igp, igr, gT, gn, ge, gf, go, ff_prefactor(),
                       rgg_scratch(c));
    hipEventRecord(c->ev[21], c->stream);
    HIPCHK(hipGetLastError());
    if (out_grad_temp) HIPCHK(hipMemcpyAsync(out_grad_temp, gT, nc * 8, hipMemcpyDeviceToHost, c->stream));
    if (out_grad_natom) HIPCHK(hipMemcpyAsync(out_grad_natom, gn, nc * 8, hipMemcpyDeviceToHost, c->stream));
    if (out_grad_nelec) HIPCHK(hipMemcpyAsync(out_grad_nelec, ge, nc * 8, hipMemcpyDeviceToHost, c->stream));
    if (out_grad_natomfr) HIPCHK(hipMemcpyAsync(out_grad_natomfr, gf, (size_t)30 * nc * 8, hipMemcpyDeviceToHost, c->stream));
    if (out_grad_oc && nion > 0) HIPCHK(hipMemcpyAsync(out_grad_oc, go, (size_t)nion * nc * 8, hipMemcpyDeviceToHost, c->stream));
    rc = tm.stop();
    if (!rc) {
        float kms = 0.f;
        hipEventElapsedTime(&kms, c->ev[20], c->ev[21]);
        c->t_ms[10] += kms;
    }
    dfree(c, dgc); dfree(c, dgp); dfree(c, dgr); dfree(c, out);
    return rc;
}
}  // namespace

int slgpu_opacity_vjp(slgpu_ctx *c, const double *grad_cap64, const double *grad_capgrey,
                      const double *grad_capross, double *out_grad_temp, double *out_grad_natom,
                      double *out_grad_nelec, double *out_grad_natomfr, double *out_grad_oc)
{
    return opacity_vjp_impl(c, false, grad_cap64, grad_capgrey, grad_capross, out_grad_temp,
                            out_grad_natom, out_grad_nelec, out_grad_natomfr, out_grad_oc);
}

int slgpu_opacity_vjp_dev(slgpu_ctx *c, const double *d_grad_cap64, const double *d_grad_capgrey,
                          const double *d_grad_capross, double *out_grad_temp, double *out_grad_natom,
                          double *out_grad_nelec, double *out_grad_natomfr, double *out_grad_oc)
{
    return opacity_vjp_impl(c, true, d_grad_cap64, d_grad_capgrey, d_grad_capross, out_grad_temp,
                            out_grad_natom, out_grad_nelec, out_grad_natomfr, out_grad_oc);
}

int slgpu_opacity_outputs_dev(slgpu_ctx *c, double *d_cap64, double *d_capgrey, double *d_capross)
{
    if (!c) return fail(SLGPU_EINVAL, "null context");
    if (!c->have_opacity) return fail(SLGPU_EORDER, "an opacity build must run first");
    const size_t nc = c->nx, cg = nc * c->ng;
    if (d_cap64) HIPCHK(hipMemcpyAsync(d_cap64, c->d.cap64, cg * 8, hipMemcpyDeviceToDevice, c->stream));
    if (d_capgrey) HIPCHK(hipMemcpyAsync(d_capgrey, c->d.capgrey, nc * 8, hipMemcpyDeviceToDevice, c->stream));
    if (d_capross) HIPCHK(hipMemcpyAsync(d_capross, c->d.capross, nc * 8, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int slgpu_set_unstaged(slgpu_ctx *c, int on)
{
    if (!c) return fail(SLGPU_EINVAL, "null context");
    if (on != 0 && on != 1) return fail(SLGPU_EINVAL, "on must be 0 or 1");
    c->force_unstaged = on;
    return SLGPU_OK;
}

int slgpu_sample_groups(slgpu_ctx *c, int table, int direct, int64_t n, const double *r,
                        const int *cell, int *out_group)
{
    if (!c) return fail(SLGPU_EINVAL, "null context");
    if (!c->have_prep) return fail(SLGPU_EORDER, "slgpu_ddmc_prep must run first");
    if (table < 0 || table >= NTAB) return fail(SLGPU_EINVAL, "table must be in 0..4");
    if (n < 0 || (n > 0 && (!r || !cell || !out_group))) return fail(SLGPU_EINVAL, "bad sample arrays");
    if (n == 0) return 0;
    for (int64_t k = 0; k < n; ++k)
        if (!(r[k] > 0.0 && r[k] <= 1.0)) return fail(SLGPU_EINVAL, "draws must lie in (0, 1]");
    double *dr = nullptr;
    int *dc = nullptr, *dg = nullptr;
    if (hipMalloc(&dr, n * 8) != hipSuccess || hipMalloc(&dc, n * 4) != hipSuccess ||
        hipMalloc(&dg, n * 4) != hipSuccess) {
        hipFree(dr); hipFree(dc); hipFree(dg);
        return fail(SLGPU_ENOMEM, "sample buffers");
    }
    int rc = 0;
    if (hipMemcpyAsync(dr, r, n * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(dc, cell, n * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess) rc = 1;
    if (!rc) {
        launch_sample_groups(c, table, direct, n, dr, dc, dg);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(out_group, dg, n * 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) rc = 1;
    }
    hipFree(dr); hipFree(dc); hipFree(dg);
    return rc ? fail(SLGPU_EHIP, "sample_groups: device error") : 0;
}

int slgpu_set_eos_data(slgpu_ctx *c, const int *ni, const double *e, const double *q,
                       const double *g1, const double *amass)
{
    if (!c || !ni || !e || !q || !g1 || !amass) return fail(SLGPU_EINVAL, "null argument");
    for (int iz = 1; iz <= 30; ++iz) {
        const int n = ni[iz - 1];
        if (n < 1 || n > iz + 1) return fail(SLGPU_EINVAL, "ni(iz) must be in 1..iz+1 (ionsmod.f:102)");
        if (!(amass[iz - 1] > 0.0)) return fail(SLGPU_EINVAL, "element masses must be positive");
        for (int ii = 1; ii <= n; ++ii)
            if (!(q[(iz - 1) * 31 + ii - 1] > 0.0))
                return fail(SLGPU_EINVAL, "partition functions must be positive for every stage");
    }
    HIPCHK(hipMemcpyAsync(c->d.eos_ni, ni, 30 * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d.eos_e, e, 30 * 31 * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d.eos_q, q, 30 * 31 * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d.eos_g1, g1, 30 * 31 * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d.eos_amass, amass, 30 * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->have_eos_data = true;
    c->have_eos = false;
    return 0;
}

int slgpu_eos(slgpu_ctx *c, const double *temp, const double *mass, const double *massfr,
              double acc, double *nelec, double *out_natom, double *out_natomfr, double *out_oc,
              double *out_ginv, int *out_niter)
{
    if (!c || !temp || !mass || !massfr || !nelec) return fail(SLGPU_EINVAL, "null argument");
    if (!(c->have_grid && c->have_ions && c->have_eos_data))
        return fail(SLGPU_EORDER, "grid, ions and EOS data must be set first");
    const int nc = c->nx, nion = c->nion;
    for (int i = 0; i < nc; ++i) {
        if (mass[i] <= 0.0) continue;
        if (!(temp[i] > 0.0) || !(nelec[i] > 0.0))
            return fail(SLGPU_EINVAL, "temperatures and start electron fractions must be positive");
        double s = 0.0;
        for (int l = 0; l < 30; ++l) {
            if (massfr[(size_t)i * 30 + l] < 0.0) return fail(SLGPU_EINVAL, "massfr2natomfr: negative mass fractions");
            s += massfr[(size_t)i * 30 + l];
        }
        if (!(s > 0.0)) return fail(SLGPU_EINVAL, "massfr2natomfr: all mass fractions zero");
    }
    if (!(acc > 0.0)) acc = 1e-8;                            // ionsmod.f:183
    PhaseTimer tm(c, 8);
    HIPCHK(hipMemsetAsync(c->d.flags, 0, 4, c->stream));
    HIPCHK(hipMemcpyAsync(c->eos_temp, temp, nc * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->eos_mass, mass, nc * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->eos_massfr, massfr, (size_t)30 * nc * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->eos_nelec, nelec, nc * 8, hipMemcpyHostToDevice, c->stream));
    launch_eos(c, c->eos_temp, c->eos_mass, c->eos_massfr, acc, c->eos_nelec, c->eos_natom,
               c->eos_natomfr, c->eos_oc, c->eos_ginv, c->eos_niter);
    HIPCHK(hipGetLastError());
    int flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, c->d.flags, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(nelec, c->eos_nelec, nc * 8, hipMemcpyDeviceToHost, c->stream));
    if (out_natom) HIPCHK(hipMemcpyAsync(out_natom, c->eos_natom, nc * 8, hipMemcpyDeviceToHost, c->stream));
    if (out_natomfr) HIPCHK(hipMemcpyAsync(out_natomfr, c->eos_natomfr, (size_t)30 * nc * 8, hipMemcpyDeviceToHost, c->stream));
    if (out_oc && nion > 0) HIPCHK(hipMemcpyAsync(out_oc, c->eos_oc, (size_t)nion * nc * 8, hipMemcpyDeviceToHost, c->stream));
    if (out_ginv && nion > 0) HIPCHK(hipMemcpyAsync(out_ginv, c->eos_ginv, (size_t)nion * nc * 8, hipMemcpyDeviceToHost, c->stream));
    if (out_niter) HIPCHK(hipMemcpyAsync(out_niter, c->eos_niter, nc * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    int rc = tm.stop();
    if (rc) return rc;
    if (flag) {
        c->have_eos = false;
        return fail(SLGPU_EINVAL, "ions_solve_eos: nelec <= 0 or invalid Saha normalisation (ionsmod.f:291,338-349)");
    }
    c->have_eos = true;
    return 0;
}

int slgpu_eos_nlte_ndens(slgpu_ctx *c, double *out_ndens)
{
    if (!c || !out_ndens) return fail(SLGPU_EINVAL, "null argument");
    if (!c->have_eos) return fail(SLGPU_EORDER, "slgpu_eos must run first");
    const int nc = c->nx;
    double *dn = nullptr;
    int rc = dalloc(c, &dn, (size_t)2 * nc);
    if (rc) return rc;
    launch_eos_ndens(c, c->eos_temp, c->eos_mass, c->eos_natom, c->eos_natomfr, dn);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out_ndens, dn, (size_t)2 * nc * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    dfree(c, dn);
    return 0;
}

int slgpu_eos_vjp(slgpu_ctx *c, const double *grad_natom, const double *grad_natomfr,
                  const double *grad_nelec, const double *grad_oc, double *out_grad_temp,
                  double *out_grad_mass, double *out_grad_massfr)
{
    if (!c) return fail(SLGPU_EINVAL, "null context");
    if (!c->have_eos) return fail(SLGPU_EORDER, "slgpu_eos must run first");
    const int nc = c->nx, nion = c->nion;
    PhaseTimer tm(c, 9);
    double *buf = nullptr;
    const size_t nin = (size_t)nc * (2 + 30 + std::max(nion, 1)), nout = (size_t)nc * 32;
    int rc = dalloc(c, &buf, nin + nout);
    if (rc) return rc;
    double *gna = buf, *gne = buf + nc, *gfr = buf + 2 * (size_t)nc, *goc = buf + 32 * (size_t)nc;
    double *oT = buf + nin, *om = oT + nc, *omf = om + nc;
    if (grad_natom) HIPCHK(hipMemcpyAsync(gna, grad_natom, nc * 8, hipMemcpyHostToDevice, c->stream));
    if (grad_nelec) HIPCHK(hipMemcpyAsync(gne, grad_nelec, nc * 8, hipMemcpyHostToDevice, c->stream));
    if (grad_natomfr) HIPCHK(hipMemcpyAsync(gfr, grad_natomfr, (size_t)30 * nc * 8, hipMemcpyHostToDevice, c->stream));
    if (grad_oc && nion > 0) HIPCHK(hipMemcpyAsync(goc, grad_oc, (size_t)nion * nc * 8, hipMemcpyHostToDevice, c->stream));
    launch_eos_vjp(c, c->eos_temp, c->eos_mass, c->eos_massfr, c->eos_nelec, c->eos_natom,
                   c->eos_natomfr, c->eos_oc, gna, gfr, gne, goc, oT, om, omf);
    HIPCHK(hipGetLastError());
    if (out_grad_temp) HIPCHK(hipMemcpyAsync(out_grad_temp, oT, nc * 8, hipMemcpyDeviceToHost, c->stream));
    if (out_grad_mass) HIPCHK(hipMemcpyAsync(out_grad_mass, om, nc * 8, hipMemcpyDeviceToHost, c->stream));
    if (out_grad_massfr) HIPCHK(hipMemcpyAsync(out_grad_massfr, omf, (size_t)30 * nc * 8, hipMemcpyDeviceToHost, c->stream));
    rc = tm.stop();
    dfree(c, buf);
    return rc;
}

int slgpu_timers(slgpu_ctx *c, double *ms, int n, int reset)
{
    if (!c) return fail(SLGPU_EINVAL, "null context");
    for (int k = 0; k < n && k < 12; ++k) ms[k] = c->t_ms[k];
    if (reset)
        for (double &t : c->t_ms) t = 0.0;
    return 0;
}

}  // extern "C"
